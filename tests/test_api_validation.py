"""Argument checks of the Python binding (host logic, no GPU): the C entry
points take device pointers without sizes, so the binding rejects density
matrices and output tensors whose shape, dtype, device or layout do not
match the ctx basis before any library call."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import nukexc_amd as nk  # noqa: E402
from nukexc_amd.api import Context  # noqa: E402


class _NoLib:
    def __getattr__(self, name):
        raise AssertionError(f"{name} must not be called for rejected arguments")


@pytest.fixture
def ctx(monkeypatch):
    """A Context whose device is the CPU and whose basis has 4 functions;
    any library call fails the test."""
    c = object.__new__(Context)
    c.h = None
    c.device = torch.device("cpu")
    c._nbf = lambda: 4
    c._molset_nbf = [4, 3]
    monkeypatch.setattr("nukexc_amd.api.lib", lambda: _NoLib())
    return c


def _status(exc):
    return exc.value.code


def test_vxc_rejects_wrong_sized_P(ctx):
    with pytest.raises(nk.NukexcError) as e:
        ctx.vxc(np.zeros((3, 3)))
    assert _status(e) == 1 and "4 x 4" in str(e.value)
    with pytest.raises(nk.NukexcError):
        ctx.vxc(np.zeros((4, 5)))


@pytest.mark.parametrize("bad", [
    torch.zeros((4, 4), dtype=torch.float32),
    torch.zeros((5, 5), dtype=torch.float64),
    torch.zeros((4, 8), dtype=torch.float64)[:, ::2],
    np.zeros((4, 4)),
])
def test_vxc_rejects_bad_output(ctx, bad):
    with pytest.raises(nk.NukexcError) as e:
        ctx.vxc(np.eye(4), out=bad)
    assert _status(e) == 1


def test_polarized_and_gradient_reject_wrong_sizes(ctx):
    with pytest.raises(nk.NukexcError):
        ctx.vxc_polarized(np.eye(4), np.eye(3))
    with pytest.raises(nk.NukexcError):
        ctx.xc_grad(np.eye(5))
    with pytest.raises(nk.NukexcError):
        ctx.xc_grad_polarized(np.eye(4), np.eye(2))
    with pytest.raises(nk.NukexcError):
        ctx.autotune(np.eye(2))


def test_graph_capture_rejects_bad_tensors(ctx):
    good = torch.zeros((4, 4), dtype=torch.float64)
    with pytest.raises(nk.NukexcError):
        ctx.graph_capture(good, torch.zeros((4, 4), dtype=torch.float32))
    with pytest.raises(nk.NukexcError):
        ctx.graph_capture(torch.zeros((3, 3), dtype=torch.float64), good)
    with pytest.raises(nk.NukexcError):
        ctx.graph_capture_polarized(good, good, good, good.t()[:, :3])


def test_molset_rejects_wrong_count_and_sizes(ctx):
    with pytest.raises(nk.NukexcError):
        ctx.molset_vxc([np.eye(4)])
    with pytest.raises(nk.NukexcError):
        ctx.molset_vxc([np.eye(4), np.eye(4)])
    good4, good3 = (torch.zeros((k, k), dtype=torch.float64) for k in (4, 3))
    with pytest.raises(nk.NukexcError):
        ctx.molset_graph_capture([good4, good3], [good4, good4])
