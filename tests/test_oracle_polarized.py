"""Pins the oracle's spin-polarised functionals and V_xc (libxc XC_POLARIZED,
called by compute_lsda / compute_gga_lsda, xc_integrals.hpp:169-262,
:543-718).  libxc is not installed, so the checks are exact identities that a
wrong sign, a dropped term or a misplaced spin index would break:
spin scaling onto the (separately pinned) unpolarised forms, the PBE
gradient expansion, the large-gradient limit, finite differences, and
dE/dP_s = V_s.  The PW92 parameter sets and PBE's beta and gamma are pinned
in test_oracle_pw92_pbe_c.py."""
import math

import numpy as np
import pytest

from oracle import oracle as O

IDS = [1, 12, 13, 101, 130]
RNG = np.random.default_rng(11)


def _pol(ids, ra, rb, saa=0.0, sab=0.0, sbb=0.0):
    ra, rb = np.broadcast_arrays(np.atleast_1d(np.asarray(ra, float)),
                                 np.atleast_1d(np.asarray(rb, float)))
    rho = np.column_stack([ra, rb])
    n = rho.shape[0]
    sig = np.column_stack([np.broadcast_to(saa, n), np.broadcast_to(sab, n),
                           np.broadcast_to(sbb, n)]).astype(float)
    return O.xc_eval_polarized(ids, rho, sig)


def _edens(ids, ra, rb, saa=0.0, sab=0.0, sbb=0.0):
    exc, _, _ = _pol(ids, ra, rb, saa, sab, sbb)
    return exc * (np.asarray(ra, float) + np.asarray(rb, float))


@pytest.mark.parametrize("fid", IDS)
def test_spin_scaling_to_unpolarised(fid):
    """rho_a = rho_b = rho/2, every sigma = sigma/4 is the unpolarised point:
    same exc, vrho_s = vrho, (vs_aa + vs_ab + vs_bb)/4 = vsigma."""
    rho = 10 ** RNG.uniform(-6, 2, 400)
    sig = rho ** (8 / 3) * 10 ** RNG.uniform(-3, 2, 400)
    e0, v0, s0 = O.xc_eval([fid], rho, sig)
    e1, v1, s1 = _pol([fid], rho / 2, rho / 2, sig / 4, sig / 4, sig / 4)
    assert np.allclose(e1, e0, rtol=1e-13, atol=0)
    assert np.allclose(v1[:, 0], v0, rtol=1e-12, atol=0)
    assert np.allclose(v1[:, 1], v0, rtol=1e-12, atol=0)
    if fid in (101, 130):
        assert np.allclose(s1.sum(1) / 4, s0, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("fid", [1, 101])
def test_exchange_spin_scaling_identity(fid):
    """E_x[ra, rb] = (E_x[2 ra] + E_x[2 rb]) / 2 with sigma_ss -> 4 sigma_ss,
    evaluated through the unpolarised oracle; sigma_ab does not enter."""
    ra, rb = 10 ** RNG.uniform(-5, 1, 200), 10 ** RNG.uniform(-5, 1, 200)
    saa, sbb = ra ** (8 / 3) * 10 ** RNG.uniform(-2, 2, 200), rb ** (8 / 3) * 10 ** RNG.uniform(-2, 2, 200)
    sab = 0.3 * np.sqrt(saa * sbb)
    ea, va, sa = O.xc_eval([fid], 2 * ra, 4 * saa)
    eb, vb, sb = O.xc_eval([fid], 2 * rb, 4 * sbb)
    ref = 0.5 * (ea * 2 * ra + eb * 2 * rb)
    e, v, s = _pol([fid], ra, rb, saa, sab, sbb)
    assert np.allclose(e * (ra + rb), ref, rtol=1e-13, atol=0)
    assert np.allclose(v[:, 0], va, rtol=1e-12) and np.allclose(v[:, 1], vb, rtol=1e-12)
    if fid == 101:
        assert np.allclose(s[:, 0], 2 * sa, rtol=1e-12) and np.allclose(s[:, 2], 2 * sb, rtol=1e-12)
        assert np.all(s[:, 1] == 0.0)


def test_lda_x_fully_polarised_closed_form():
    """e_x(rho, 0) = 2^(1/3) e_x^unpol(rho) = -(3/4)(6/pi)^(1/3) rho^(4/3)."""
    rho = 10 ** RNG.uniform(-4, 2, 50)
    e = _edens([1], rho, 0.0)
    assert np.allclose(e, -0.75 * (6 / math.pi) ** (1 / 3) * rho ** (4 / 3), rtol=1e-14)


def test_pbe_c_zero_gradient_is_pw92_mod():
    ra, rb = 10 ** RNG.uniform(-4, 1, 100), 10 ** RNG.uniform(-4, 1, 100)
    e130 = _edens([130], ra, rb, 0.0, 0.0, 0.0)
    e13 = _edens([13], ra, rb)
    assert np.allclose(e130, e13, rtol=1e-12, atol=0)


def test_pbe_c_gradient_expansion():
    """Small t: rho H -> rho beta phi^3 t^2, so d e / d sigma_tot -> beta phi /
    (4 ks^2 rho) with ks^2 = 4 kF / pi, kF = (3 pi^2 rho)^(1/3), phi =
    ((1+z)^(2/3) + (1-z)^(2/3))/2 (PBE, PRL 77, 3865, Eqs. 3-8); and
    vsigma_ab = 2 vsigma_aa = 2 vsigma_bb."""
    beta = 0.06672455060314922
    for rho, z in [(0.05, 0.0), (0.3, 0.4), (2.0, -0.7), (1.0, 0.95)]:
        ra, rb = rho * (1 + z) / 2, rho * (1 - z) / 2
        tiny = 1e-10 * rho ** (8 / 3)
        _, _, s = _pol([130], ra, rb, tiny, tiny, tiny)
        phi = ((1 + z) ** (2 / 3) + (1 - z) ** (2 / 3)) / 2
        ks2 = 4 * (3 * math.pi ** 2 * rho) ** (1 / 3) / math.pi
        ref = beta * phi / (4 * ks2 * rho)
        assert abs(s[0, 0] - ref) < 1e-6 * ref
        assert abs(s[0, 2] - ref) < 1e-6 * ref
        assert abs(s[0, 1] - 2 * ref) < 1e-6 * ref


def test_pbe_c_large_gradient_limit():
    """t -> infinity: H -> -eps_c, so the correlation energy density vanishes."""
    rho = np.array([0.01, 0.5, 3.0])
    e_big = _edens([130], 0.6 * rho, 0.4 * rho, 1e12 * rho, 0.0, 1e12 * rho)
    e_pw = _edens([13], 0.6 * rho, 0.4 * rho)
    assert np.all(np.abs(e_big) < 1e-6 * np.abs(e_pw))


@pytest.mark.parametrize("fid", IDS)
def test_polarised_derivatives_fd(fid):
    """vrho_s and vsigma_k are the derivatives of e = rho exc (fourth-order
    central differences, relative step 1e-3 for rho_s, 1e-2 for sigma_k)."""
    n = 40
    ra, rb = 10 ** RNG.uniform(-2, 1, n), 10 ** RNG.uniform(-2, 1, n)
    saa, sbb = ra ** (8 / 3) * 10 ** RNG.uniform(-1, 1, n), rb ** (8 / 3) * 10 ** RNG.uniform(-1, 1, n)
    sab = 0.5 * np.sqrt(saa * sbb) * RNG.uniform(-1, 1, n)
    x = [ra, rb, saa, sab, sbb]
    _, v, s = _pol([fid], *x)
    ana = np.column_stack([v, s])
    for k in range(5 if fid in (101, 130) else 2):
        # sigma enters e only through the gradient correction: a larger step
        # keeps the difference above rounding of the LDA part
        h = (1e-3 if k < 2 else 1e-2) * (np.sqrt(saa * sbb) if k == 3 else x[k])

        def e_at(t):
            xs = [a.copy() for a in x]
            xs[k] = xs[k] + t * h
            return _edens([fid], *xs)

        e0 = e_at(0)
        fd = (-e_at(2) + 8 * e_at(1) - 8 * e_at(-1) + e_at(-2)) / (12 * h)
        # truncation O(h^4) plus rounding of the four energies, ~ eps |e| / h
        tol = 1e-8 * np.abs(ana[:, k]) + 20 * 2.2e-16 * np.abs(e0) / h
        assert np.all(np.abs(fd - ana[:, k]) < tol), (k, np.max(np.abs(fd - ana[:, k]) / tol))


@pytest.mark.parametrize("fid", IDS)
def test_spin_swap_symmetry(fid):
    ra, rb = 10 ** RNG.uniform(-3, 1, 50), 10 ** RNG.uniform(-3, 1, 50)
    saa, sbb = ra ** (8 / 3), 2 * rb ** (8 / 3)
    sab = 0.2 * np.sqrt(saa * sbb)
    e1, v1, s1 = _pol([fid], ra, rb, saa, sab, sbb)
    e2, v2, s2 = _pol([fid], rb, ra, sbb, sab, saa)
    assert np.allclose(e1, e2, rtol=1e-14)
    assert np.allclose(v1, v2[:, ::-1], rtol=1e-12)
    assert np.allclose(s1, s2[:, ::-1], rtol=1e-12, atol=1e-300)


def test_polarised_clamps():
    """libxc input clamps: a zero channel is evaluated at rho_s = thr and
    its exchange drops out; below the total threshold everything is zero."""
    e, v, s = _pol([1, 12], [0.3, 1e-16], [0.0, 1e-16])
    ex = _edens([1], 0.3, 0.0)[0]
    assert v[1, 0] == v[1, 1] == 0.0 and e[1] == 0.0
    # exchange of the empty channel contributes nothing
    e_x, v_x, _ = _pol([1], 0.3, 0.0)
    assert v_x[0, 1] == 0.0 and abs(e_x[0] * 0.3 - ex) < 1e-16


# ------------------------------------------------------ V_xc (oracle) ------
def _atom_pair_setup():
    """Two-centre STO basis and a small grid (H2-like at 1.4 bohr)."""
    xyz = np.array([[0.0, 0.0, 0.0], [1.4, 0.0, 0.0]])
    from conftest import lebedev
    P, W, _ = O.make_flat_grid(xyz, [1, 1], 30, lebedev(17), scheme=0, alpha=0.6)
    fns = [(1, 0, 0, 1.2), (2, 1, 0, 1.0), (2, 1, 1, 0.9)]
    n, l, m, z, c = [], [], [], [], []
    for a in range(2):
        for (nn, ll, mm, zz) in fns:
            n.append(nn); l.append(ll); m.append(mm); z.append(zz); c.append(xyz[a])
    basis = dict(n=n, l=l, m=m, zeta=z, norm=[O.sto_norm_cutoff(a, b)[0] for a, b in zip(n, z)],
                 centers=np.array(c))
    phi, gx, gy, gz = O.collocation(basis, P, grad=True)
    return phi, gx, gy, gz, W


def _spin_dms(nbf, rng):
    """Positive-definite spin density matrices, so that rho_s > 0 everywhere
    and a small change of one element changes rho_s by a small fraction."""
    A = rng.normal(size=(nbf, nbf))
    B = rng.normal(size=(nbf, nbf))
    I = np.eye(nbf)
    return 0.3 * (A @ A.T / nbf + 0.2 * I), 0.2 * (B @ B.T / nbf + 0.2 * I)


@pytest.mark.parametrize("ids", [[1, 12], [101, 130]])
def test_vxc_polarized_spin_symmetric_is_unpolarised(ids):
    phi, gx, gy, gz, W = _atom_pair_setup()
    nbf = phi.shape[0]
    Pa, _ = _spin_dms(nbf, np.random.default_rng(2))
    grad = dict(gx=gx, gy=gy, gz=gz)
    E0, N0, V0 = O.vxc_from_collocation(phi, W, ids, P=2 * Pa, **grad)
    E1, N1, Va, Vb = O.vxc_polarized(phi, W, ids, Pa, Pa, **grad)
    assert abs(E1 - E0) < 1e-12 * abs(E0)
    assert abs(N1.sum() - N0) < 1e-12 * N0
    assert np.abs(Va - V0).max() < 1e-11 * np.abs(V0).max()
    assert np.abs(Vb - V0).max() < 1e-11 * np.abs(V0).max()


@pytest.mark.parametrize("ids", [[1, 12], [101, 130]])
def test_vxc_polarized_is_gradient_of_energy(ids):
    """V_s = dE/dP_s: symmetric finite differences of E_xc in P_alpha and
    P_beta (off-diagonal perturbations carry a factor 2)."""
    phi, gx, gy, gz, W = _atom_pair_setup()
    nbf = phi.shape[0]
    rng = np.random.default_rng(5)
    Pa, Pb = _spin_dms(nbf, rng)
    grad = dict(gx=gx, gy=gy, gz=gz)
    _, _, Va, Vb = O.vxc_polarized(phi, W, ids, Pa, Pb, **grad)
    h = 1e-5
    for (mu, nu) in [(0, 0), (1, 4), (2, 5), (3, 3)]:
        for spin, V in ((0, Va), (1, Vb)):
            D = np.zeros((nbf, nbf))
            D[mu, nu] += h
            D[nu, mu] += h if mu != nu else 0.0
            Pp = [Pa.copy(), Pb.copy()]
            Pm = [Pa.copy(), Pb.copy()]
            Pp[spin] += D
            Pm[spin] -= D
            Ep = O.vxc_polarized(phi, W, ids, *Pp, **grad)[0]
            Em = O.vxc_polarized(phi, W, ids, *Pm, **grad)[0]
            fd = (Ep - Em) / (2 * h)
            ref = V[mu, nu] * (2 if mu != nu else 1)
            assert abs(fd - ref) < 1e-6 * max(1.0, abs(ref)), (ids, mu, nu, spin, fd, ref)
