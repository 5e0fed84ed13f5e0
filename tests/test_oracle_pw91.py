"""Pins of the oracle's PW91 exchange and correlation (GGA_X_PW91 = 109,
GGA_C_PW91 = 134; tests/standalone_helpers.hpp:201-202, 217-218), which the
reference evaluates through libxc (not installed here).  Each check probes
the oracle's function against numbers fixed outside it:

* exchange at small s: F(s) = 1 + (10/81) s^2 + O(s^4), the gradient
  expansion of exchange (PW91 builds c + d = 10/81 into its parameters);
* exchange at intermediate s is Becke 88 (pinned to Becke's H and He
  exchange energies in test_oracle_blyp.py) times (1 + q)/(1 + q + f s^4):
  PW91 takes a, b and c from B88's beta = 0.0042, so the two forms differ
  only by PW91's f s^4 term once exp(-100 s^2) has vanished;
* exchange at large s: F s^2 -> c/f (PW91's decay of the enhancement);
* correlation at t -> 0 is PW92 with the PW_MOD parameters (pinned in
  test_oracle_pw92_pbe_c.py);
* correlation's gradient coefficient at high density: H/t^2 -> nu (C_c(0)
  - 3 C_x / 7) phi^3 as t -> 0 and rs -> 0, with C_c(0) = beta/nu from the
  Ma-Brueckner coefficient that PBE also uses (beta = 0.066725, Perdew,
  Burke and Ernzerhof, PRL 77, 3865 (1996)) and C_x Sham's exchange
  gradient coefficient -7 / (432 pi (3 pi^2)^(1/3));
* correlation at t -> infinity vanishes (H -> -eps_c);
* spin: the polarised form at zeta = 0 is the unpolarised one, and exchange
  obeys spin scaling.

Left unpinned: the Rasolt-Geldart Pade C_xc(rs) away from rs = 0 and the
exponent of H1 (they enter only at finite rs and t), and libxc's input
thresholds."""
import math

import numpy as np
import pytest

from oracle import oracle as O

KF = (3 * math.pi ** 2) ** (1 / 3)


def _sigma_for_s(rho, s):
    """sigma = (2 kF rho s)^2 at density rho."""
    return (2 * KF * rho ** (4 / 3) * s) ** 2


def _F(fid, rho, s):
    """Enhancement factor over LDA exchange at (rho, s), unpolarised."""
    e = O.xc_eval([fid], np.array([rho]), np.array([_sigma_for_s(rho, s)]))[0][0]
    ex = O.xc_eval([1], np.array([rho]), np.array([0.0]))[0][0]
    return e / ex


def test_pw91_x_gradient_expansion():
    for rho in (0.01, 1.0, 30.0):
        s = 1e-3
        mu = (_F(109, rho, s) - 1) / s ** 2
        print(f"rho {rho}: (F - 1)/s^2 = {mu:.6f} (10/81 = {10 / 81:.6f})")
        assert abs(mu - 10 / 81) < 1e-4


@pytest.mark.parametrize("s", [0.8, 1.2, 2.0, 4.0])
def test_pw91_x_is_b88_with_the_f_term(s):
    """F_B88 / F_PW91 = 1 + f s^4 / (1 + q) = 1 + (f / c) s^2 (F_B88 - 1),
    with F_B88 = 1 + c s^2 / (1 + q), c = 0.2743 and f = 0.004 (PW91's
    published values)."""
    rho = 0.7
    fb, fp = _F(106, rho, s), _F(109, rho, s)
    want = fb / (1 + 0.004 / 0.2743 * s * s * (fb - 1))
    print(f"s {s}: F_B88 {fb:.8f}, F_PW91 {fp:.8f}, predicted {want:.8f}")
    # PW91's a, b, c are B88's beta in five figures: F agrees to ~1e-5
    assert abs(fp - want) < 3e-5
    # and the f term is really there: without it F_PW91 would equal F_B88
    assert abs(fp - fb) > 30 * abs(fp - want)


def test_pw91_x_large_s():
    s = 1e5
    f = _F(109, 1.0, s)
    print(f"F s^2 at s = 1e5: {f * s * s:.5f} (c/f = {0.2743 / 0.004:.5f})")
    assert abs(f * s * s - 0.2743 / 0.004) < 2e-3 * 0.2743 / 0.004


def _H(rho, t2):
    """PW91 H per particle at (rho, t^2), unpolarised: e/rho - eps_c(PW_MOD)."""
    ks2 = 4 * KF * rho ** (1 / 3) / math.pi
    sig = 4 * ks2 * rho * rho * t2
    e = O.xc_eval([134], np.array([rho]), np.array([sig]))[0][0]
    ec = O.xc_eval([13], np.array([rho]), np.array([0.0]))[0][0]
    return e - ec, ec


def test_pw91_c_reduces_to_pw92_mod_at_t0():
    for rho in (1e-3, 0.1, 10.0):
        h, ec = _H(rho, 1e-14)
        assert abs(h) < 1e-12 * abs(ec)


def test_pw91_c_gradient_coefficient_high_density():
    beta_mb = 0.066725
    nu = 16 / math.pi * KF
    cx = -7 / (432 * math.pi * KF)
    want = nu * (beta_mb / nu - 3 * cx / 7)
    rs = 1e-4
    rho = 3 / (4 * math.pi * rs ** 3)
    t2 = 1e-10
    h, _ = _H(rho, t2)
    got = h / t2
    print(f"H/t^2 at rs = 1e-4: {got:.6f}, nu (C_c(0) - 3 C_x/7) = {want:.6f}")
    assert abs(got - want) < 2e-4 * want


def test_pw91_c_vanishes_for_large_t():
    for rho in (0.01, 1.0):
        h, ec = _H(rho, 1e6)
        print(f"rho {rho}: eps_c {ec:.6e}, eps_c + H at t^2 = 1e6: {ec + h:.3e}")
        assert abs(ec + h) < 1e-5 * abs(ec)


def test_pw91_spin_forms():
    rng = np.random.default_rng(3)
    rho = 10 ** rng.uniform(-3, 1, 200)
    sig = rho ** (8 / 3) * 10 ** rng.uniform(-2, 1, 200)
    for fid in (109, 134):
        e0 = O.xc_eval([fid], rho, sig)[0]
        r2 = np.stack([rho / 2, rho / 2], 1)
        s3 = np.stack([sig / 4, sig / 4, sig / 4], 1)
        e1 = O.xc_eval_polarized([fid], r2, s3)[0]
        assert np.allclose(e1, e0, rtol=1e-13, atol=0)
    # spin scaling of exchange: E_x[ra, 0] = E_x[2 ra] / 2 (per volume)
    r2 = np.stack([rho, 0 * rho], 1)
    s3 = np.stack([sig, 0 * sig, 0 * sig], 1)
    ep = O.xc_eval_polarized([109], r2, s3)[0] * rho
    eu = O.xc_eval([109], 2 * rho, 4 * sig)[0] * 2 * rho
    # (the polarised exc divides by rho_a + thr)
    assert np.allclose(ep, eu / 2, rtol=1e-11, atol=0)


@pytest.mark.parametrize("fid", [109, 134])
def test_pw91_potentials_are_derivatives(fid):
    """vrho, vsigma against central differences of rho e."""
    rng = np.random.default_rng(5)
    rho = 10 ** rng.uniform(-2, 1, 50)
    sig = rho ** (8 / 3) * 10 ** rng.uniform(-2, 0.5, 50)
    exc, vr, vs = O.xc_eval([fid], rho, sig)
    h = 1e-6
    f = lambda r, s: O.xc_eval([fid], r, s)[0] * r  # noqa: E731
    dr = (f(rho * (1 + h), sig) - f(rho * (1 - h), sig)) / (2 * h * rho)
    ds = (f(rho, sig * (1 + h)) - f(rho, sig * (1 - h))) / (2 * h * sig)
    assert np.allclose(vr, dr, rtol=1e-6, atol=1e-10)
    assert np.allclose(vs, ds, rtol=1e-5, atol=1e-12)
