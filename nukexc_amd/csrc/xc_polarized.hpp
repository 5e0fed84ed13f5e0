// Spin-polarised exchange-correlation functionals (FP64), libxc layout:
//   rho = (rho_a, rho_b), sigma = (sigma_aa, sigma_ab, sigma_bb),
//   vrho = (de/drho_a, de/drho_b), vsigma = (de/dsigma_aa, de/dsigma_ab, de/dsigma_bb).
// The reference evaluates these with libxc XC_POLARIZED on the host
// (xc_integrals.hpp:203, :617).  The energy density of each functional is
// written once over a forward-mode dual number carrying the five first
// derivatives, so the potentials are the exact derivatives of the energy.
//
// Exchange uses spin scaling, E_x[ra, rb] = (E_x[2 ra] + E_x[2 rb]) / 2.
// PW92 correlation interpolates between the paramagnetic, ferromagnetic and
// spin-stiffness fits with f(zeta) (libxc f_pw); PBE correlation adds
// H(rs, zeta, t) with phi(zeta) = ((1+z)^(2/3) + (1-z)^(2/3)) / 2 and
// t^2 = sigma_tot / (4 phi^2 ks^2 rho^2).  1 +- zeta is formed as
// 2 rho_a / rho and 2 rho_b / rho (no cancellation near |zeta| = 1) and floored
// at the machine epsilon (libxc zeta_threshold).
#pragma once

#include <hip/hip_runtime.h>

#include "xc_dual.hpp"
#include "xc_functionals.hpp"

namespace nkx {

namespace xcp {

constexpr double kPi = 3.14159265358979323846;
constexpr double kZetaThr = 2.220446049250313e-16;

// unpolarised exchange energy density of density n, gradient^2 s (D5 inputs)
__host__ __device__ inline D5 lda_x_unpol(D5 n) {
  return xcd::kCx * dpow(n, 4.0 / 3.0);
}
__host__ __device__ inline D5 pbe_x_unpol(D5 n, D5 s) {
  const double mu = xcd::kBeta * kPi * kPi / 3.0;
  const double cs = 1.0 / (4.0 * xcd::kKf * xcd::kKf);
  D5 n43 = dpow(n, 4.0 / 3.0);
  D5 s2 = cs * s / (n43 * n43);
  D5 F = (1.0 + xcd::kKappa) - xcd::kKappa / (1.0 + (mu / xcd::kKappa) * s2);
  return xcd::kCx * n43 * F;
}

struct PW {
  double A, a1, b1, b2, b3, b4;
};
__host__ __device__ inline D5 pw_g(D5 rs, const PW &p) {
  D5 s = dsqrt(rs);
  D5 den = 2.0 * p.A * (p.b1 * s + p.b2 * rs + p.b3 * rs * s + p.b4 * rs * rs);
  return -2.0 * p.A * (1.0 + p.a1 * rs) * dlog(1.0 + 1.0 / den);
}

// PW92 correlation energy per particle eps_c(rs, zeta); mod selects the
// libxc PW_MOD parameters (also used inside GGA_C_PBE).
__host__ __device__ inline D5 pw92_eps(D5 rs, D5 zeta, D5 opz, D5 omz, bool mod) {
  const PW p0 = {mod ? 0.0310907 : 0.031091, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294};
  const PW p1 = {mod ? 0.01554535 : 0.015545, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517};
  const PW p2 = {mod ? 0.0168869 : 0.016887, 0.11125, 10.357, 3.6231, 0.88026, 0.49671};
  const double fz20 = mod ? 1.709920934161365617563962776245 : 1.709921;
  D5 g0 = pw_g(rs, p0), g1 = pw_g(rs, p1), g2 = pw_g(rs, p2);
  D5 fz = (dpow(opz, 4.0 / 3.0) + dpow(omz, 4.0 / 3.0) - 2.0) * (1.0 / (pow(2.0, 4.0 / 3.0) - 2.0));
  D5 z2 = zeta * zeta;
  D5 z4 = z2 * z2;
  return g0 + z4 * fz * (g1 - g0 + g2 * (1.0 / fz20)) - fz * g2 * (1.0 / fz20);
}

__host__ __device__ inline D5 pbe_c_eps(D5 rho, D5 zeta, D5 opz, D5 omz, D5 sig_tot) {
  const double gam = 0.031090690869654895;  // (1 - ln 2)/pi^2
  const double bg = xcd::kBeta / gam;
  D5 rs = xcd::kRs / dpow(rho, 1.0 / 3.0);
  D5 ec = pw92_eps(rs, zeta, opz, omz, true);
  D5 phi = 0.5 * (dpow(opz, 2.0 / 3.0) + dpow(omz, 2.0 / 3.0));
  D5 phi3 = phi * phi * phi;
  D5 kF = xcd::kKf * dpow(rho, 1.0 / 3.0);
  D5 t2 = sig_tot * kPi / (16.0 * phi * phi * kF * rho * rho);
  D5 A = bg / (dexp(-1.0 * ec / (gam * phi3)) - 1.0);
  // y = bg t2 (1 + u) / (1 + u + u^2), u = A t2, with its partial
  // derivatives in closed form: dy/dt2 = bg (1 + 2u) / D^2 and
  // dy/dA = -bg t2^2 u (2 + u) / D^2 (D = 1 + u + u^2).  Differentiating the
  // quotient term by term cancels to a relative 1/u^2 where the enhancement
  // saturates (low density, large gradients).
  D5 y;
  {
    const double u = A.v * t2.v, D = 1.0 + u + u * u;
    const double dt = bg * (1.0 + 2.0 * u) / (D * D);
    const double dA = -bg * t2.v * t2.v * u * (2.0 + u) / (D * D);
    y.v = bg * t2.v * (1.0 + u) / D;
    for (int k = 0; k < 5; ++k) y.d[k] = dt * t2.d[k] + dA * A.d[k];
  }
  D5 H = gam * phi3 * dlog(1.0 + y);
  return ec + H;
}

}  // namespace xcp

// Sum over the functional list at one point, libxc XC_POLARIZED semantics:
// a functional contributes nothing where rho_a + rho_b < thr; otherwise the
// inputs are clamped (rho_s >= thr, sigma_ss >= thr^(8/3),
// |sigma_ab| <= (sigma_aa + sigma_bb) / 2) and exc is the energy per particle
// of the clamped point; the potentials are the derivatives there.  The caller
// integrates w (rho_a + rho_b) exc.
__host__ __device__ inline void xc_eval_polarized(int n, const int *ids, const double *thr,
                                                  const double rho[2], const double sigma[3],
                                                  double &exc, double vrho[2], double vsigma[3]) {
  exc = 0.0;
  vrho[0] = vrho[1] = 0.0;
  vsigma[0] = vsigma[1] = vsigma[2] = 0.0;
  for (int k = 0; k < n; ++k) {
    const double t = thr[k];
    if (!(rho[0] + rho[1] >= t)) continue;
    const double st = t * t * cbrt(t * t);
    const double xa = fmax(rho[0], t), xb = fmax(rho[1], t);
    const double xaa = fmax(sigma[0], st), xbb = fmax(sigma[2], st);
    const double ave = 0.5 * (xaa + xbb);
    const double xab = fmin(fmax(sigma[1], -ave), ave);
    D5 ra = dvar(xa, 0), rb = dvar(xb, 1);
    D5 saa = dvar(xaa, 2), sab = dvar(xab, 3), sbb = dvar(xbb, 4);
    D5 r;
    switch (ids[k]) {
      case 1:  // spin-scaled Slater exchange; a channel at or below thr drops out
        r = dconst(0.0);
        if (xa > t) r = r + 0.5 * xcp::lda_x_unpol(2.0 * ra);
        if (xb > t) r = r + 0.5 * xcp::lda_x_unpol(2.0 * rb);
        break;
      case 101:
        r = dconst(0.0);
        if (xa > t) r = r + 0.5 * xcp::pbe_x_unpol(2.0 * ra, 4.0 * saa);
        if (xb > t) r = r + 0.5 * xcp::pbe_x_unpol(2.0 * rb, 4.0 * sbb);
        break;
      case 12:
      case 13: {
        D5 rr = ra + rb;
        D5 rs = xcd::kRs / dpow(rr, 1.0 / 3.0);
        D5 opz = dmax(2.0 * ra / rr, xcp::kZetaThr), omz = dmax(2.0 * rb / rr, xcp::kZetaThr);
        r = rr * xcp::pw92_eps(rs, (ra - rb) / rr, opz, omz, ids[k] == 13);
        break;
      }
      case 106:  // B88 is spin-resolved; a channel at or below thr drops out
        r = dconst(0.0);
        if (xa > t) r = r + xcb::b88_channel(ra, saa);
        if (xb > t) r = r + xcb::b88_channel(rb, sbb);
        break;
      case 131:
        r = xcb::lyp(ra, rb, saa, sab, sbb);
        break;
      case 109:  // spin-scaled; a channel at or below thr drops out
        r = dconst(0.0);
        if (xa > t) r = r + 0.5 * xcb::pw91_x_unpol(2.0 * ra, 4.0 * saa);
        if (xb > t) r = r + 0.5 * xcb::pw91_x_unpol(2.0 * rb, 4.0 * sbb);
        break;
      case 134:
        r = xcb::pw91_c(ra, rb, saa + 2.0 * sab + sbb, xcp::kZetaThr);
        break;
      case 7:
        r = xcb::vwn(ra, rb, xcp::kZetaThr);
        break;
      case 130: {
        D5 rr = ra + rb;
        D5 opz = dmax(2.0 * ra / rr, xcp::kZetaThr), omz = dmax(2.0 * rb / rr, xcp::kZetaThr);
        r = rr * xcp::pbe_c_eps(rr, (ra - rb) / rr, opz, omz, saa + 2.0 * sab + sbb);
        break;
      }
      default:
        continue;
    }
    exc += r.v / (xa + xb);
    vrho[0] += r.d[0];
    vrho[1] += r.d[1];
    vsigma[0] += r.d[2];
    vsigma[1] += r.d[3];
    vsigma[2] += r.d[4];
  }
}

}  // namespace nkx
