// On-device exchange-correlation functionals (FP64), unpolarised.
//
// The reference evaluates these on the host through libxc 7.0.0
// (xc_integrals.hpp:113, :471-472).  Here they run per grid point inside the
// density kernels.  Each routine returns the energy density e = rho * eps and
// the partial derivatives de/drho, de/dsigma, summed over a functional list.
//
// Parameters follow libxc: LDA_X (1), LDA_C_PW (12), LDA_C_PW_MOD (13),
// GGA_X_PBE (101: kappa 0.804, mu = beta pi^2/3, beta 0.06672455060314922),
// GGA_C_PBE (130: same beta, gamma = (1 - ln 2)/pi^2, PW92 with the PW_MOD
// parameter set), GGA_X_B88 (106), GGA_C_LYP (131), LDA_C_VWN (7), GGA_X_PW91
// (109) and GGA_C_PW91 (134) (xc_extra.hpp).  Threshold handling: rho < thr contributes nothing; sigma
// is floored at thr^(8/3).
#pragma once

#include <hip/hip_runtime.h>

#include "xc_extra.hpp"

namespace nkx {

struct XcPoint {
  double e, vrho, vsigma;
};

namespace xcd {
// B88 (106), LYP (131), VWN (7), PW91 (109, 134) at the unpolarised point: the spin-resolved forms
// at rho_a = rho_b = rho / 2, sigma_ss = sigma_ab = sigma / 4, so
// de/drho = (e_a + e_b) / 2 and de/dsigma = (e_aa + e_ab + e_bb) / 4.
// Not inlined: its dual-number code would otherwise set the register count
// of every kernel that evaluates the functional list, PBE included.
__host__ __device__ inline __attribute__((noinline)) void blyp_unpol(int id, double rho,
                                                                    double sigma, XcPoint &o) {
  const D5 ra = dvar(0.5 * rho, 0), rb = dvar(0.5 * rho, 1);
  const D5 saa = dvar(0.25 * sigma, 2), sab = dvar(0.25 * sigma, 3), sbb = dvar(0.25 * sigma, 4);
  D5 e;
  switch (id) {
    case 106: e = xcb::b88_channel(ra, saa) + xcb::b88_channel(rb, sbb); break;
    case 7: e = xcb::vwn(ra, rb, 2.220446049250313e-16); break;
    case 109:
      e = 0.5 * (xcb::pw91_x_unpol(2.0 * ra, 4.0 * saa) + xcb::pw91_x_unpol(2.0 * rb, 4.0 * sbb));
      break;
    case 134: e = xcb::pw91_c(ra, rb, saa + 2.0 * sab + sbb, 2.220446049250313e-16); break;
    default: e = xcb::lyp(ra, rb, saa, sab, sbb); break;
  }
  o.e += e.v;
  o.vrho += 0.5 * (e.d[0] + e.d[1]);
  o.vsigma += 0.25 * (e.d[2] + e.d[3] + e.d[4]);
}
}  // namespace xcd

namespace xcd {

constexpr double kPi = 3.14159265358979323846;
// -(3/4)(3/pi)^(1/3)
constexpr double kCx = -0.73855876638202240588;
// (3 pi^2)^(1/3)
constexpr double kKf = 3.09366772628013593097;
// (3/(4 pi))^(1/3)
constexpr double kRs = 0.62035049089940001667;
constexpr double kBeta = 0.06672455060314922;
constexpr double kKappa = 0.8040;

__host__ __device__ inline void slater_x(double rho, double r13, XcPoint &o) {
  o.e += kCx * rho * r13;
  o.vrho += (4.0 / 3.0) * kCx * r13;
}

// PW92 paramagnetic correlation energy per particle and its rs-derivative.
__host__ __device__ inline void pw92(double rs, double A, double &eps, double &deps) {
  const double a1 = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382, b4 = 0.49294;
  const double s = sqrt(rs);
  const double den = 2.0 * A * (b1 * s + b2 * rs + b3 * rs * s + b4 * rs * rs);
  const double dden = A * (b1 / s + 2.0 * b2 + 3.0 * b3 * s + 4.0 * b4 * rs);
  const double lg = log(1.0 + 1.0 / den);  // libxc's form (not log1p)
  const double pre = -2.0 * A * (1.0 + a1 * rs);
  eps = pre * lg;
  deps = -2.0 * A * a1 * lg - pre * dden / (den * (den + 1.0));
}

__host__ __device__ inline void pw_c(double rho, double r13, double A, XcPoint &o) {
  const double rs = kRs / r13;
  double eps, deps;
  pw92(rs, A, eps, deps);
  o.e += rho * eps;
  o.vrho += eps - rs * deps / 3.0;
}

__host__ __device__ inline void pbe_x(double rho, double r13, double sigma, XcPoint &o) {
  const double mu = kBeta * kPi * kPi / 3.0;
  // s^2 = sigma / (4 kF^2 rho^2) = sigma / (4 (3pi^2)^(2/3) rho^(8/3))
  const double r43 = rho * r13;
  const double inv83 = 1.0 / (r43 * r43);
  const double cs = 1.0 / (4.0 * kKf * kKf);
  const double s2 = cs * sigma * inv83;
  const double q = 1.0 / (1.0 + mu * s2 / kKappa);
  const double F = 1.0 + kKappa - kKappa * q;
  const double dF = mu * q * q;
  const double ex = kCx * r43;
  o.e += ex * F;
  o.vrho += (4.0 / 3.0) * kCx * r13 * F - (8.0 / 3.0) * ex * dF * s2 / rho;
  o.vsigma += ex * dF * cs * inv83;
}

__host__ __device__ inline void pbe_c(double rho, double r13, double sigma, XcPoint &o) {
  const double gam = 0.031090690869654895;  // (1 - ln 2) / pi^2
  const double bg = kBeta / gam;
  const double rs = kRs / r13;
  double ec, decr;
  pw92(rs, 0.0310907, ec, decr);
  const double dec = -decr * rs / (3.0 * rho);  // d eps_c / d rho
  // t^2 = sigma pi / (16 kF rho^2)
  const double kF = kKf * r13;
  const double ct = kPi / (16.0 * kF * rho * rho);
  const double t2 = sigma * ct;
  const double ex = exp(-ec / gam);
  const double A = bg / (ex - 1.0);
  const double dA = bg * ex / (gam * (ex - 1.0) * (ex - 1.0));  // dA / d eps_c
  const double At2 = A * t2;
  const double num = 1.0 + At2;
  const double dnm = 1.0 + At2 * (1.0 + At2);
  const double inv = 1.0 / dnm;
  const double y = bg * t2 * num * inv;
  const double H = gam * log(1.0 + y);
  const double Hy = gam / (1.0 + y);
  // dy/dt2 = bg (1 + 2 A t2) / dnm^2 ;  dy/dA = -bg t2^2 (A t2)(2 + A t2) / dnm^2
  const double dyt = bg * (1.0 + 2.0 * At2) * inv * inv;
  const double dyA = -bg * t2 * t2 * At2 * (2.0 + At2) * inv * inv;
  const double dt2 = -(7.0 / 3.0) * t2 / rho;
  o.e += rho * (ec + H);
  o.vrho += ec + H + rho * (dec + Hy * (dyA * dA * dec + dyt * dt2));
  o.vsigma += rho * Hy * dyt * ct;
}

}  // namespace xcd

// Functional list, evaluated point by point.  ids are libxc ids; thr the
// density thresholds (one per functional).
__host__ __device__ inline XcPoint xc_eval(int n, const int *ids, const double *thr,
                                           double rho, double sigma) {
  XcPoint o{0.0, 0.0, 0.0};
  for (int k = 0; k < n; ++k) {
    const double t = thr[k];
    if (!(rho >= t)) continue;
    const double r13 = cbrt(rho);
    const double st = t * t * cbrt(t * t);  // (t^(4/3))^2
    const double s = sigma > st ? sigma : st;
    switch (ids[k]) {
      case 1: xcd::slater_x(rho, r13, o); break;
      case 12: xcd::pw_c(rho, r13, 0.031091, o); break;
      case 13: xcd::pw_c(rho, r13, 0.0310907, o); break;
      case 101: xcd::pbe_x(rho, r13, s, o); break;
      case 130: xcd::pbe_c(rho, r13, s, o); break;
      case 7:
      case 106:
      case 109:
      case 131:
      case 134: xcd::blyp_unpol(ids[k], rho, s, o); break;
      default: break;
    }
  }
  return o;
}

}  // namespace nkx
