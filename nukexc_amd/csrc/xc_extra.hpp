// Becke 88 exchange (libxc GGA_X_B88, id 106), Lee-Yang-Parr correlation
// (libxc GGA_C_LYP, id 131) and Vosko-Wilk-Nusair correlation (libxc LDA_C_VWN,
// id 7) as spin-resolved energy densities over D5.
//
// B88 (A. D. Becke, Phys. Rev. A 38, 3098 (1988), Eq. 8), per spin channel:
//   e_s = rho_s^(4/3) [c_x - beta x_s^2 / (1 + 6 beta x_s asinh x_s)],
//   x_s = |grad rho_s| / rho_s^(4/3), c_x = -(3/2)(3/(4 pi))^(1/3), beta = 0.0042.
// LYP in the closed form of Miehlich, Savin, Stoll and Preuss (Chem. Phys.
// Lett. 157, 200 (1989), Eq. 2), a = 0.04918, b = 0.132, c = 0.2533,
// d = 0.349, C_F = (3/10)(3 pi^2)^(2/3):
//   e = -4a/(1 + d rho^-1/3) rho_a rho_b / rho
//       - a b w { rho_a rho_b [2^(11/3) C_F (rho_a^(8/3) + rho_b^(8/3))
//                 + (47/18 - 7 del/18) |grad rho|^2 - (5/2 - del/18)(s_aa + s_bb)
//                 - (del - 11)/9 (rho_a/rho s_aa + rho_b/rho s_bb)]
//               - 2/3 rho^2 |grad rho|^2 + (2/3 rho^2 - rho_a^2) s_bb
//               + (2/3 rho^2 - rho_b^2) s_aa },
//   w = exp(-c rho^-1/3) / (1 + d rho^-1/3) rho^(-11/3),
//   del = c rho^-1/3 + d rho^-1/3 / (1 + d rho^-1/3).
//
// PW91 exchange (libxc GGA_X_PW91, id 109; Perdew et al., Phys. Rev. B 46,
// 6671 (1992), Eq. 8): spin scaling of the unpolarised
//   e = A_x n^(4/3) F(s),  F = 1 + [(c + d e^(-alpha s^2)) s^2 - f s^4]
//                               / [1 + a s asinh(b s) + f s^4],
// s = |grad n| / (2 k_F n), a = 0.19645, b = 7.7956, c = 0.2743, d = -0.1508,
// f = 0.004, alpha = 100 (a, b, c are Becke's beta = 0.0042 in the s variable,
// c + d the gradient-expansion coefficient 10/81).
// PW91 correlation (libxc GGA_C_PW91, id 134; same paper, Eqs. 13-15) on
// PW92 with the PW_MOD parameters: eps_c + H0 + H1,
//   H0 = phi^3 beta^2/(2 alpha) ln(1 + (2 alpha/beta)(t^2 + A t^4)/(1 + A t^2 + A^2 t^4)),
//   A  = (2 alpha/beta) / (exp(-2 alpha eps_c / (phi^3 beta^2)) - 1),
//   H1 = nu (C_xc(rs) - C_xc(0) - 3 C_x/7) phi^3 t^2 exp(-100 phi^4 (k_s^2/k_F^2) t^2),
// alpha = 0.09, nu = (16/pi)(3 pi^2)^(1/3), beta = nu C_c(0), C_c(0) = 0.004235,
// C_x = -0.001667, C_xc(rs) the Rasolt-Geldart Pade
//   (2.568 + 23.266 rs + 0.007389 rs^2) / (1000 (1 + 8.723 rs + 0.472 rs^2 + 0.07389 rs^3)).
#pragma once

#include "xc_dual.hpp"

namespace nkx {
namespace xcb {

constexpr double kB88Beta = 0.0042;
constexpr double kCxSpin = -0.9305257363491002;  // -(3/2)(3/(4 pi))^(1/3)
constexpr double kLypA = 0.04918, kLypB = 0.132, kLypC = 0.2533, kLypD = 0.349;
constexpr double kCF = 2.871234000188191;      // (3/10)(3 pi^2)^(2/3)
constexpr double k2p113 = 12.699208415745595;  // 2^(11/3)

// one spin channel of B88: density r, sigma_ss s
__host__ __device__ inline D5 b88_channel(D5 r, D5 s) {
  const D5 r43 = dpow(r, 4.0 / 3.0);
  const D5 x = dsqrt(s) / r43;
  const D5 x2 = s / (r43 * r43);
  return r43 * (kCxSpin - kB88Beta * x2 / (1.0 + 6.0 * kB88Beta * x * dasinh(x)));
}

__host__ __device__ inline D5 lyp(D5 ra, D5 rb, D5 saa, D5 sab, D5 sbb) {
  const D5 rr = ra + rb;
  const D5 rm13 = dpow(rr, -1.0 / 3.0);
  const D5 den = 1.0 + kLypD * rm13;
  const D5 w = dexp(-kLypC * rm13) / den * dpow(rr, -11.0 / 3.0);
  const D5 del = kLypC * rm13 + kLypD * rm13 / den;
  const D5 sig = saa + 2.0 * sab + sbb;
  const D5 rr2 = rr * rr;
  const D5 inner = k2p113 * kCF * (dpow(ra, 8.0 / 3.0) + dpow(rb, 8.0 / 3.0)) +
                   (47.0 / 18.0 - (7.0 / 18.0) * del) * sig -
                   (2.5 - del / 18.0) * (saa + sbb) -
                   (del - 11.0) / 9.0 * (ra / rr * saa + rb / rr * sbb);
  const D5 brace = ra * rb * inner - (2.0 / 3.0) * rr2 * sig +
                   ((2.0 / 3.0) * rr2 - ra * ra) * sbb + ((2.0 / 3.0) * rr2 - rb * rb) * saa;
  return -4.0 * kLypA / den * ra * rb / rr - kLypA * kLypB * w * brace;
}

// VWN correlation (S. H. Vosko, L. Wilk, M. Nusair, Can. J. Phys. 58, 1200
// (1980), the Ceperley-Alder fit "V"): with x = sqrt(rs), X(x) = x^2 + b x + c,
// Q = sqrt(4c - b^2),
//   G(x) = A { ln(x^2/X) + 2b/Q atan(Q/(2x + b))
//            - b x0/X(x0) [ln((x - x0)^2/X) + 2(b + 2 x0)/Q atan(Q/(2x + b))] }
// for the paramagnetic (P), ferromagnetic (F) and spin-stiffness (alpha)
// parameter sets, and eps = P + alpha f(z)/f''(0) (1 - z^4) + (F - P) f(z) z^4.
struct VwnSet {
  double A, x0, b, c;
};
__host__ __device__ inline D5 vwn_g(D5 x, const VwnSet &p) {
  const double Q = sqrt(4.0 * p.c - p.b * p.b);
  const double X0 = p.x0 * p.x0 + p.b * p.x0 + p.c;
  const D5 X = x * x + p.b * x + p.c;
  const D5 at = datan(Q / (2.0 * x + p.b));
  return p.A * (dlog(x * x / X) + (2.0 * p.b / Q) * at -
                (p.b * p.x0 / X0) * (dlog((x - p.x0) * (x - p.x0) / X) +
                                     (2.0 * (p.b + 2.0 * p.x0) / Q) * at));
}

// energy density rho eps_c; 1 +- zeta passed as 2 rho_s / rho, floored
__host__ __device__ inline D5 vwn(D5 ra, D5 rb, double zthr) {
  const VwnSet P{0.0310907, -0.10498, 3.72744, 12.9352};
  const VwnSet F{0.01554535, -0.32500, 7.06042, 18.0578};
  const VwnSet Al{-1.0 / (6.0 * 3.14159265358979323846 * 3.14159265358979323846), -0.0047584,
                  1.13107, 13.0045};
  const double fpp = 1.709920934161365617563962776245;  // f''(0) = 4 / (9 (2^(1/3) - 1))
  const D5 rr = ra + rb;
  const D5 x = dsqrt(0.62035049089940001667 / dpow(rr, 1.0 / 3.0));  // sqrt(rs)
  const D5 opz = dmax(2.0 * ra / rr, zthr), omz = dmax(2.0 * rb / rr, zthr);
  const D5 z = (ra - rb) / rr;
  const D5 f = (dpow(opz, 4.0 / 3.0) + dpow(omz, 4.0 / 3.0) - 2.0) / 0.5198420997897464;
  const D5 z4 = z * z * z * z;
  const D5 eP = vwn_g(x, P), eF = vwn_g(x, F), eA = vwn_g(x, Al);
  return rr * (eP + eA * f * (1.0 - z4) / fpp + (eF - eP) * f * z4);
}

// PW91 exchange of an unpolarised density n with gradient square g
__host__ __device__ inline D5 pw91_x_unpol(D5 n, D5 g) {
  constexpr double a = 0.19645, b = 7.7956, c = 0.2743, d = -0.1508, f = 0.004, al = 100.0;
  constexpr double cs = 0.026121172985233605;  // 1 / (4 (3 pi^2)^(2/3))
  const D5 n43 = dpow(n, 4.0 / 3.0);
  const D5 s2 = cs * g / (n43 * n43);
  const D5 sv = dsqrt(s2);
  const D5 s4 = s2 * s2;
  const D5 F = 1.0 + ((c + d * dexp(-al * s2)) * s2 - f * s4) /
                         (1.0 + a * sv * dasinh(b * sv) + f * s4);
  return -0.73855876638202240588 * n43 * F;
}

// y = (beta/g) t2 (1 + u) / (1 + u + u^2), u = A t2, with the partials in
// closed form (dy/dt2 = (beta/g)(1 + 2u)/D^2, dy/dA = -(beta/g) t2^2 u (2 + u)/D^2,
// D = 1 + u + u^2): the quotient differentiated term by term cancels to a
// relative 1/u^2 where the enhancement saturates.
__host__ __device__ inline D5 gga_c_y(D5 t2, D5 A, double bg) {
  D5 y;
  const double u = A.v * t2.v, D = 1.0 + u + u * u;
  const double dt = bg * (1.0 + 2.0 * u) / (D * D);
  const double dA = -bg * t2.v * t2.v * u * (2.0 + u) / (D * D);
  y.v = bg * t2.v * (1.0 + u) / D;
  for (int k = 0; k < 5; ++k) y.d[k] = dt * t2.d[k] + dA * A.d[k];
  return y;
}

// PW91 correlation energy per particle given eps_c (PW_MOD), phi, rs, t^2
__host__ __device__ inline D5 pw91_c_h(D5 ec, D5 phi, D5 rs, D5 t2) {
  constexpr double alpha = 0.09, Cc0 = 4.235e-3, Cxc0 = 2.568e-3, Cx = -0.001667;
  constexpr double nu = 15.755920349483144;  // (16/pi)(3 pi^2)^(1/3)
  constexpr double beta = nu * Cc0;
  constexpr double gam = beta * beta / (2.0 * alpha);
  // -100 (k_s/k_F)^2 / rs = -100 (4/pi) (4/(9 pi))^(1/3)
  constexpr double ha = -66.34364396064501;
  const D5 phi2 = phi * phi, phi3 = phi2 * phi;
  const D5 A = (beta / gam) / (dexp(-1.0 * ec / (gam * phi3)) - 1.0);
  const D5 H0 = gam * phi3 * dlog(1.0 + gga_c_y(t2, A, beta / gam));
  const D5 rs2 = rs * rs;
  const D5 Cxc = (2.568 + 23.266 * rs + 0.007389 * rs2) /
                 (1000.0 * (1.0 + 8.723 * rs + 0.472 * rs2 + 0.07389 * rs2 * rs));
  const D5 H1 = nu * (Cxc - (Cxc0 + 3.0 * Cx / 7.0)) * phi3 * t2 *
                dexp(ha * rs * phi2 * phi2 * t2);
  return H0 + H1;
}

// PW92 energy per particle, PW_MOD parameters (the correlation under PBE
// and PW91): G0 + z^4 f(z) (G1 - G0 + G2/f''(0)) - f(z) G2/f''(0)
__host__ __device__ inline D5 pw92_mod_g(D5 rs, double A, double a1, double b1, double b2,
                                         double b3, double b4) {
  const D5 s = dsqrt(rs);
  const D5 den = 2.0 * A * (b1 * s + b2 * rs + b3 * rs * s + b4 * rs * rs);
  return -2.0 * A * (1.0 + a1 * rs) * dlog(1.0 + 1.0 / den);
}
__host__ __device__ inline D5 pw92_mod_eps(D5 rs, D5 z, D5 opz, D5 omz) {
  const D5 g0 = pw92_mod_g(rs, 0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294);
  const D5 g1 = pw92_mod_g(rs, 0.01554535, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517);
  const D5 g2 = pw92_mod_g(rs, 0.0168869, 0.11125, 10.357, 3.6231, 0.88026, 0.49671);
  const double ifz20 = 1.0 / 1.709920934161365617563962776245;
  const D5 fz = (dpow(opz, 4.0 / 3.0) + dpow(omz, 4.0 / 3.0) - 2.0) * (1.0 / 0.5198420997897464);
  const D5 z2 = z * z;
  const D5 z4 = z2 * z2;
  return g0 + z4 * fz * (g1 - g0 + g2 * ifz20) - fz * g2 * ifz20;
}

// PW91 correlation energy density rho (eps_c + H0 + H1); 1 +- zeta as
// 2 rho_s / rho floored at zthr; stot = sigma_aa + 2 sigma_ab + sigma_bb
__host__ __device__ inline D5 pw91_c(D5 ra, D5 rb, D5 stot, double zthr) {
  constexpr double kPi = 3.14159265358979323846;
  const D5 rr = ra + rb;
  const D5 r13 = dpow(rr, 1.0 / 3.0);
  const D5 rs = 0.62035049089940001667 / r13;
  const D5 opz = dmax(2.0 * ra / rr, zthr), omz = dmax(2.0 * rb / rr, zthr);
  const D5 ec = pw92_mod_eps(rs, (ra - rb) / rr, opz, omz);
  const D5 phi = 0.5 * (dpow(opz, 2.0 / 3.0) + dpow(omz, 2.0 / 3.0));
  // t^2 = stot / (4 phi^2 ks^2 rho^2), ks^2 = 4 kF / pi, kF = (3 pi^2 rho)^(1/3)
  const D5 kF = 3.09366772628013593097 * r13;
  const D5 t2 = stot * kPi / (16.0 * phi * phi * kF * rr * rr);
  return rr * (ec + pw91_c_h(ec, phi, rs, t2));
}

}  // namespace xcb
}  // namespace nkx
