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

}  // namespace xcb
}  // namespace nkx
