// oracle.cpp -- FP64 CPU restatement of the NuKEXC exchange-correlation path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by the
// tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
// by nothing else.  The product (nukexc_amd/, libnukexc_hip.so) never links,
// loads or calls it.
//
// Every function below restates the reference algorithm it cites
// (paths relative to the NuKEXC source tree).  The restatement is written for
// clarity, not speed, except that the point loops are OpenMP-parallel so the
// CPU baseline can run on all host cores.
//
// Pins (tests/test_oracle_*.py):
//   radial grids + Lebedev + capped Becke + collocation
//       latex/Data/convergence_h2plus.csv (S_AB to ~1e-13, npts)
//   solid harmonics      closed forms for l<=2 (test_spherical_harmonics.cxx:
//                         108-145), orthonormality on Lebedev spheres,
//                         finite-difference gradients
//   STO norm / cutoff    S=1 (test_hydrogen.cxx:115), closed-form root check
//   Becke                sum_i p_i = 1 (test_partition.cxx:197-258), midplane
//                         0.5 (test_partition.cxx:140-195), brute force
//   density              int rho = 1, int sigma = 1/(2 pi) for H 1s and
//                         2 zeta^5/pi for He (test_density.cxx:252-372)
//   LDA_X                E = -0.2127415030860106 (test_hydrogen.cxx:815-857)
//   GGA_X_PBE            E, V00 (test_hydrogen.cxx:859-912,
//                         test_helium.cxx:67-121)
//   v_rho, v_sigma       central finite differences of e(rho, sigma)
//   LDA_C_PW(_MOD)       high-density expansion (exact c0, PW92's published
//                         c1, no sqrt(rs) term) of all three parameter sets,
//                         the spin interpolation with the closed-form f''(0),
//                         the Perdew-Zunger fit, LSD correlation of H (exact
//                         density) and He (Hartree-Fock density)
//   GGA_C_PBE            gradient expansion (published beta and gamma =
//                         (1 - ln 2)/pi^2 from the t^2 and t^4 terms),
//                         high-density scaling limit, PBE correlation of H
//                         and He, PW92-mod at zero gradient
//   libxc's density / sigma / zeta thresholds are restated, not pinned.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Radial quadratures.  IntegratorXX (pinned commit 1369be58, not vendored)
// maps Chebyshev second-kind nodes x_i = cos(i pi/(n+1)), i = 1..n, onto
// [0, inf).  The weights include the Jacobian dr/dx and r^2, so that
// sum_i w_i f(r_i) ~ int_0^inf f(r) r^2 dr.  Call site: grid.hpp:247-265.
// ---------------------------------------------------------------------------
void radial_ta(int n, double xi, double alpha, double *r, double *w) {
  // Treutler-Ahlrichs M4: r = xi/ln2 (1+x)^alpha ln(2/(1-x))  (grid.hpp:104-107)
  const double ln2 = std::log(2.0);
  for (int i = 1; i <= n; ++i) {
    const double t = i * kPi / (n + 1);
    const double x = std::cos(t);
    const double wch = kPi / (n + 1) * std::sin(t) * std::sin(t) /
                       std::sqrt(1.0 - x * x);
    const double lg = std::log(2.0 / (1.0 - x));
    const double rr = xi / ln2 * std::pow(1.0 + x, alpha) * lg;
    const double drdx = xi / ln2 *
                        (alpha * std::pow(1.0 + x, alpha - 1.0) * lg +
                         std::pow(1.0 + x, alpha) / (1.0 - x));
    r[i - 1] = rr;
    w[i - 1] = wch * drdx * rr * rr;
  }
}

void radial_becke(int n, double R, double *r, double *w) {
  // Becke: r = R (1+x)/(1-x)  (grid.hpp:231-234, latex/main.tex Eq. Becke_radial)
  for (int i = 1; i <= n; ++i) {
    const double t = i * kPi / (n + 1);
    const double x = std::cos(t);
    const double wch = kPi / (n + 1) * std::sin(t) * std::sin(t) /
                       std::sqrt(1.0 - x * x);
    const double rr = R * (1.0 + x) / (1.0 - x);
    const double drdx = 2.0 * R / ((1.0 - x) * (1.0 - x));
    r[i - 1] = rr;
    w[i - 1] = wch * drdx * rr * rr;
  }
}

// Element tables, grid.hpp:71-102.
const double kTaXi[] = {0.00, 0.8, 0.9, 1.8, 1.4, 1.3, 1.1, 0.9, 0.9, 0.9,
                        0.9,  1.4, 1.3, 1.3, 1.2, 1.1, 1.0, 1.0, 1.0, 1.5,
                        1.4,  1.3, 1.2, 1.2, 1.2, 1.2, 1.2, 1.2, 1.1, 1.1,
                        1.1,  1.1, 1.0, 0.9, 0.9, 0.9, 0.9};
const double kBeckeSlater[] = {0.00, 0.35, 0.35, 1.45, 1.05, 0.85, 0.70,
                               0.65, 0.60, 0.50, 0.35, 1.80, 1.50, 1.25,
                               1.10, 1.00, 1.00, 1.00, 0.70};
constexpr double kAngToBohr = 1.8897261254578281;  // atomic_properties.hpp:49

int periodic_row(int Z) {  // grid.hpp:113-127
  if (Z <= 2) return 1;
  if (Z <= 10) return 2;
  if (Z <= 18) return 3;
  if (Z <= 36) return 4;
  if (Z <= 54) return 5;
  if (Z <= 86) return 6;
  return 7;
}

// ---------------------------------------------------------------------------
// Becke partition, partitioning.hpp:44-55 (mu, p, s), :63-120 (uncapped),
// :126-213 (capped, R_ij -> min(R_ij, 5)).
// ---------------------------------------------------------------------------
inline double becke_mu(double ri, double rj, double Rij) {
  double mu = (ri - rj) / Rij;
  return mu > 1.0 ? 1.0 : (mu < -1.0 ? -1.0 : mu);
}
// x^3 as (x x) x (the reference's std::pow(x, 3) rounds once; near mu = 1 the
// last bit of p decides whether 1 - p(p(p(mu))) is exactly zero, so the
// oracle fixes one rounding order, which the device kernel repeats)
inline double becke_p(double x) { return (1.5 * x) - (0.5 * (x * x * x)); }
inline double becke_s(double f) { return 0.5 * (1.0 - f); }

inline double dist3(const double *a, const double *b) {
  const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}

// ---------------------------------------------------------------------------
// Real solid harmonics S_lm = r^l Y_lm (no Condon-Shortley phase), stored as
// homogeneous polynomials in (x, y, z).  Definition:
// python/precompute_solid_harmonics.py:63-102 (poly_P, poly_A, poly_B,
// solid_harmonic); spherical_harmonics.hpp:212-232.  Expanded here into
// monomial coefficients once, then evaluated with their exact derivatives.
// ---------------------------------------------------------------------------
constexpr int kLMax = 7;
struct Poly {
  double c[kLMax + 1][kLMax + 1][kLMax + 1];  // c[a][b][d] x^a y^b z^d
};
std::vector<Poly> g_harm;  // index l*l + l + m

double fact(int n) {
  double r = 1.0;
  for (int i = 2; i <= n; ++i) r *= i;
  return r;
}
double binom(int n, int k) { return fact(n) / (fact(k) * fact(n - k)); }

void build_harmonics() {
  if (!g_harm.empty()) return;
  g_harm.resize((kLMax + 1) * (kLMax + 1));
  for (int l = 0; l <= kLMax; ++l) {
    for (int m = -l; m <= l; ++m) {
      Poly &P = g_harm[l * l + l + m];
      std::memset(&P, 0, sizeof(P));
      const int am = std::abs(m);
      // A_m (m > 0) = Re (x + i y)^m, B_m (m < 0) = Im (x + i y)^m, 1 (m = 0)
      double ab[kLMax + 1][kLMax + 1] = {};  // ab[p][q] x^p y^q
      if (m == 0) {
        ab[0][0] = 1.0;
      } else {
        for (int p = 0; p <= am; ++p) {
          const int q = am - p;  // power of (i y)
          // i^q: q%4 == 0 -> 1, 1 -> i, 2 -> -1, 3 -> -i
          double re = 0, im = 0;
          switch (q % 4) {
            case 0: re = 1; break;
            case 1: im = 1; break;
            case 2: re = -1; break;
            case 3: im = -1; break;
          }
          ab[p][q] += binom(am, p) * (m > 0 ? re : im);
        }
      }
      const double norm =
          (m == 0) ? std::sqrt((2.0 * l + 1.0) / (4.0 * kPi))
                   : std::sqrt((2.0 * l + 1.0) / (2.0 * kPi));
      const double pref = std::sqrt(fact(l - am) / fact(l + am));
      // Pi_l^m(r, z) = pref * sum_k (-1)^k 2^-l C(l,k) C(2l-2k,l)
      //                 (l-2k)!/(l-2k-m)! r^{2k} z^{l-2k-m}
      for (int k = 0; 2 * k <= l - am; ++k) {
        const double ck = ((k % 2) ? -1.0 : 1.0) * std::ldexp(1.0, -l) *
                          binom(l, k) * binom(2 * l - 2 * k, l) *
                          fact(l - 2 * k) / fact(l - 2 * k - am) * pref *
                          norm;
        const int zp = l - 2 * k - am;
        // r^{2k} = sum_{i+j+h=k} k!/(i! j! h!) x^{2i} y^{2j} z^{2h}
        for (int i = 0; i <= k; ++i)
          for (int j = 0; i + j <= k; ++j) {
            const int h = k - i - j;
            const double mult = fact(k) / (fact(i) * fact(j) * fact(h));
            for (int p = 0; p <= am; ++p)
              for (int q = 0; p + q <= am; ++q) {
                if (ab[p][q] == 0.0) continue;
                P.c[2 * i + p][2 * j + q][2 * h + zp] += ck * mult * ab[p][q];
              }
          }
      }
    }
  }
}

void eval_harmonic(int l, int m, double x, double y, double z, double *val,
                   double *gx, double *gy, double *gz) {
  double px[kLMax + 1], py[kLMax + 1], pz[kLMax + 1];
  px[0] = py[0] = pz[0] = 1.0;
  for (int i = 1; i <= kLMax; ++i) {
    px[i] = px[i - 1] * x;
    py[i] = py[i - 1] * y;
    pz[i] = pz[i - 1] * z;
  }
  const Poly &P = g_harm[l * l + l + m];
  double v = 0, dx = 0, dy = 0, dz = 0;
  for (int a = 0; a <= l; ++a)
    for (int b = 0; a + b <= l; ++b) {
      const int d = l - a - b;
      const double c = P.c[a][b][d];
      if (c == 0.0) continue;
      v += c * px[a] * py[b] * pz[d];
      if (a) dx += c * a * px[a - 1] * py[b] * pz[d];
      if (b) dy += c * b * px[a] * py[b - 1] * pz[d];
      if (d) dz += c * d * px[a] * py[b] * pz[d - 1];
    }
  *val = v;
  if (gx) {
    *gx = dx;
    *gy = dy;
    *gz = dz;
  }
}

// ---------------------------------------------------------------------------
// Functionals (libxc 7.0.0 semantics, unpolarised, spin-scaling form).
// Each returns e = rho*eps and the partial derivatives de/drho, de/dsigma.
// ---------------------------------------------------------------------------
struct XcOut {
  double e, vr, vs;
};

// Dirac/Slater exchange, libxc LDA_X (id 1).
XcOut lda_x(double rho) {
  const double Cx = -0.75 * std::cbrt(3.0 / kPi);
  const double r13 = std::cbrt(rho);
  return {Cx * rho * r13, 4.0 / 3.0 * Cx * r13, 0.0};
}

// PW92 G(rs) and dG/drs (Perdew & Wang, PRB 45, 13244 (1992), Eq. 10).
void pw92_G(double rs, double A, double a1, double b1, double b2, double b3,
            double b4, double *G, double *dG) {
  const double srs = std::sqrt(rs);
  const double Q0 = -2.0 * A * (1.0 + a1 * rs);
  const double Q1 = 2.0 * A * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs);
  const double dQ1 = A * (b1 / srs + 2.0 * b2 + 3.0 * b3 * srs + 4.0 * b4 * rs);
  const double L = std::log(1.0 + 1.0 / Q1);
  *G = Q0 * L;
  *dG = -2.0 * A * a1 * L - Q0 * dQ1 / (Q1 * Q1 + Q1);
}

// eps_c^PW92(rs, zeta = 0) with the libxc LDA_C_PW (id 12) or LDA_C_PW_MOD
// (id 13, used inside GGA_C_PBE) parameter sets.
void pw92_eps(double rs, bool mod, double *eps, double *deps_drs) {
  const double A = mod ? 0.0310907 : 0.031091;
  pw92_G(rs, A, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294, eps, deps_drs);
}

XcOut lda_c_pw(double rho, bool mod) {
  const double rs = std::cbrt(3.0 / (4.0 * kPi * rho));
  double eps, de;
  pw92_eps(rs, mod, &eps, &de);
  // d(rho eps)/drho = eps + rho deps/drs drs/drho, drs/drho = -rs/(3 rho)
  return {rho * eps, eps - rs / 3.0 * de, 0.0};
}

// PBE exchange, libxc GGA_X_PBE (id 101): kappa = 0.804,
// mu = beta pi^2 / 3 with beta = 0.06672455060314922.
XcOut gga_x_pbe(double rho, double sigma) {
  const double kappa = 0.8040;
  const double mu = 0.06672455060314922 * kPi * kPi / 3.0;
  const double Ax = -0.75 * std::cbrt(3.0 / kPi);
  const double c = 1.0 / (4.0 * std::pow(3.0 * kPi * kPi, 2.0 / 3.0));
  const double r13 = std::cbrt(rho);
  const double r43 = rho * r13;
  const double s2 = c * sigma / (r43 * r43);  // sigma / (4 kF^2 rho^2)
  const double den = 1.0 + mu * s2 / kappa;
  const double F = 1.0 + kappa - kappa / den;
  const double dF = mu / (den * den);  // dF/ds2
  const double e = Ax * r43 * F;
  const double vr = 4.0 / 3.0 * Ax * r13 * F + Ax * r43 * dF * (-8.0 / 3.0 * s2 / rho);
  const double vs = Ax * r43 * dF * c / (r43 * r43);
  return {e, vr, vs};
}

// PBE correlation, libxc GGA_C_PBE (id 130), unpolarised (phi = 1):
// H = gamma ln(1 + beta/gamma t^2 (1 + A t^2)/(1 + A t^2 + A^2 t^4)),
// A = beta/gamma / (exp(-eps_c/gamma) - 1), with the PW_MOD eps_c.
XcOut gga_c_pbe(double rho, double sigma) {
  const double beta = 0.06672455060314922;
  const double gamma = (1.0 - std::log(2.0)) / (kPi * kPi);
  const double rs = std::cbrt(3.0 / (4.0 * kPi * rho));
  double ec, dec_drs;
  pw92_eps(rs, true, &ec, &dec_drs);
  const double drs_drho = -rs / (3.0 * rho);
  const double dec = dec_drs * drs_drho;  // d eps_c / d rho
  const double kF = std::cbrt(3.0 * kPi * kPi * rho);
  // t^2 = sigma / (4 ks^2 rho^2), ks^2 = 4 kF / pi
  const double ct = kPi / (16.0 * kF * rho * rho);
  const double t2 = sigma * ct;
  const double ex = std::exp(-ec / gamma);
  const double A = beta / gamma / (ex - 1.0);
  const double dA_dec = beta / gamma * ex / (gamma * (ex - 1.0) * (ex - 1.0));
  const double At2 = A * t2;
  const double num = 1.0 + At2;
  const double dnm = 1.0 + At2 + At2 * At2;
  const double y = beta / gamma * t2 * num / dnm;
  const double H = gamma * std::log(1.0 + y);
  const double dH_dy = gamma / (1.0 + y);
  // dy/dt2 and dy/dA from the quotient rule, simplified so that nothing
  // cancels when A t^2 >> 1:  with u = A t^2,
  //   d/dt2 [t2 (1+u)/(1+u+u^2)] = (1 + 2u) / (1+u+u^2)^2
  //   d/dA  [t2 (1+u)/(1+u+u^2)] = -t2^2 u (2 + u) / (1+u+u^2)^2
  const double bg = beta / gamma;
  const double dy_dt2 = bg * (1.0 + 2.0 * At2) / (dnm * dnm);
  const double dy_dA = -bg * t2 * t2 * At2 * (2.0 + At2) / (dnm * dnm);
  const double dt2_drho = -7.0 / 3.0 * t2 / rho;
  const double dH_drho = dH_dy * (dy_dA * dA_dec * dec + dy_dt2 * dt2_drho);
  const double e = rho * (ec + H);
  const double vr = ec + H + rho * (dec + dH_drho);
  const double vs = rho * dH_dy * dy_dt2 * ct;
  return {e, vr, vs};
}

// B88 exchange (GGA_X_B88, id 106; Becke, PRA 38, 3098 (1988), Eq. 8) and
// LYP correlation (GGA_C_LYP, id 131; Miehlich, Savin, Stoll, Preuss, CPL 157,
// 200 (1989), Eq. 2), spin-resolved energy densities for a generic number type
// (real, or complex for complex-step derivatives).  Parity unpinned against
// libxc (not installed): pinned by Becke's and Lee-Yang-Parr's published atom
// values and the one-electron zero of LYP.
template <typename T>
T b88_spin(T r, T s) {  // one channel: density r, sigma_ss s
  const double beta = 0.0042;
  const double cx = -1.5 * std::cbrt(3.0 / (4.0 * kPi));
  T r43 = std::pow(r, 4.0 / 3.0);
  T x = std::sqrt(s) / r43;
  return r43 * (cx - beta * x * x / (1.0 + 6.0 * beta * x * std::asinh(x)));
}

template <typename T>
T lyp_e(T ra, T rb, T saa, T sab, T sbb) {
  const double a = 0.04918, b = 0.132, c = 0.2533, d = 0.349;
  const double cf = 0.3 * std::pow(3.0 * kPi * kPi, 2.0 / 3.0);
  T n = ra + rb;
  T n13 = std::pow(n, 1.0 / 3.0);
  T g = 1.0 / (1.0 + d / n13);
  T omega = std::exp(-c / n13) * g / std::pow(n, 11.0 / 3.0);
  T delta = (c + d * g) / n13;
  T grad2 = saa + 2.0 * sab + sbb;
  T t_same = std::pow(2.0, 11.0 / 3.0) * cf * (std::pow(ra, 8.0 / 3.0) + std::pow(rb, 8.0 / 3.0));
  T bracket = ra * rb * (t_same + (47.0 - 7.0 * delta) / 18.0 * grad2 -
                         (2.5 - delta / 18.0) * (saa + sbb) -
                         (delta - 11.0) / 9.0 * (ra * saa + rb * sbb) / n);
  bracket += (2.0 / 3.0) * n * n * (saa + sbb - grad2) - ra * ra * sbb - rb * rb * saa;
  return -4.0 * a * g * ra * rb / n - a * b * omega * bracket;
}

// VWN correlation (LDA_C_VWN, id 7; Vosko, Wilk, Nusair, Can. J. Phys. 58,
// 1200 (1980), fit V) in rs: per parameter set (A, x0, b, c),
// G = A [ln(rs/X) + 2b/Q atan(Q/(2 sqrt rs + b)) - b x0/X0 (ln((sqrt rs - x0)^2/X)
// + 2(b + 2 x0)/Q atan(Q/(2 sqrt rs + b)))], X = rs + b sqrt rs + c; spin
// interpolation through f(z) and the stiffness set.  Pinned against PW92
// (a separate fit of the same Ceperley-Alder data) and the z = 1 limit.
template <typename T>
T vwn_set(T rs, double A, double x0, double b, double c) {
  const double Q = std::sqrt(4.0 * c - b * b), X0 = x0 * x0 + b * x0 + c;
  T sx = std::sqrt(rs);
  T X = rs + b * sx + c;
  T at = std::atan(Q / (2.0 * sx + b));
  return A * (std::log(rs / X) + 2.0 * b / Q * at -
              b * x0 / X0 * (std::log((sx - x0) * (sx - x0) / X) + 2.0 * (b + 2.0 * x0) / Q * at));
}

template <typename T>
T vwn_e(T ra, T rb) {
  T n = ra + rb;
  T rs = std::pow(3.0 / (4.0 * kPi * n), 1.0 / 3.0);
  T opz = T(2.0) * ra / n, omz = T(2.0) * rb / n;
  if (std::real(opz) < 2.220446049250313e-16) opz = 2.220446049250313e-16;
  if (std::real(omz) < 2.220446049250313e-16) omz = 2.220446049250313e-16;
  T z = (ra - rb) / n;
  T f = (std::pow(opz, 4.0 / 3.0) + std::pow(omz, 4.0 / 3.0) - 2.0) /
        (2.0 * std::cbrt(2.0) - 2.0);
  const double fpp = 4.0 / (9.0 * (std::cbrt(2.0) - 1.0));
  T eP = vwn_set(rs, 0.0310907, -0.10498, 3.72744, 12.9352);
  T eF = vwn_set(rs, 0.01554535, -0.32500, 7.06042, 18.0578);
  T eA = vwn_set(rs, -1.0 / (6.0 * kPi * kPi), -0.0047584, 1.13107, 13.0045);
  T z4 = z * z * z * z;
  return n * (eP + eA * f / fpp * (1.0 - z4) + (eF - eP) * f * z4);
}

// PW91 exchange (GGA_X_PW91, id 109; Perdew et al., PRB 46, 6671 (1992),
// Eq. 8) of an unpolarised density n with gradient square g:
//   A_x n^(4/3) [1 + ((c + d exp(-alpha s^2)) s^2 - f s^4) / (1 + a s asinh(b s) + f s^4)],
// s^2 = g / (4 (3 pi^2)^(2/3) n^(8/3)), with libxc's parameters.
template <typename T>
T pw91_x_n(T n, T g) {
  const double a = 0.19645, b = 7.7956, c = 0.2743, d = -0.1508, f = 0.004, alpha = 100.0;
  const double Ax = -0.75 * std::cbrt(3.0 / kPi);
  T n43 = std::pow(n, 4.0 / 3.0);
  T s2 = g / (4.0 * std::pow(3.0 * kPi * kPi, 2.0 / 3.0) * n43 * n43);
  T s = std::sqrt(s2);
  T num = (c + d * std::exp(-alpha * s2)) * s2 - f * s2 * s2;
  T den = 1.0 + a * s * std::asinh(b * s) + f * s2 * s2;
  return Ax * n43 * (1.0 + num / den);
}

// defined with the polarised functionals below
template <typename T>
T pw92_pol(T rs, T z, T opz, T omz, bool mod);
template <typename T>
T floor_eps(T x);
inline double pbe_y(double t2, double A, double bg);
inline std::complex<double> pbe_y(std::complex<double> t2, std::complex<double> A, double bg);

// PW91 correlation (GGA_C_PW91, id 134; same paper, Eqs. 13-15): energy
// density rho (eps_c + H0 + H1) with eps_c the PW_MOD PW92 energy,
//   H0 = phi^3 beta^2/(2 alpha) ln(1 + 2 alpha/beta (t^2 + A t^4)/(1 + A t^2 + A^2 t^4)),
//   A  = 2 alpha/beta / (exp(-2 alpha eps_c/(phi^3 beta^2)) - 1),
//   H1 = nu (C_xc(rs) - C_xc(0) - 3 C_x / 7) phi^3 t^2 exp(-100 phi^4 (ks/kF)^2 t^2),
// nu = 16/pi (3 pi^2)^(1/3), beta = nu C_c(0), C_c(0) = 4.235e-3,
// C_x = -1.667e-3, alpha = 0.09, C_xc the Rasolt-Geldart Pade (PRB 34, 1325
// (1986)): (2.568 + 23.266 rs + 0.007389 rs^2)/(1000 (1 + 8.723 rs + 0.472 rs^2
// + 0.07389 rs^3)).  t^2 = |grad rho|^2 / (2 phi ks rho)^2 as for PBE.
template <typename T>
T pw91_c_e(T ra, T rb, T stot) {
  const double alpha = 0.09, Cc0 = 4.235e-3, Cx = -1.667e-3;
  const double nu = 16.0 / kPi * std::cbrt(3.0 * kPi * kPi);
  const double beta = nu * Cc0;
  T rho = ra + rb;
  T z = (ra - rb) / rho;
  T opz = 2.0 * ra / rho, omz = 2.0 * rb / rho;
  T rs = std::pow(3.0 / (4.0 * kPi * rho), 1.0 / 3.0);
  T ec = pw92_pol(rs, z, opz, omz, true);
  T phi = 0.5 * (std::pow(floor_eps(opz), 2.0 / 3.0) + std::pow(floor_eps(omz), 2.0 / 3.0));
  T kF = std::pow(3.0 * kPi * kPi * rho, 1.0 / 3.0);
  T ks2 = 4.0 * kF / kPi;
  T t2 = stot / (4.0 * phi * phi * ks2 * rho * rho);
  T phi3 = phi * phi * phi;
  T A = (2.0 * alpha / beta) / (std::exp(-2.0 * alpha * ec / (phi3 * beta * beta)) - 1.0);
  T H0 = phi3 * beta * beta / (2.0 * alpha) *
         std::log(1.0 + pbe_y(t2, A, 2.0 * alpha / beta));
  auto cxc = [](T r) {
    return (2.568 + 23.266 * r + 0.007389 * r * r) /
           (1000.0 * (1.0 + 8.723 * r + 0.472 * r * r + 0.07389 * r * r * r));
  };
  T H1 = nu * (cxc(rs) - cxc(T(0.0)) - 3.0 * Cx / 7.0) * phi3 * t2 *
         std::exp(-100.0 * phi3 * phi * (ks2 / (kF * kF)) * t2);
  return rho * (ec + H0 + H1);
}

template <typename T>
T blyp_e(int id, T ra, T rb, T saa, T sab, T sbb, bool a_on, bool b_on) {
  if (id == 7) return vwn_e(ra, rb);
  if (id == 134) return pw91_c_e(ra, rb, saa + 2.0 * sab + sbb);
  if (id == 109) {
    T e = 0.0;
    if (a_on) e += 0.5 * pw91_x_n(T(2.0) * ra, T(4.0) * saa);
    if (b_on) e += 0.5 * pw91_x_n(T(2.0) * rb, T(4.0) * sbb);
    return e;
  }
  if (id == 106) {
    T e = 0.0;
    if (a_on) e += b88_spin(ra, saa);
    if (b_on) e += b88_spin(rb, sbb);
    return e;
  }
  return lyp_e(ra, rb, saa, sab, sbb);
}

// Unpolarised B88 / LYP: the spin-resolved form at rho_s = rho / 2,
// sigma_ss = sigma_ab = sigma / 4, differentiated by complex step along rho
// and along sigma.
XcOut blyp_unpol(int id, double rho, double sigma) {
  using C = std::complex<double>;
  const double h = 1e-30;
  auto ev = [&](C r, C s) { return blyp_e<C>(id, 0.5 * r, 0.5 * r, 0.25 * s, 0.25 * s, 0.25 * s,
                                             true, true); };
  const C e0 = ev(C(rho, 0.0), C(sigma, 0.0));
  const C er = ev(C(rho, h), C(sigma, 0.0));
  const C es = ev(C(rho, 0.0), C(sigma, h));
  return {e0.real(), er.imag() / h, es.imag() / h};
}

// libxc thresholds: points with rho below dens_threshold contribute nothing;
// sigma is floored at sigma_threshold^2 with sigma_threshold =
// dens_threshold^(4/3).  Default dens_threshold per functional follows libxc's
// xc_func_info tables.  Parity unpinned (libxc not installed).
double default_dens_threshold(int id) {
  switch (id) {
    case 130:
    case 134: return 1e-12;
    default: return 1e-15;
  }
}

bool xc_eval_point(int id, double rho, double sigma, double thr, XcOut *o) {
  if (!(rho >= thr)) {
    *o = {0, 0, 0};
    return true;
  }
  const double sthr = std::pow(thr, 4.0 / 3.0);
  const double s = std::max(sigma, sthr * sthr);
  switch (id) {
    case 1: *o = lda_x(rho); return true;
    case 12: *o = lda_c_pw(rho, false); return true;
    case 13: *o = lda_c_pw(rho, true); return true;
    case 101: *o = gga_x_pbe(rho, s); return true;
    case 130: *o = gga_c_pbe(rho, s); return true;
    case 7:
    case 106:
    case 109:
    case 131:
    case 134: *o = blyp_unpol(id, rho, s); return true;
    default: return false;
  }
}

bool is_gga(int id) {
  return id == 101 || id == 130 || id == 106 || id == 131 || id == 109 || id == 134;
}

// ---------------------------------------------------------------------------
// Spin-polarised functionals (libxc XC_POLARIZED, called by compute_lsda /
// compute_gga_lsda, xc_integrals.hpp:203, :617).  Energy densities are written
// for a generic number type; the potentials come from complex-step
// differentiation, d e / d x = Im e(x + i h) / h with h = 1e-30, which is
// exact to rounding (no subtractive cancellation).
// Exchange: spin scaling E_x[a, b] = (E_x[2a] + E_x[2b]) / 2.
// PW92 (Perdew & Wang 1992, Eq. 8 with the libxc parameter sets):
//   eps = G0 - G2 f(z) (1 - z^4) / f''(0) + (G1 - G0) f(z) z^4.
// PBE correlation (Perdew, Burke & Ernzerhof 1996, Eqs. 7-8) with
//   phi = ((1+z)^(2/3) + (1-z)^(2/3)) / 2, t^2 = |grad rho|^2 / (2 phi ks rho)^2.
// 1 +- z are formed as 2 rho_a / rho, 2 rho_b / rho and floored at DBL_EPSILON
// (libxc's zeta threshold).
// ---------------------------------------------------------------------------
using cplx = std::complex<double>;

template <typename T>
T pw92_g(T rs, double A, double a1, double b1, double b2, double b3, double b4) {
  T srs = std::sqrt(rs);
  T Q1 = 2.0 * A * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs);
  return -2.0 * A * (1.0 + a1 * rs) * std::log(1.0 + 1.0 / Q1);
}

template <typename T>
T floor_eps(T x) { return std::real(x) < 2.220446049250313e-16 ? T(2.220446049250313e-16) : x; }

template <typename T>
T pw92_pol(T rs, T z, T opz, T omz, bool mod) {
  T G0 = pw92_g(rs, mod ? 0.0310907 : 0.031091, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294);
  T G1 = pw92_g(rs, mod ? 0.01554535 : 0.015545, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517);
  T G2 = pw92_g(rs, mod ? 0.0168869 : 0.016887, 0.11125, 10.357, 3.6231, 0.88026, 0.49671);
  const double fpp0 = mod ? 1.709920934161365617563962776245 : 1.709921;
  T f = (std::pow(floor_eps(opz), 4.0 / 3.0) + std::pow(floor_eps(omz), 4.0 / 3.0) - 2.0) /
        (std::pow(2.0, 4.0 / 3.0) - 2.0);
  T z4 = z * z * z * z;
  return G0 - G2 * f * (1.0 - z4) / fpp0 + (G1 - G0) * f * z4;
}

// y = bg t2 (1 + u) / (1 + u + u^2), u = A t2.  For the complex step the
// imaginary part is formed from the closed-form partials dy/dt2 = bg (1 + 2u)
// / D^2 and dy/dA = -bg t2^2 u (2 + u) / D^2 (D = 1 + u + u^2): complex
// arithmetic on the quotient would cancel to a relative 1/u^2 where the
// enhancement saturates.
inline double pbe_y(double t2, double A, double bg) {
  const double u = A * t2;
  return bg * t2 * (1.0 + u) / (1.0 + u + u * u);
}
inline cplx pbe_y(cplx t2, cplx A, double bg) {
  const double t = t2.real(), a = A.real(), u = a * t, D = 1.0 + u + u * u;
  const double dt = bg * (1.0 + 2.0 * u) / (D * D);
  const double da = -bg * t * t * u * (2.0 + u) / (D * D);
  return cplx(pbe_y(t, a, bg), dt * t2.imag() + da * A.imag());
}

template <typename T>
T pbe_c_pol(T ra, T rb, T stot) {
  const double beta = 0.06672455060314922, gamma = (1.0 - std::log(2.0)) / (kPi * kPi);
  T rho = ra + rb;
  T z = (ra - rb) / rho;
  T opz = 2.0 * ra / rho, omz = 2.0 * rb / rho;
  T rs = std::pow(3.0 / (4.0 * kPi * rho), 1.0 / 3.0);
  T ec = pw92_pol(rs, z, opz, omz, true);
  T phi = 0.5 * (std::pow(floor_eps(opz), 2.0 / 3.0) + std::pow(floor_eps(omz), 2.0 / 3.0));
  T kF = std::pow(3.0 * kPi * kPi * rho, 1.0 / 3.0);
  T ks2 = 4.0 * kF / kPi;
  T t2 = stot / (4.0 * phi * phi * ks2 * rho * rho);
  T g3 = gamma * phi * phi * phi;
  T A = beta / gamma / (std::exp(-ec / g3) - 1.0);
  T H = g3 * std::log(1.0 + pbe_y(t2, A, beta / gamma));
  return rho * (ec + H);
}

template <typename T>
T x_unpol(int id, T n, T s) {
  const double Ax = -0.75 * std::cbrt(3.0 / kPi);
  T n43 = std::pow(n, 4.0 / 3.0);
  if (id == 1) return Ax * n43;
  const double kappa = 0.8040, mu = 0.06672455060314922 * kPi * kPi / 3.0;
  T s2 = s / (4.0 * std::pow(3.0 * kPi * kPi, 2.0 / 3.0) * n43 * n43);
  return Ax * n43 * (1.0 + kappa - kappa / (1.0 + mu * s2 / kappa));
}

// Energy density of one functional at an input point already clamped by
// clamp_polarized: (ra, rb, saa, sab, sbb).  Spin-scaled exchange drops a
// channel whose density is at or below the threshold (libxc screen_dens_zeta).
template <typename T>
T e_polarized(int id, const T *x, double thr) {
  T ra = x[0], rb = x[1], saa = x[2], sab = x[3], sbb = x[4];
  switch (id) {
    case 1:
    case 101: {
      T e = 0.0;
      if (std::real(ra) > thr) e += 0.5 * x_unpol(id, T(2.0) * ra, T(4.0) * saa);
      if (std::real(rb) > thr) e += 0.5 * x_unpol(id, T(2.0) * rb, T(4.0) * sbb);
      return e;
    }
    case 12:
    case 13: {
      T rho = ra + rb;
      T rs = std::pow(3.0 / (4.0 * kPi * rho), 1.0 / 3.0);
      return rho * pw92_pol(rs, (ra - rb) / rho, T(2.0) * ra / rho, T(2.0) * rb / rho, id == 13);
    }
    case 130:
      return pbe_c_pol(ra, rb, saa + 2.0 * sab + sbb);
    case 7:
    case 106:
    case 109:
    case 131:
    case 134:
      return blyp_e(id, ra, rb, saa, sab, sbb, std::real(ra) > thr, std::real(rb) > thr);
  }
  return T(0.0);
}

// libxc's polarised input clamps: rho_s >= thr, sigma_ss >= thr^(8/3),
// |sigma_ab| <= (sigma_aa + sigma_bb) / 2.  Potentials are the derivatives at
// the clamped point.
void clamp_polarized(double *x, double thr) {
  const double st = std::pow(thr, 8.0 / 3.0);
  x[0] = std::max(x[0], thr);
  x[1] = std::max(x[1], thr);
  x[2] = std::max(x[2], st);
  x[4] = std::max(x[4], st);
  const double ave = 0.5 * (x[2] + x[4]);
  x[3] = std::min(std::max(x[3], -ave), ave);
}

}  // namespace

extern "C" {

int orc_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// Radial rule: scheme 0 = Treutler-Ahlrichs (alpha), 1 = Becke.
int orc_radial(int scheme, int n, double R, double alpha, double *r, double *w) {
  if (n <= 0) return 1;
  if (scheme == 0)
    radial_ta(n, R, alpha, r, w);
  else if (scheme == 1)
    radial_becke(n, R, r, w);
  else
    return 1;
  return 0;
}

// Per-element radial scale R: TA xi or half the Bragg-Slater radius (grid.hpp:227-235).
double orc_radial_scale(int scheme, int Z) {
  if (scheme == 0) return (Z >= 0 && Z < (int)(sizeof(kTaXi) / sizeof(double))) ? kTaXi[Z] : 1.0;
  return (Z >= 0 && Z < (int)(sizeof(kBeckeSlater) / sizeof(double)))
             ? 0.5 * kBeckeSlater[Z] * kAngToBohr
             : 1.0;
}

// Becke partition: w[g] *= P_owner / sum_i P_i.  capped != 0 selects the
// partition_becke_team form (R_ij -> min(R_ij, 5)), 0 the partition_becke form.
void orc_becke(int natoms, const double *centers, long npts, const double *pts,
               const int *owner, double *w, int capped) {
  std::vector<double> R((size_t)natoms * natoms);
  for (int i = 0; i < natoms; ++i)
    for (int j = 0; j < natoms; ++j) {
      double d = dist3(centers + 3 * i, centers + 3 * j);
      R[(size_t)i * natoms + j] = capped ? std::min(d, 5.0) : d;
    }
#pragma omp parallel
  {
    std::vector<double> r(natoms);
#pragma omp for schedule(static)
    for (long g = 0; g < npts; ++g) {
      for (int i = 0; i < natoms; ++i) r[i] = dist3(pts + 3 * g, centers + 3 * i);
      double wown = 0, norm = 0;
      for (int i = 0; i < natoms; ++i) {
        double Pi = 1.0;
        for (int j = 0; j < natoms; ++j) {
          if (i == j) continue;
          const double mu = becke_mu(r[i], r[j], R[(size_t)i * natoms + j]);
          Pi *= becke_s(becke_p(becke_p(becke_p(mu))));
        }
        if (i == owner[g]) wown = Pi;
        norm += Pi;
      }
      w[g] *= wown / norm;
    }
  }
}

// make_flat_grid (grid.hpp:172-348) with the angular rule supplied by the
// caller (ang: nang rows of x, y, z, w with sum w = 4 pi).  Outputs must hold
// natoms * nrad_max * nang entries.  Returns the number of points kept
// (w > w_thr), in generation order, or -1 on a negative weight.
// Angular pruning (grid.hpp:164-166, 256-265; IntegratorXX's Treutler and
// Robust schemes): counting radial shells from the nucleus outwards, shells
// 0 .. nr/3 (Treutler) or 0 .. nr/4 (Robust) use the low rule, shells up to
// nr/2 the medium rule, the rest the full rule.  Pinned by the reference's
// pruned point counts (latex/Data/convergence_pruning.csv).
int prune_rule(int pruning, int k, int nr) {  // k: shell rank from the nucleus
  if (pruning == 0) return 0;
  const int b1 = (pruning == 1 ? nr / 3 : nr / 4) + 1, b2 = nr / 2 + 1;
  return k < b1 ? 1 : (k < b2 ? 2 : 0);
}

// rules: angs[q] with nangs[q] points for q = 0 full, 1 low, 2 medium.
long orc_make_flat_grid_pruned(int natoms, const double *xyz, const int *Z, int nrad,
                               const int *nangs, const double *const *angs, int pruning,
                               int scheme, double alpha, double w_thr, int capped,
                               int pyscf_sizing, double *pts_out, double *w_out,
                               int *owner_out) {
  std::vector<double> P, W;
  std::vector<int> O;
  for (int a = 0; a < natoms; ++a) {
    int nr = nrad;
    if (pyscf_sizing) {
      static const double ratio[7] = {1.0, 1.5, 1.6, 1.8, 1.9, 2.0, 2.1};
      nr = (int)(nrad * ratio[periodic_row(Z[a]) - 1] + 0.5);
    }
    std::vector<double> r(nr), wr(nr);
    orc_radial(scheme, nr, orc_radial_scale(scheme, Z[a]), alpha, r.data(), wr.data());
    for (int i = 0; i < nr; ++i) {
      const int q = prune_rule(pruning, r[0] < r[nr - 1] ? i : nr - 1 - i, nr);
      const int nang = nangs[q];
      const double *ang = angs[q];
      for (int j = 0; j < nang; ++j) {
        P.push_back(xyz[3 * a + 0] + r[i] * ang[4 * j + 0]);
        P.push_back(xyz[3 * a + 1] + r[i] * ang[4 * j + 1]);
        P.push_back(xyz[3 * a + 2] + r[i] * ang[4 * j + 2]);
        const double ww = wr[i] * ang[4 * j + 3];
        if (ww < 0) return -1;
        W.push_back(ww);
        O.push_back(a);
      }
    }
  }
  const long n = (long)W.size();
  orc_becke(natoms, xyz, n, P.data(), O.data(), W.data(), capped);
  long k = 0;
  for (long g = 0; g < n; ++g) {
    if (W[g] > w_thr) {
      pts_out[3 * k + 0] = P[3 * g + 0];
      pts_out[3 * k + 1] = P[3 * g + 1];
      pts_out[3 * k + 2] = P[3 * g + 2];
      w_out[k] = W[g];
      owner_out[k] = O[g];
      ++k;
    }
  }
  return k;
}

long orc_make_flat_grid(int natoms, const double *xyz, const int *Z, int nrad,
                        int nang, const double *ang, int scheme, double alpha,
                        double w_thr, int capped, int pyscf_sizing,
                        double *pts_out, double *w_out, int *owner_out) {
  const int nangs[3] = {nang, nang, nang};
  const double *angs[3] = {ang, ang, ang};
  return orc_make_flat_grid_pruned(natoms, xyz, Z, nrad, nangs, angs, 0, scheme, alpha, w_thr,
                                   capped, pyscf_sizing, pts_out, w_out, owner_out);
}

// STO normalisation N = (2 zeta)^{n+1/2} / sqrt((2n)!)  (stobasis.hpp:200-201)
// and the cutoff radius of compute_cutoff (stobasis.hpp:90-173): the root of
// ln N + (n-1) ln r - zeta r = ln tol beyond the radial peak, by safeguarded
// Newton steps with bisection fallback.
void orc_sto_norm_cutoff(int n, double zeta, double tol, double *norm_out,
                         double *cut_out) {
  const double norm = std::pow(2.0 * zeta, n + 0.5) / std::sqrt(fact(2 * n));
  *norm_out = norm;
  if (norm <= tol) {
    *cut_out = 0.0;
    return;
  }
  const double lt = std::log(tol), ln = std::log(norm);
  auto g = [&](double r) { return ln + (n - 1) * std::log(r) - zeta * r - lt; };
  const double peak = (n > 1) ? (n - 1) / zeta : 0.0;
  double r = std::max(peak + (ln - lt) / zeta, peak + 1e-6);
  double lo = peak + 1e-10, hi = r;
  for (int k = 0; k < 64; ++k) {
    if (g(hi) <= 0.0) break;
    hi *= 2.0;
  }
  r = hi;
  for (int k = 0; k < 50; ++k) {
    const double gv = g(r);
    if (gv > 0.0)
      lo = r;
    else
      hi = r;
    const double gp = (n - 1) / r - zeta;
    double nx;
    if (gp >= 0.0) {
      nx = 0.5 * (lo + hi);
    } else {
      nx = r - gv / gp;
      if (nx <= lo || nx >= hi) nx = 0.5 * (lo + hi);
    }
    const double d = std::fabs(nx - r);
    r = nx;
    if (d < 1e-10 * r) break;
  }
  *cut_out = r;
}

// Solid harmonic value and gradient (l <= 7).
int orc_solid_harmonic(int l, int m, double x, double y, double z, double *out4) {
  if (l < 0 || l > kLMax || m < -l || m > l) return 1;
  build_harmonics();
  eval_harmonic(l, m, x, y, z, out4, out4 + 1, out4 + 2, out4 + 3);
  return 0;
}

// Collocation Phi(mu, g) and its gradient (stobasis.hpp:494-584).  Outputs are
// row-major [nbf][npts] (the reference's LayoutLeft (N_bf, N_pts) with mu
// fastest is the transpose; the tests transpose where they compare).  Any
// gradient pointer may be NULL.
void orc_collocation(int nbf, const int *n, const int *l, const int *m,
                     const double *zeta, const double *norm, const double *O,
                     long npts, const double *pts, double *phi, double *gx,
                     double *gy, double *gz) {
  build_harmonics();
#pragma omp parallel for schedule(static)
  for (long g = 0; g < npts; ++g) {
    for (int i = 0; i < nbf; ++i) {
      const double dx = pts[3 * g + 0] - O[3 * i + 0];
      const double dy = pts[3 * g + 1] - O[3 * i + 1];
      const double dz = pts[3 * g + 2] - O[3 * i + 2];
      const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
      const int k = n[i] - l[i] - 1;
      const double rk = (k == 0) ? 1.0 : std::pow(r, (double)k);
      const double R = norm[i] * rk * std::exp(-zeta[i] * r);
      double S, Sx, Sy, Sz;
      eval_harmonic(l[i], m[i], dx, dy, dz, &S, &Sx, &Sy, &Sz);
      const size_t idx = (size_t)i * npts + g;
      if (phi) phi[idx] = R * S;
      if (gx) {
        const double common = (k / r - zeta[i]) * R / r;
        gx[idx] = R * Sx + S * (dx * common);
        gy[idx] = R * Sy + S * (dy * common);
        gz[idx] = R * Sz + S * (dz * common);
      }
    }
  }
}

// Density from a density matrix P (row-major nbf x nbf, symmetric):
// rho_g = sum_uv Phi_ug P_uv Phi_vg, grad rho_g = 2 sum_uv dPhi_ug P_uv Phi_vg.
// This equals the MO route of density.hpp:188-255 when P = C diag(f) C^T.
// grad may be NULL (LDA).  sigma = |grad rho|^2 when grad is given.
void orc_density_P(int nbf, long npts, const double *phi, const double *gx,
                   const double *gy, const double *gz, const double *P,
                   double *rho, double *grad3, double *sigma) {
#pragma omp parallel
  {
    std::vector<double> X(nbf);
#pragma omp for schedule(static)
    for (long g = 0; g < npts; ++g) {
      for (int u = 0; u < nbf; ++u) {
        double s = 0;
        for (int v = 0; v < nbf; ++v) s += P[(size_t)u * nbf + v] * phi[(size_t)v * npts + g];
        X[u] = s;
      }
      double r = 0, a = 0, b = 0, c = 0;
      for (int u = 0; u < nbf; ++u) {
        r += phi[(size_t)u * npts + g] * X[u];
        if (grad3) {
          a += gx[(size_t)u * npts + g] * X[u];
          b += gy[(size_t)u * npts + g] * X[u];
          c += gz[(size_t)u * npts + g] * X[u];
        }
      }
      rho[g] = r;
      if (grad3) {
        grad3[3 * g + 0] = 2.0 * a;
        grad3[3 * g + 1] = 2.0 * b;
        grad3[3 * g + 2] = 2.0 * c;
        if (sigma) sigma[g] = 4.0 * (a * a + b * b + c * c);
      }
    }
  }
}

// MO route, density.hpp:150-186 and :188-255: psi = C^T Phi,
// rho = sum_i f_i psi_i^2, grad rho = 2 sum_i f_i psi_i grad psi_i.
// C is row-major [nbf][nocc] (the reference's LayoutLeft (N_bf, N_occ)).
void orc_density_C(int nbf, int nocc, long npts, const double *phi,
                   const double *gx, const double *gy, const double *gz,
                   const double *C, const double *f, double *rho,
                   double *grad3, double *sigma) {
#pragma omp parallel
  {
    std::vector<double> ps(4 * (size_t)nocc);
#pragma omp for schedule(static)
    for (long g = 0; g < npts; ++g) {
      std::fill(ps.begin(), ps.end(), 0.0);
      for (int u = 0; u < nbf; ++u) {
        const double p0 = phi[(size_t)u * npts + g];
        const double p1 = grad3 ? gx[(size_t)u * npts + g] : 0;
        const double p2 = grad3 ? gy[(size_t)u * npts + g] : 0;
        const double p3 = grad3 ? gz[(size_t)u * npts + g] : 0;
        for (int i = 0; i < nocc; ++i) {
          const double c = C[(size_t)u * nocc + i];
          ps[i] += c * p0;
          ps[nocc + i] += c * p1;
          ps[2 * nocc + i] += c * p2;
          ps[3 * nocc + i] += c * p3;
        }
      }
      double r = 0, a = 0, b = 0, cc = 0;
      for (int i = 0; i < nocc; ++i) {
        r += f[i] * ps[i] * ps[i];
        a += f[i] * ps[nocc + i] * ps[i];
        b += f[i] * ps[2 * nocc + i] * ps[i];
        cc += f[i] * ps[3 * nocc + i] * ps[i];
      }
      rho[g] = r;
      if (grad3) {
        grad3[3 * g + 0] = 2.0 * a;
        grad3[3 * g + 1] = 2.0 * b;
        grad3[3 * g + 2] = 2.0 * cc;
        if (sigma) sigma[g] = 4.0 * (a * a + b * b + cc * cc);
      }
    }
  }
}

// Functional evaluation, summed over the listed functionals (the reference's
// driver sums exchange and correlation, standalone_helpers.hpp:232-259).
// exc is energy per particle (libxc convention: e = rho * exc).
// dens_thr <= 0 selects each functional's libxc default.  vsigma may be NULL
// for LDA-only lists.  Returns nonzero for an unknown id.
int orc_xc_eval(int nfunc, const int *ids, double dens_thr, long npts,
                const double *rho, const double *sigma, double *exc,
                double *vrho, double *vsigma) {
  for (int k = 0; k < nfunc; ++k) {
    XcOut o;
    if (!xc_eval_point(ids[k], 1.0, 1.0, 0.5, &o)) return 1;
  }
#pragma omp parallel for schedule(static)
  for (long g = 0; g < npts; ++g) {
    double e = 0, vr = 0, vs = 0;
    for (int k = 0; k < nfunc; ++k) {
      const double thr = dens_thr > 0 ? dens_thr : default_dens_threshold(ids[k]);
      XcOut o;
      xc_eval_point(ids[k], rho[g], sigma ? sigma[g] : 0.0, thr, &o);
      e += o.e;
      vr += o.vr;
      vs += o.vs;
    }
    exc[g] = rho[g] > 0 ? e / rho[g] : 0.0;
    vrho[g] = vr;
    if (vsigma) vsigma[g] = vs;
  }
  return 0;
}

int orc_is_gga(int id) { return is_gga(id) ? 1 : 0; }

// Polarised functional evaluation, libxc layout: rho[2 g + s],
// sigma[3 g + {aa, ab, bb}], vrho[2 g + s], vsigma[3 g + k]; exc per particle
// (the caller integrates w (rho_a + rho_b) exc, xc_integrals.hpp:251-258).
int orc_xc_eval_polarized(int nfunc, const int *ids, double dens_thr, long npts,
                          const double *rho, const double *sigma, double *exc, double *vrho,
                          double *vsigma) {
  for (int k = 0; k < nfunc; ++k) {
    XcOut o;
    if (!xc_eval_point(ids[k], 1.0, 1.0, 0.5, &o)) return 1;
  }
  const double h = 1e-30;
#pragma omp parallel for schedule(static)
  for (long g = 0; g < npts; ++g) {
    double x[5] = {rho[2 * g], rho[2 * g + 1], sigma ? sigma[3 * g] : 0.0,
                   sigma ? sigma[3 * g + 1] : 0.0, sigma ? sigma[3 * g + 2] : 0.0};
    double e = 0, d[5] = {0, 0, 0, 0, 0};
    for (int k = 0; k < nfunc; ++k) {
      const double thr = dens_thr > 0 ? dens_thr : default_dens_threshold(ids[k]);
      if (!(x[0] + x[1] >= thr)) continue;
      double xc[5] = {x[0], x[1], x[2], x[3], x[4]};
      clamp_polarized(xc, thr);
      cplx cx[5];
      for (int j = 0; j < 5; ++j) cx[j] = cplx(xc[j], 0.0);
      // libxc returns the energy per particle of the clamped point
      e += std::real(e_polarized(ids[k], cx, thr)) / (xc[0] + xc[1]);
      for (int j = 0; j < 5; ++j) {
        cx[j] = cplx(xc[j], h);
        d[j] += std::imag(e_polarized(ids[k], cx, thr)) / h;
        cx[j] = cplx(xc[j], 0.0);
      }
    }
    exc[g] = e;
    vrho[2 * g] = d[0];
    vrho[2 * g + 1] = d[1];
    if (vsigma) {
      vsigma[3 * g] = d[2];
      vsigma[3 * g + 1] = d[3];
      vsigma[3 * g + 2] = d[4];
    }
  }
  return 0;
}

// E_xc, N_el and V_xc from collocation data (compute_lda xc_integrals.hpp:
// 73-166, compute_gga :424-541): Z_ug = w_g (1/2 Phi_ug vrho_g
// + 2 (grad rho . grad Phi_ug) vsigma_g), Vbar = Phi Z^T, V = Vbar + Vbar^T,
// E = sum_g w rho exc.  Density from P (C == NULL) or from (C, f, nocc).
// grad pointers NULL => LDA route.  V row-major nbf x nbf, accumulated (+=).
int orc_vxc_from_collocation(int nbf, long npts, const double *phi,
                             const double *gx, const double *gy,
                             const double *gz, const double *w,
                             const double *P, int nocc, const double *C,
                             const double *f, int nfunc, const int *ids,
                             double dens_thr, double *E, double *N, double *V) {
  const bool gga = gx != nullptr;
  for (int k = 0; k < nfunc; ++k)
    if (is_gga(ids[k]) && !gga) return 2;  // GGA needs gradients
  std::vector<double> rho(npts), grad(gga ? 3 * npts : 0), sig(gga ? npts : 0);
  if (C)
    orc_density_C(nbf, nocc, npts, phi, gx, gy, gz, C, f, rho.data(),
                  gga ? grad.data() : nullptr, gga ? sig.data() : nullptr);
  else
    orc_density_P(nbf, npts, phi, gx, gy, gz, P, rho.data(),
                  gga ? grad.data() : nullptr, gga ? sig.data() : nullptr);
  std::vector<double> exc(npts), vr(npts), vs(npts);
  if (orc_xc_eval(nfunc, ids, dens_thr, npts, rho.data(), gga ? sig.data() : nullptr,
                  exc.data(), vr.data(), vs.data()))
    return 1;
  double e = 0, nel = 0;
  for (long g = 0; g < npts; ++g) {
    e += w[g] * rho[g] * exc[g];
    nel += w[g] * rho[g];
  }
  *E += e;
  *N += nel;
  // Z[u][g]
  std::vector<double> Z((size_t)nbf * npts);
#pragma omp parallel for schedule(static)
  for (long g = 0; g < npts; ++g) {
    for (int u = 0; u < nbf; ++u) {
      const size_t i = (size_t)u * npts + g;
      double z = 0.5 * phi[i] * vr[g];
      if (gga)
        z += 2.0 * vs[g] *
             (grad[3 * g] * gx[i] + grad[3 * g + 1] * gy[i] + grad[3 * g + 2] * gz[i]);
      Z[i] = w[g] * z;
    }
  }
#pragma omp parallel for schedule(dynamic)
  for (int u = 0; u < nbf; ++u)
    for (int v = 0; v <= u; ++v) {
      double s = 0;
      const double *a = phi + (size_t)u * npts, *b = Z.data() + (size_t)v * npts;
      const double *c = phi + (size_t)v * npts, *d = Z.data() + (size_t)u * npts;
      for (long g = 0; g < npts; ++g) s += a[g] * b[g] + c[g] * d[g];
      V[(size_t)u * nbf + v] += s;
      if (u != v) V[(size_t)v * nbf + u] += s;
    }
  return 0;
}

// Spin-polarised E_xc and V_alpha, V_beta (compute_lsda xc_integrals.hpp:
// 169-262, compute_gga_lsda :543-718) from spin density matrices:
// Z_s = w (1/2 Phi vrho_s + vec_s . grad Phi), vec_a = 2 vsigma_aa grad rho_a
// + vsigma_ab grad rho_b, V_s = Phi Z_s^T + Z_s Phi^T, E = sum w (rho_a +
// rho_b) exc.  Accumulates into E, N[2], Va, Vb.
int orc_vxc_polarized(int nbf, long npts, const double *phi, const double *gx,
                      const double *gy, const double *gz, const double *w,
                      const double *Pa, const double *Pb, int nfunc, const int *ids,
                      double dens_thr, double *E, double *N, double *Va, double *Vb) {
  const bool gga = gx != nullptr;
  for (int k = 0; k < nfunc; ++k)
    if (is_gga(ids[k]) && !gga) return 2;
  std::vector<double> rs[2], gs[2], tmp(gga ? npts : 0);
  const double *Ps[2] = {Pa, Pb};
  for (int s = 0; s < 2; ++s) {
    rs[s].resize(npts);
    gs[s].resize(gga ? 3 * npts : 0);
    orc_density_P(nbf, npts, phi, gx, gy, gz, Ps[s], rs[s].data(),
                  gga ? gs[s].data() : nullptr, gga ? tmp.data() : nullptr);
  }
  std::vector<double> rho2(2 * npts), sig3(gga ? 3 * npts : 0);
  for (long g = 0; g < npts; ++g) {
    rho2[2 * g] = rs[0][g];
    rho2[2 * g + 1] = rs[1][g];
    if (gga) {
      const double *a = &gs[0][3 * g], *b = &gs[1][3 * g];
      sig3[3 * g] = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
      sig3[3 * g + 1] = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
      sig3[3 * g + 2] = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
    }
  }
  std::vector<double> exc(npts), vr(2 * npts), vs(3 * npts);
  if (orc_xc_eval_polarized(nfunc, ids, dens_thr, npts, rho2.data(),
                            gga ? sig3.data() : nullptr, exc.data(), vr.data(),
                            gga ? vs.data() : nullptr))
    return 1;
  double e = 0, na = 0, nb = 0;
  for (long g = 0; g < npts; ++g) {
    e += w[g] * (rs[0][g] + rs[1][g]) * exc[g];
    na += w[g] * rs[0][g];
    nb += w[g] * rs[1][g];
  }
  *E += e;
  N[0] += na;
  N[1] += nb;
  std::vector<double> Z((size_t)nbf * npts);
  for (int s = 0; s < 2; ++s) {
    const int o = 1 - s;
#pragma omp parallel for schedule(static)
    for (long g = 0; g < npts; ++g) {
      double vec[3] = {0, 0, 0};
      if (gga)
        for (int c = 0; c < 3; ++c)
          vec[c] = 2.0 * vs[3 * g + 2 * s] * gs[s][3 * g + c] + vs[3 * g + 1] * gs[o][3 * g + c];
      for (int u = 0; u < nbf; ++u) {
        const size_t i = (size_t)u * npts + g;
        double z = 0.5 * phi[i] * vr[2 * g + s];
        if (gga) z += vec[0] * gx[i] + vec[1] * gy[i] + vec[2] * gz[i];
        Z[i] = w[g] * z;
      }
    }
    double *V = s ? Vb : Va;
#pragma omp parallel for schedule(dynamic)
    for (int u = 0; u < nbf; ++u)
      for (int v = 0; v <= u; ++v) {
        double acc = 0;
        const double *a = phi + (size_t)u * npts, *b = Z.data() + (size_t)v * npts;
        const double *c = phi + (size_t)v * npts, *d = Z.data() + (size_t)u * npts;
        for (long g = 0; g < npts; ++g) acc += a[g] * b[g] + c[g] * d[g];
        V[(size_t)u * nbf + v] += acc;
        if (u != v) V[(size_t)v * nbf + u] += acc;
      }
  }
  return 0;
}

// Numerical overlap S = Phi diag(w) Phi^T (core_hamiltonian.hpp:62-119),
// accumulated into S (row-major nbf x nbf).
void orc_overlap(int nbf, long npts, const double *phi, const double *w, double *S) {
#pragma omp parallel for schedule(dynamic)
  for (int u = 0; u < nbf; ++u)
    for (int v = 0; v <= u; ++v) {
      double s = 0;
      const double *a = phi + (size_t)u * npts, *b = phi + (size_t)v * npts;
      for (long g = 0; g < npts; ++g) s += a[g] * w[g] * b[g];
      S[(size_t)u * nbf + v] += s;
      if (u != v) S[(size_t)v * nbf + u] += s;
    }
}

}  // extern "C"
