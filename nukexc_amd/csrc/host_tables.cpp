// Host-side tables of the XC path: STO normalisation and cutoff radii, real
// solid harmonics as monomial expansions, Lebedev rules, radial rules.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

#include "nkx_internal.hpp"

namespace nkx {

namespace {
#include "lebedev_tables.inc"

constexpr double kPi = 3.14159265358979323846;

double factorial(int n) {
  double r = 1.0;
  for (int i = 2; i <= n; ++i) r *= i;
  return r;
}
}  // namespace

// make_manual_basis: N = (2 zeta)^{n+1/2} / sqrt((2n)!)  (stobasis.hpp:200-201).
// compute_cutoff (stobasis.hpp:90-173): the radius beyond the radial maximum
// where N r^{n-1} e^{-zeta r} falls to tol, found by Newton iteration on
// g(r) = ln N + (n-1) ln r - zeta r - ln tol, kept inside a shrinking bracket
// and bisecting whenever a Newton step would leave it.
void sto_norm_cutoff(int n, double zeta, double tol, double *norm, double *cut) {
  NKX_REQUIRE(n >= 1 && zeta > 0 && tol > 0, NKX_ERR_INVALID_ARGUMENT,
              "sto_norm_cutoff: need n >= 1, zeta > 0, tol > 0");
  const double N = std::pow(2.0 * zeta, n + 0.5) / std::sqrt(factorial(2 * n));
  *norm = N;
  if (N <= tol) {
    *cut = 0.0;
    return;
  }
  const double lnN = std::log(N), lnt = std::log(tol);
  const double peak = n > 1 ? (n - 1) / zeta : 0.0;
  auto g = [&](double r) { return lnN + (n - 1) * std::log(r) - zeta * r - lnt; };
  double lo = peak + 1e-10;
  double hi = std::max(peak + (lnN - lnt) / zeta, peak + 1e-6);
  for (int it = 0; it < 64 && g(hi) > 0.0; ++it) hi *= 2.0;
  double r = hi;
  for (int it = 0; it < 50; ++it) {
    const double gv = g(r);
    (gv > 0.0 ? lo : hi) = r;
    const double dg = (n - 1) / r - zeta;
    double next = 0.5 * (lo + hi);
    if (dg < 0.0) {
      const double newton = r - gv / dg;
      if (newton > lo && newton < hi) next = newton;
    }
    const double step = std::fabs(next - r);
    r = next;
    if (step < 1e-10 * r) break;
  }
  *cut = r;
}

// Real solid harmonics S_lm = r^l Y_lm in the reference's convention
// (python/precompute_solid_harmonics.py:63-102: no Condon-Shortley phase,
// m > 0 ~ Re (x+iy)^m, m < 0 ~ Im (x+iy)^m).  Built here from the standard
// recurrences for Racah-normalised solid harmonics (S_00 = 1):
//   S_{l+1,+-(l+1)} from x, y times S_{l,+-l}
//   S_{l+1,m} = ((2l+1) z S_lm - sqrt((l+m)(l-m)) r^2 S_{l-1,m})
//               / sqrt((l+m+1)(l-m+1)),
// then scaled by sqrt((2l+1)/(4 pi)).
namespace {
constexpr int L = 7;
struct Poly {
  double c[L + 1][L + 1][L + 1];
  Poly() { std::memset(c, 0, sizeof(c)); }
};
Poly mul_xyz(const Poly &p, int axis) {
  Poly r;
  for (int a = 0; a <= L; ++a)
    for (int b = 0; b <= L; ++b)
      for (int d = 0; d <= L; ++d) {
        if (p.c[a][b][d] == 0.0) continue;
        const int na = a + (axis == 0), nb = b + (axis == 1), nd = d + (axis == 2);
        if (na > L || nb > L || nd > L) continue;
        r.c[na][nb][nd] += p.c[a][b][d];
      }
  return r;
}
Poly axpy(double s, const Poly &x, double t, const Poly &y) {
  Poly r;
  for (int a = 0; a <= L; ++a)
    for (int b = 0; b <= L; ++b)
      for (int d = 0; d <= L; ++d) r.c[a][b][d] = s * x.c[a][b][d] + t * y.c[a][b][d];
  return r;
}
Poly mul_r2(const Poly &p) {
  Poly x2 = mul_xyz(mul_xyz(p, 0), 0), y2 = mul_xyz(mul_xyz(p, 1), 1),
       z2 = mul_xyz(mul_xyz(p, 2), 2);
  return axpy(1.0, axpy(1.0, x2, 1.0, y2), 1.0, z2);
}

std::once_flag g_harm_once;
std::vector<std::vector<HarmTerm>> g_harm;  // index l*l + l + m

void build_harmonics() {
  std::vector<std::map<int, Poly>> S(L + 1);
  Poly one;
  one.c[0][0][0] = 1.0;
  S[0][0] = one;
  for (int l = 0; l < L; ++l) {
    const double f = std::sqrt((l == 0 ? 2.0 : 1.0) * (2.0 * l + 1.0) / (2.0 * l + 2.0));
    const Poly &Sp = S[l][l];
    const Poly &Sm = S[l][-l];
    if (l == 0) {
      S[1][1] = axpy(f, mul_xyz(Sp, 0), 0.0, Sp);
      S[1][-1] = axpy(f, mul_xyz(Sp, 1), 0.0, Sp);
    } else {
      S[l + 1][l + 1] = axpy(f, mul_xyz(Sp, 0), -f, mul_xyz(Sm, 1));
      S[l + 1][-(l + 1)] = axpy(f, mul_xyz(Sp, 1), f, mul_xyz(Sm, 0));
    }
    for (int m = -l; m <= l; ++m) {
      Poly zt = mul_xyz(S[l][m], 2);
      Poly low;
      if (l >= 1 && std::abs(m) <= l - 1) low = mul_r2(S[l - 1][m]);
      const double den = std::sqrt((double)(l + m + 1) * (l - m + 1));
      const double cl = std::sqrt((double)(l + m) * (l - m));
      S[l + 1][m] = axpy((2.0 * l + 1.0) / den, zt, -cl / den, low);
    }
  }
  g_harm.resize((L + 1) * (L + 1));
  for (int l = 0; l <= L; ++l)
    for (int m = -l; m <= l; ++m) {
      const double s = std::sqrt((2.0 * l + 1.0) / (4.0 * kPi));
      std::vector<HarmTerm> &t = g_harm[l * l + l + m];
      const Poly &P = S[l][m];
      for (int a = 0; a <= l; ++a)
        for (int b = 0; a + b <= l; ++b) {
          const int d = l - a - b;
          const double c = P.c[a][b][d];
          if (std::fabs(c) < 1e-15) continue;
          t.push_back(HarmTerm{c * s, a, b, d, 0});
        }
    }
}
}  // namespace

const std::vector<HarmTerm> &solid_harmonic_terms(int l, int m) {
  NKX_REQUIRE(l >= 0 && l <= L && m >= -l && m <= l, NKX_ERR_INVALID_ARGUMENT,
              "solid harmonic needs 0 <= l <= 7 and |m| <= l");
  std::call_once(g_harm_once, build_harmonics);
  return g_harm[l * l + l + m];
}

// ---- Lebedev --------------------------------------------------------------
int lebedev_lookup(int order, int *npts, int *actual) {
  for (int i = 0; i < kLebNumRules; ++i)
    if (kLebOrder[i] >= order) {
      if (npts) *npts = kLebNpts[i];
      if (actual) *actual = kLebOrder[i];
      return i;
    }
  throw Error{NKX_ERR_INVALID_ARGUMENT, "Lebedev order above 131 is not tabulated"};
}

// Expands the stored octahedral orbits: every signed permutation of
// (a, b, c), duplicates removed.  Rows: x, y, z, w.
void lebedev_expand(int order, double *xyzw) {
  const int rule = lebedev_lookup(order, nullptr, nullptr);
  static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2},
                                  {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  int k = 0;
  for (int o = 0; o < kLebOrbitCount[rule]; ++o) {
    const double *rep = kLebOrbits[kLebOrbitOffset[rule] + o];
    double seen[48][3];
    int ns = 0;
    for (int p = 0; p < 6; ++p)
      for (int s = 0; s < 8; ++s) {
        double v[3];
        for (int d = 0; d < 3; ++d) {
          const double x = rep[perms[p][d]];
          v[d] = ((s >> d) & 1) ? -x : x;
          if (v[d] == 0.0) v[d] = 0.0;  // fold -0.0
        }
        bool dup = false;
        for (int q = 0; q < ns && !dup; ++q)
          dup = seen[q][0] == v[0] && seen[q][1] == v[1] && seen[q][2] == v[2];
        if (dup) continue;
        std::memcpy(seen[ns++], v, sizeof(v));
        xyzw[4 * k + 0] = v[0];
        xyzw[4 * k + 1] = v[1];
        xyzw[4 * k + 2] = v[2];
        xyzw[4 * k + 3] = rep[3];
        ++k;
      }
  }
  NKX_REQUIRE(k == kLebNpts[rule], NKX_ERR_HIP, "Lebedev expansion size mismatch");
}

// ---- radial rules -----------------------------------------------------------
// Chebyshev second-kind nodes x_i = cos(i pi/(n+1)) mapped to [0, inf); the
// weight carries the node weight for int dx, the Jacobian and r^2.
void radial_rule(int scheme, int n, double R, double alpha, double *r, double *w) {
  NKX_REQUIRE(n > 0, NKX_ERR_INVALID_ARGUMENT, "radial rule needs n > 0");
  NKX_REQUIRE(scheme == NKX_RADIAL_TREUTLER_AHLRICHS || scheme == NKX_RADIAL_BECKE,
              NKX_ERR_INVALID_ARGUMENT, "unknown radial scheme");
  const double ln2 = std::log(2.0), inv_ln2 = 1.0 / ln2;
  for (int i = 0; i < n; ++i) {
    const double t = (i + 1) * kPi / (n + 1);
    const double x = std::cos(t);
    const double st = std::sin(t);
    const double wx = kPi / (n + 1) * st;  // sin^2 t / sqrt(1 - x^2) = sin t
    double rr, jac;
    if (scheme == NKX_RADIAL_TREUTLER_AHLRICHS) {
      // r = R/ln2 (1+x)^alpha ln(2/(1-x))   (Treutler & Ahlrichs M4), rounded
      // as ((R / ln2) (1+x)^alpha) ln(2/(1-x)) -- the nodes fix which points
      // the partition zeroes, so their rounding is kept to one stated order
      const double lg = std::log(2.0 / (1.0 - x));
      const double pa = std::pow(1.0 + x, alpha);
      rr = R / ln2 * pa * lg;
      jac = R * inv_ln2 * pa * (alpha * lg / (1.0 + x) + 1.0 / (1.0 - x));
    } else {
      // r = R (1+x)/(1-x)   (Becke)
      rr = R * (1.0 + x) / (1.0 - x);
      jac = 2.0 * R / ((1.0 - x) * (1.0 - x));
    }
    r[i] = rr;
    w[i] = wx * jac * rr * rr;
  }
}

// grid.hpp:71-102: TA xi per element; Becke uses half the Bragg-Slater radius.
double radial_scale(int scheme, int Z) {
  static const double ta_xi[] = {0.00, 0.8, 0.9, 1.8, 1.4, 1.3, 1.1, 0.9, 0.9, 0.9,
                                 0.9,  1.4, 1.3, 1.3, 1.2, 1.1, 1.0, 1.0, 1.0, 1.5,
                                 1.4,  1.3, 1.2, 1.2, 1.2, 1.2, 1.2, 1.2, 1.1, 1.1,
                                 1.1,  1.1, 1.0, 0.9, 0.9, 0.9, 0.9};
  static const double bragg[] = {
      0.00, 0.35, 0.35, 1.45, 1.05, 0.85, 0.70, 0.65, 0.60, 0.50, 0.35, 1.80, 1.50,
      1.25, 1.10, 1.00, 1.00, 1.00, 0.70, 2.20, 1.80, 1.60, 1.45, 1.40, 1.35, 1.40,
      1.35, 1.35, 1.35, 1.35, 1.35, 1.30, 1.25, 1.15, 1.10, 1.00, 0.90, 2.35, 2.00,
      1.80, 1.55, 1.45, 1.45, 1.35, 1.35, 1.40, 1.40, 1.40, 1.45, 1.45, 1.45, 1.45,
      1.40, 1.35, 1.20, 2.60, 2.15, 1.95, 1.85, 1.85, 1.85, 1.85, 1.85, 1.80, 1.75,
      1.75, 1.75, 1.75, 1.75, 1.75, 1.75, 1.75, 1.70, 1.60, 1.45, 1.35, 1.35, 1.35,
      1.35, 1.35, 1.40, 1.45, 1.50, 1.50, 1.45, 1.40, 1.30, 2.80, 2.35, 2.15, 1.95,
      1.80, 1.80, 1.75, 1.75, 1.75, 1.75, 1.75, 1.75, 1.75, 1.75, 1.75, 1.75, 1.75,
      1.75, 1.75, 1.75, 1.75, 1.75, 1.75, 1.75, 1.75, 1.75, 1.75, 1.75, 1.75, 1.75,
      1.75, 1.75};
  const double ang_to_bohr = 1.8897261254578281;
  if (scheme == NKX_RADIAL_TREUTLER_AHLRICHS)
    return (Z >= 0 && Z < (int)(sizeof(ta_xi) / sizeof(double))) ? ta_xi[Z] : 1.0;
  return (Z >= 0 && Z < (int)(sizeof(bragg) / sizeof(double)))
             ? 0.5 * bragg[Z] * ang_to_bohr
             : 1.0;
}

int periodic_row(int Z) {
  if (Z <= 2) return 1;
  if (Z <= 10) return 2;
  if (Z <= 18) return 3;
  if (Z <= 36) return 4;
  if (Z <= 54) return 5;
  if (Z <= 86) return 6;
  return 7;
}

}  // namespace nkx
