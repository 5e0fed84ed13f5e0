// Device evaluation of one STO basis function and its gradient at a point.
//   phi(r) = N r^k e^{-zeta r} S_lm(x, y, z),  k = n - l - 1
//   grad phi = R grad S + S (dR/dr) rvec / r,  dR/dr = (k / r - zeta) R
// (stobasis.hpp:494-584 and :599-638).  S_lm comes from its monomial
// expansion (HarmTerm list).
#pragma once

#include "nkx_internal.hpp"

namespace nkx {

template <typename T>
__device__ __forceinline__ T ipow(T x, int e) {
  T r = T(1);
  for (int i = 0; i < e; ++i) r *= x;
  return r;
}

template <typename T>
__device__ __forceinline__ T dev_exp(T x);
template <>
__device__ __forceinline__ double dev_exp<double>(double x) { return exp(x); }
template <>
__device__ __forceinline__ float dev_exp<float>(float x) { return expf(x); }

template <typename T>
__device__ __forceinline__ T dev_sqrt(T x);
template <>
__device__ __forceinline__ double dev_sqrt<double>(double x) { return sqrt(x); }
template <>
__device__ __forceinline__ float dev_sqrt<float>(float x) { return sqrtf(x); }

// Value only.
template <typename T>
__device__ __forceinline__ T sto_value(const FuncRec &f, const HarmTerm *terms, T x,
                                       T y, T z) {
  const T dx = x - T(f.cx), dy = y - T(f.cy), dz = z - T(f.cz);
  const T r = dev_sqrt<T>(dx * dx + dy * dy + dz * dz);
  const T R = T(f.norm) * ipow(r, f.k) * dev_exp<T>(-T(f.zeta) * r);
  T S = T(0);
  for (int t = 0; t < f.term_cnt; ++t) {
    const HarmTerm h = terms[f.term_off + t];
    S += T(h.coef) * ipow(dx, h.a) * ipow(dy, h.b) * ipow(dz, h.c);
  }
  return R * S;
}

// Value and gradient.
template <typename T>
__device__ __forceinline__ void sto_value_grad(const FuncRec &f, const HarmTerm *terms,
                                               T x, T y, T z, T &v, T &gx, T &gy,
                                               T &gz) {
  const T dx = x - T(f.cx), dy = y - T(f.cy), dz = z - T(f.cz);
  const T r = dev_sqrt<T>(dx * dx + dy * dy + dz * dz);
  const T R = T(f.norm) * ipow(r, f.k) * dev_exp<T>(-T(f.zeta) * r);
  T S = T(0), Sx = T(0), Sy = T(0), Sz = T(0);
  for (int t = 0; t < f.term_cnt; ++t) {
    const HarmTerm h = terms[f.term_off + t];
    const T c = T(h.coef);
    const T xa1 = ipow(dx, h.a > 0 ? h.a - 1 : 0);
    const T yb1 = ipow(dy, h.b > 0 ? h.b - 1 : 0);
    const T zc1 = ipow(dz, h.c > 0 ? h.c - 1 : 0);
    const T xa = h.a > 0 ? xa1 * dx : T(1);
    const T yb = h.b > 0 ? yb1 * dy : T(1);
    const T zc = h.c > 0 ? zc1 * dz : T(1);
    S += c * xa * yb * zc;
    if (h.a > 0) Sx += c * T(h.a) * xa1 * yb * zc;
    if (h.b > 0) Sy += c * T(h.b) * xa * yb1 * zc;
    if (h.c > 0) Sz += c * T(h.c) * xa * yb * zc1;
  }
  const T common = (T(f.k) / r - T(f.zeta)) * R / r;
  v = R * S;
  gx = R * Sx + S * (dx * common);
  gy = R * Sy + S * (dy * common);
  gz = R * Sz + S * (dz * common);
}

// Value, gradient and Hessian (FP64), for the nuclear gradient of GGA
// functionals.  With R(r) = N r^k e^{-zeta r}, R' = (k/r - zeta) R,
// R'' = ((k/r - zeta)^2 - k/r^2) R and u = rvec / r:
//   d_i d_j phi = R S_ij + (R'/r)(x_i S_j + x_j S_i)
//                 + S [(R'' - R'/r) x_i x_j / r^2 + (R'/r) delta_ij].
// h = (xx, xy, xz, yy, yz, zz).
__device__ __forceinline__ void sto_value_grad_hess(const FuncRec &f, const HarmTerm *terms,
                                                    double x, double y, double z, double &v,
                                                    double g[3], double h[6]) {
  const double d[3] = {x - f.cx, y - f.cy, z - f.cz};
  const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const double R = f.norm * ipow(r, f.k) * exp(-f.zeta * r);
  double S = 0.0, Sg[3] = {0.0, 0.0, 0.0}, Sh[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int t = 0; t < f.term_cnt; ++t) {
    const HarmTerm ht = terms[f.term_off + t];
    const int e[3] = {ht.a, ht.b, ht.c};
    double p0[3], p1[3], p2[3];  // d^0, d^1, d^2 of d_i^{e_i}
    for (int i = 0; i < 3; ++i) {
      p0[i] = ipow(d[i], e[i]);
      p1[i] = e[i] > 0 ? e[i] * ipow(d[i], e[i] - 1) : 0.0;
      p2[i] = e[i] > 1 ? e[i] * (e[i] - 1) * ipow(d[i], e[i] - 2) : 0.0;
    }
    const double c = ht.coef;
    S += c * p0[0] * p0[1] * p0[2];
    Sg[0] += c * p1[0] * p0[1] * p0[2];
    Sg[1] += c * p0[0] * p1[1] * p0[2];
    Sg[2] += c * p0[0] * p0[1] * p1[2];
    Sh[0] += c * p2[0] * p0[1] * p0[2];
    Sh[1] += c * p1[0] * p1[1] * p0[2];
    Sh[2] += c * p1[0] * p0[1] * p1[2];
    Sh[3] += c * p0[0] * p2[1] * p0[2];
    Sh[4] += c * p0[0] * p1[1] * p1[2];
    Sh[5] += c * p0[0] * p0[1] * p2[2];
  }
  // one reciprocal instead of four divisions by r and r^2
  const double ir = 1.0 / r, ir2 = ir * ir;
  const double q = f.k * ir - f.zeta;
  const double R1r = q * R * ir;                              // R' / r
  const double A = ((q * q - f.k * ir2) * R - R1r) * ir2;     // (R'' - R'/r) / r^2
  v = R * S;
  for (int i = 0; i < 3; ++i) g[i] = R * Sg[i] + S * d[i] * R1r;
  const int ii[6] = {0, 0, 0, 1, 1, 2}, jj[6] = {0, 1, 2, 1, 2, 2};
  for (int m = 0; m < 6; ++m) {
    const int i = ii[m], j = jj[m];
    h[m] = R * Sh[m] + R1r * (d[i] * Sg[j] + d[j] * Sg[i]) + S * (A * d[i] * d[j] + (i == j ? R1r : 0.0));
  }
}

// Mixed-form collocation: values evaluated in FP64 and handed to the caller
// in FP64 (they are split into three bf16 terms there, ~2^-27 relative);
// gradients are rounded once to FP32.  Evaluating in FP32 instead (zeta r,
// normalisation or harmonic coefficients rounded to FP32, or an FP32
// exponential) shifts every value of a function by a common relative amount
// of 1e-8 .. 1e-7; such a systematic error survives the sum over millions of
// grid points and alone exceeds the mixed-form budget (|dE_xc| <= 1e-6 Eh) on
// large systems.
// The exponential uses Cody-Waite argument reduction, zeta r = k ln2 + f,
// |f| <= ln2/2, with ln2 = hi + lo, hi = ln2 truncated to its leading 40
// significand bits (k hi is exact for |k| < 2^13) and lo = ln2 - hi rounded
// to FP64, then a degree-11 Taylor polynomial for e^{-f} (truncation < 3e-13).
__host__ __device__ inline double exp_neg_mixed(double y) {
  const double k = rint(y * 1.4426950408889634);  // y / ln2
  double f = fma(-k, 0x1.62e42fefa2000p-1, y);  // hi
  f = fma(-k, 0x1.9ef35793c7673p-41, f);         // lo
  const double x = -f;
  double p = 1.0 / 39916800.0;
  p = fma(p, x, 1.0 / 3628800.0);
  p = fma(p, x, 1.0 / 362880.0);
  p = fma(p, x, 1.0 / 40320.0);
  p = fma(p, x, 1.0 / 5040.0);
  p = fma(p, x, 1.0 / 720.0);
  p = fma(p, x, 1.0 / 120.0);
  p = fma(p, x, 1.0 / 24.0);
  p = fma(p, x, 1.0 / 6.0);
  p = fma(p, x, 0.5);
  p = fma(p, x, 1.0);
  p = fma(p, x, 1.0);
  return ldexp(p, -(int)k);
}

__device__ __forceinline__ double sto_value_mixed(const FuncRec &f, const HarmTerm *terms,
                                                  double x, double y, double z) {
  const double dx = x - f.cx, dy = y - f.cy, dz = z - f.cz;
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  const double R = f.norm * ipow(r, f.k) * exp_neg_mixed(f.zeta * r);
  double S = 0.0;
  for (int t = 0; t < f.term_cnt; ++t) {
    const HarmTerm h = terms[f.term_off + t];
    S += h.coef * ipow(dx, h.a) * ipow(dy, h.b) * ipow(dz, h.c);
  }
  return R * S;
}

__device__ __forceinline__ void sto_value_grad_mixed(const FuncRec &f, const HarmTerm *terms,
                                                     double x, double y, double z, double &v,
                                                     float &gx, float &gy, float &gz) {
  const double dx = x - f.cx, dy = y - f.cy, dz = z - f.cz;
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  const double R = f.norm * ipow(r, f.k) * exp_neg_mixed(f.zeta * r);
  double S = 0.0, Sx = 0.0, Sy = 0.0, Sz = 0.0;
  for (int t = 0; t < f.term_cnt; ++t) {
    const HarmTerm h = terms[f.term_off + t];
    const double xa1 = ipow(dx, h.a > 0 ? h.a - 1 : 0);
    const double yb1 = ipow(dy, h.b > 0 ? h.b - 1 : 0);
    const double zc1 = ipow(dz, h.c > 0 ? h.c - 1 : 0);
    const double xa = h.a > 0 ? xa1 * dx : 1.0;
    const double yb = h.b > 0 ? yb1 * dy : 1.0;
    const double zc = h.c > 0 ? zc1 * dz : 1.0;
    S += h.coef * xa * yb * zc;
    if (h.a > 0) Sx += h.coef * h.a * xa1 * yb * zc;
    if (h.b > 0) Sy += h.coef * h.b * xa * yb1 * zc;
    if (h.c > 0) Sz += h.coef * h.c * xa * yb * zc1;
  }
  const double common = ((double)f.k / r - f.zeta) * R / r;
  v = R * S;
  gx = (float)(R * Sx + S * (dx * common));
  gy = (float)(R * Sy + S * (dy * common));
  gz = (float)(R * Sz + S * (dz * common));
}

}  // namespace nkx
