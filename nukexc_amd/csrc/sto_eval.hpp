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

}  // namespace nkx
