// Forward-mode dual number carrying five first derivatives (rho_a, rho_b,
// sigma_aa, sigma_ab, sigma_bb): an energy density written once over D5 gives
// its potentials as exact derivatives.
#pragma once

#include <hip/hip_runtime.h>

namespace nkx {

struct D5 {
  double v, d[5];
};

__host__ __device__ inline D5 dconst(double c) { return D5{c, {0, 0, 0, 0, 0}}; }
__host__ __device__ inline D5 dvar(double c, int k) {
  D5 r = dconst(c);
  r.d[k] = 1.0;
  return r;
}
__host__ __device__ inline D5 operator+(D5 a, D5 b) {
  for (int k = 0; k < 5; ++k) a.d[k] += b.d[k];
  a.v += b.v;
  return a;
}
__host__ __device__ inline D5 operator-(D5 a, D5 b) {
  for (int k = 0; k < 5; ++k) a.d[k] -= b.d[k];
  a.v -= b.v;
  return a;
}
__host__ __device__ inline D5 operator*(D5 a, D5 b) {
  D5 r;
  r.v = a.v * b.v;
  for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
__host__ __device__ inline D5 operator/(D5 a, D5 b) {
  D5 r;
  const double ib = 1.0 / b.v;
  r.v = a.v * ib;
  for (int k = 0; k < 5; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
  return r;
}
__host__ __device__ inline D5 operator+(D5 a, double c) { a.v += c; return a; }
__host__ __device__ inline D5 operator+(double c, D5 a) { a.v += c; return a; }
__host__ __device__ inline D5 operator-(D5 a, double c) { a.v -= c; return a; }
__host__ __device__ inline D5 operator-(double c, D5 a) {
  for (int k = 0; k < 5; ++k) a.d[k] = -a.d[k];
  a.v = c - a.v;
  return a;
}
__host__ __device__ inline D5 operator*(D5 a, double c) {
  for (int k = 0; k < 5; ++k) a.d[k] *= c;
  a.v *= c;
  return a;
}
__host__ __device__ inline D5 operator*(double c, D5 a) { return a * c; }
__host__ __device__ inline D5 operator/(D5 a, double c) { return a * (1.0 / c); }
__host__ __device__ inline D5 operator/(double c, D5 b) {
  D5 r;
  const double ib = 1.0 / b.v;
  r.v = c * ib;
  for (int k = 0; k < 5; ++k) r.d[k] = -r.v * b.d[k] * ib;
  return r;
}
// unary chain rule: f(a) with f'(a) given
__host__ __device__ inline D5 dchain(D5 a, double f, double fp) {
  D5 r;
  r.v = f;
  for (int k = 0; k < 5; ++k) r.d[k] = fp * a.d[k];
  return r;
}
__host__ __device__ inline D5 dlog(D5 a) { return dchain(a, log(a.v), 1.0 / a.v); }
__host__ __device__ inline D5 dexp(D5 a) {
  const double e = exp(a.v);
  return dchain(a, e, e);
}
__host__ __device__ inline D5 dsqrt(D5 a) {
  const double s = sqrt(a.v);
  return dchain(a, s, 0.5 / s);
}
__host__ __device__ inline D5 dpow(D5 a, double p) {  // a > 0
  const double v = pow(a.v, p);
  return dchain(a, v, p * v / a.v);
}
__host__ __device__ inline D5 dmax(D5 a, double floor_) { return a.v >= floor_ ? a : dconst(floor_); }
__host__ __device__ inline D5 datan(D5 a) { return dchain(a, atan(a.v), 1.0 / (1.0 + a.v * a.v)); }
__host__ __device__ inline D5 dasinh(D5 a) {
  return dchain(a, asinh(a.v), 1.0 / sqrt(1.0 + a.v * a.v));
}

}  // namespace nkx
