// int8 slice form shared by the mixed-precision kernels (V_xc build and
// nuclear gradient): tile layout, slicing, the ten-product tile contraction
// on v_mfma_i32_16x16x64_i8 and the direct global -> LDS section copy.
//
// integer MFMA accumulation.
//
// Every operand tile of the two contractions (P blocks, collocation values,
// Z) is stored as four int8 slices with one power-of-two scale per tile row
// (an Ozaki-type splitting; Ozaki, Ogita & Oishi, Numer. Algorithms 2012):
//   x = 2^(e-6) (a0 + 2^-7 a1 + 2^-14 a2 + 2^-21 a3),   |a_s| <= 64,
// e = the exponent of the row's largest |x| (|x| < 2^e), each slice rounded
// to nearest from the exact FP64 remainder, so |x - sum| <= 2^(e-28) with
// random sign.  A tile product is
//   sum_k x_ik y_kj = 2^(eA_i + eB_j - 12) sum_d 2^(-7d) sum_(s+t=d) sum_k a_s,ik b_t,kj
// and every inner sum runs on v_mfma_i32_16x16x64_i8 in exact int32
// arithmetic (|sum| <= 4 * 64^3 < 2^31).  The ten slice products with
// s + t <= 3 are kept; the dropped ones are below 2^-28 of the row-times-
// column scale, of random sign.  The int32 sums are combined and scaled in
// FP64 per exponent group of KG 64-function blocks (density: the P row and Pm
// point exponents are shared over the group, the sums run across its steps)
// and per 64-point batch (V).  The functional, rho, grad rho, E_xc, N_el and the V
// accumulation are FP64; collocation gradients are FP32 (rounded once from
// FP64).  Exact accumulation is the point of the integer form:
// v_mfma_f32_*_bf16 rounds its FP32 sums relative to the largest term with a
// mean pulled toward zero, a bias that shows up in N_el on large grids.
//
// Tile layout: a 64 x 64 tile T[row][k] is four row fragments of 16 rows;
// lane l of fragment rt holds T[16 rt + l % 16][16 (l / 16) + e], e = 0..15,
// as 16 consecutive bytes -- the A-operand order of v_mfma_i32_16x16x64_i8,
// and, read as B[k][j] with j the row of T, its B-operand order.  A section
// is the four slice planes (4 KB each) followed by the 64 row exponents.
//
// Per live slot (bytes):   F   Phi, rows = functions, k = points   16.25 KB
//                          Pm  Phi, rows = points, k = functions   16.25 KB
//                          Z   Z,   rows = functions, k = points   16.25 KB
//                          G   dPhi/dx, dy, dz, FP32 [f][p]         48 KB (GGA)
// P is gathered into one section per 64 x 64 block pair (rows mu, k = nu).
#pragma once
#include "nkx_internal.hpp"

namespace nkx {

typedef int v4i __attribute__((ext_vector_type(4)));

// v_mfma_i32_16x16x64_i8 lane layout (checked by tests/gpu_probe): lane l
// supplies A[l%16][16(l/16) .. +15] and B[16(l/16) .. +15][l%16]; result
// register r of lane l holds D[4(l/16) + r][l%16].

namespace {
constexpr int FB = 64;
constexpr int PT = 64;
constexpr int TILE = FB * PT;
constexpr int PLANE = TILE;                  // bytes per slice plane
constexpr int SEC = 4 * PLANE + 256;         // section: 4 planes + exponents (padded)
constexpr int OFF_F = 0, OFF_PM = SEC, OFF_Z = 2 * SEC, OFF_G = 3 * SEC;
constexpr int MAXNFB = 1024;
// function blocks per exponent group: the P sections of blocks (fm, fk) with
// the same fk / KG share row exponents, and a batch's Pm sections of blocks
// with the same local index / KG share point exponents
constexpr int KG = 4;

__host__ __device__ inline int64_t slot_bytes_i8(int gga) {
  return 3 * (int64_t)SEC + (gga ? 3 * (int64_t)TILE * 4 : 0);
}

// byte offset of T[row][k] inside one plane
__device__ __forceinline__ int frag8(int row, int k) {
  return (((row >> 4) * 64) + (row & 15) + 16 * (k >> 4)) * 16 + (k & 15);
}

// exponent e with max < 2^e; rows whose largest |x| is below 2^-120 are
// stored as zero
__device__ __forceinline__ int row_exp(double m) {
  if (!(m >= 0x1p-120)) return -127;
  int e;
  (void)frexp(m, &e);
  return e;
}

__device__ __forceinline__ void slice4(double x, int e, signed char s[4]) {
  if (e == -127) {
    s[0] = s[1] = s[2] = s[3] = 0;
    return;
  }
  double y = ldexp(x, 6 - e);  // |y| < 64
  double a = rint(y);
  s[0] = (signed char)a;
  double r = y - a;  // |r| <= 1/2, exact
#pragma unroll
  for (int i = 1; i < 4; ++i) {
    y = r * 128.0;
    a = rint(y);
    s[i] = (signed char)a;
    r = y - a;
  }
}

__device__ __forceinline__ double unslice(const signed char *plane0, int off, int e) {
  if (e == -127) return 0.0;
  const double v = (double)plane0[off] + ldexp((double)plane0[PLANE + off], -7) +
                   ldexp((double)plane0[2 * PLANE + off], -14) +
                   ldexp((double)plane0[3 * PLANE + off], -21);
  return ldexp(v, e - 6);
}

__device__ __forceinline__ v4i ld16(const signed char *p) {
  return *reinterpret_cast<const v4i *>(p);
}

// The ten slice products of one 16 x 16 x 64 tile product, combined into the
// FP64 value scaled by 2^(e - 19), e = eA + eB: hi = 128 d0 + d1,
// lo = 128 d2 + d3, value = hi + 2^-14 lo.
__device__ __forceinline__ void tile_product(const v4i *A, const v4i *B, int hi[4], int lo[4]) {
  const v4i z = {0, 0, 0, 0};
  v4i d0 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[0], B[0], z, 0, 0, 0);
  v4i d1 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[0], B[1], z, 0, 0, 0);
  d1 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[1], B[0], d1, 0, 0, 0);
  v4i d2 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[0], B[2], z, 0, 0, 0);
  d2 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[1], B[1], d2, 0, 0, 0);
  d2 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[2], B[0], d2, 0, 0, 0);
  v4i d3 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[0], B[3], z, 0, 0, 0);
  d3 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[1], B[2], d3, 0, 0, 0);
  d3 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[2], B[1], d3, 0, 0, 0);
  d3 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[3], B[0], d3, 0, 0, 0);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    hi[r] = d0[r] * 128 + d1[r];
    lo[r] = d2[r] * 128 + d3[r];
  }
}

// The same ten slice products added to running sums hi, lo of earlier tile
// products with the same row and column exponents: the d1 and d3 chains start
// from hi and lo, so hi = 128 d0 + d1 and lo = 128 d2 + d3 stay exact int32
// sums.  Over n products |hi| <= n (128 + 2) 2^18 and |lo| <= n (3 * 128 + 4)
// 2^18, both below 2^31 for n <= 16.
__device__ __forceinline__ void tile_product_acc(const v4i *A, const v4i *B, v4i &hi, v4i &lo) {
  const v4i z = {0, 0, 0, 0};
  const v4i d0 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[0], B[0], z, 0, 0, 0);
  v4i d1 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[0], B[1], hi, 0, 0, 0);
  d1 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[1], B[0], d1, 0, 0, 0);
  v4i d2 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[0], B[2], z, 0, 0, 0);
  d2 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[1], B[1], d2, 0, 0, 0);
  d2 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[2], B[0], d2, 0, 0, 0);
  v4i d3 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[0], B[3], lo, 0, 0, 0);
  d3 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[1], B[2], d3, 0, 0, 0);
  d3 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[2], B[1], d3, 0, 0, 0);
  d3 = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[3], B[0], d3, 0, 0, 0);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    hi[r] = d0[r] * 128 + d1[r];
    lo[r] = d2[r] * 128 + d3[r];
  }
}

// Z = a phi + bx dphi/dx + by dphi/dy + bz dphi/dz, in one fixed order of
// roundings (every kernel that forms Z uses this)
__device__ __forceinline__ double z_value(double ph, double gx, double gy, double gz, double a,
                                          double bx, double by, double bz) {
  return fma(bz, gz, fma(by, gy, fma(bx, gx, a * ph)));
}

// Slice a 64 x 64 FP64 tile held in LDS (v[row][k], row stride 65; TR: the
// tile read transposed, v[k][row]) with per-row exponents ex[row] into a
// section at dst (fragment order); one thread per (row fragment, lane) item,
// 256 threads.  Inlined and specialised on TR: as a called function with a
// run-time flag every element paid for the select and the call for its
// register saves.  A row stored as zero (exponent -127) is sliced as 0.0
// with exponent 0, so that slice4 takes no branch; its stored exponent stays
// -127.
template <bool TR>
__device__ __forceinline__ void store_section(const double (*v)[65], const int *ex,
                                              signed char *dst) {
  const int t = threadIdx.x;
  const int rt = t >> 6, l = t & 63;
  const int row = 16 * rt + (l & 15), k0 = 16 * (l >> 4);
  const int er = ex[row];
  const bool zero = er == -127;
  __attribute__((aligned(16))) signed char out[4][16];
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const double x = TR ? v[k0 + e][row] : v[row][k0 + e];
    signed char s[4];
    slice4(zero ? 0.0 : x, zero ? 0 : er, s);
#pragma unroll
    for (int p = 0; p < 4; ++p) out[p][e] = s[p];
  }
#pragma unroll
  for (int p = 0; p < 4; ++p)
    *reinterpret_cast<v4i *>(dst + p * PLANE + t * 16) = *reinterpret_cast<const v4i *>(out[p]);
  if (t < 64) dst[4 * PLANE + t] = (signed char)ex[t];
}

// Direct global -> LDS copy of one section (4 planes + the 256-byte exponent
// area) by the NW waves of a workgroup, this one taking pieces w0, w0 + NW,
// ...: 16 pieces of 1 KB (global_load_lds_dwordx4: lane l writes its 16 bytes
// at dst + 16 l) and one of 256 bytes (dword form).  The fixed trip count
// keeps the copies straight-line code, so the compiler's waits after them
// stay exact.  The copies are tracked by vmcnt only; the consumer waits with
// wait_vm0() before the barrier that publishes the LDS.
template <int NW>
__device__ __forceinline__ void dma_section(const signed char *src, signed char *dst_lds, int w0,
                                            int lane) {
#pragma unroll
  for (int k = 0; k < (17 + NW - 1) / NW; ++k) {
    const int pc = w0 + k * NW;
    if (pc < 16)
      __builtin_amdgcn_global_load_lds(src + pc * 1024 + lane * 16, dst_lds + pc * 1024, 16, 0,
                                       0);
    else if (pc == 16)
      __builtin_amdgcn_global_load_lds(src + 16384 + lane * 4, dst_lds + 16384, 4, 0, 0);
  }
}
// s_waitcnt vmcnt(0) (expcnt and lgkmcnt left at their maxima)
__device__ __forceinline__ void wait_vm0() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// 2^e as a double, e in the normal range
__device__ __forceinline__ double pow2(int e) {
  return __longlong_as_double((long long)(e + 1023) << 52);
}
// 2^e, or 0 for the zero-row exponent marker -127
__device__ __forceinline__ double pow2_or_zero(int e) { return e == -127 ? 0.0 : pow2(e); }
}  // namespace
}  // namespace nkx
