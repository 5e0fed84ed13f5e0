// Batched V_xc build on MI355X (gfx950), split-bf16 mixed form (selected
// with NKX_PREC_MIXED_BF16; the default mixed form is kernels_vxc_i8.hip).
//
// Collocation values, Z and P are FP64 numbers split into three bf16 terms,
// x ~ x0 + x1 + x2 (x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1),
// round to nearest even, residuals formed in FP64; |x - x0 - x1 - x2| is
// about 2^-27 |x|; Henry, Tang & Heinecke, ARITH 2019, split an FP32 number
// the same way).  Both contractions run on v_mfma_f32_16x16x32_bf16 with the
// six products a0 b0 + a0 b1 + a1 b0 + a0 b2 + a1 b1 + a2 b0 accumulated in
// FP32; the dropped terms are O(2^-27) relative.  Collocation gradients are
// FP32 (rounded once from FP64).  Every FP32 accumulator covers at most one
// 64-function block (density) or one 64-point batch (V) before it is added
// into FP64 registers, so FP32 accumulation error does not grow with the
// system.  The functional and the E_xc / N_el / V reductions are FP64.
//
// Operands are stored pre-split in MFMA-fragment order, so every fragment is
// one coalesced 16-byte load per lane and the main loops need no LDS staging
// and no barriers.  For a 64 x 64 tile T[row][k] a "fragment" (rt, ks) holds
// rows 16 rt .. 16 rt + 15 and k = 32 ks .. 32 ks + 31 as 64 lanes x 8 bf16:
//   lane l, element e  <->  T[16 rt + l % 16][32 ks + 8 (l / 16) + e],
// stored at ((rt * 2 + ks) * 64 + l) * 8 + e, per plane (plane-major).
// This is the A-operand order of v_mfma_f32_16x16x32_bf16 (A[i][k]) and, read
// as B[k][j] with j the row of T, its B-operand order.
//
// Per live slot the tile store holds (bytes per slot, GGA):
//   F   Phi, rows = functions, k = points        3 planes   24 KB
//   Pm  Phi, rows = points,    k = functions     3 planes   24 KB
//   G   dPhi/dx, dPhi/dy, dPhi/dz  FP32 [f][p]              48 KB
//   Z   Z,   rows = functions, k = points        3 planes   24 KB
// (LDA: F, Pm, Z.)  P is gathered into fragment order per 64 x 64 block pair.
#include "nkx_internal.hpp"
#include "sto_eval.hpp"
#include "vxc_kernels.hpp"
#include "xc_functionals.hpp"

namespace nkx {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// v_mfma_f32_16x16x32_bf16 lane layout (checked by tests/gpu_probe): lane l
// supplies A[l%16][8(l/16) .. 8(l/16)+7] and B[8(l/16) .. +7][l%16]; result
// register r of lane l holds D[4(l/16) + r][l%16].

namespace {
constexpr int FB = 64;
constexpr int PT = 64;
constexpr int TILE = FB * PT;          // elements per tile
constexpr int PLANES = 3 * TILE;       // bf16 elements of a 3-plane tile

__device__ __forceinline__ int frag_off(int row, int k) {
  // offset of T[row][k] inside one plane of a fragment-ordered 64 x 64 tile
  const int rt = row >> 4, i = row & 15, ks = k >> 5, kq = (k >> 3) & 3, e = k & 7;
  return (((rt * 2 + ks) * 64) + i + 16 * kq) * 8 + e;
}

// float -> bf16, round to nearest even on the bits (finite inputs).  Rounding
// to nearest keeps each residual of the split symmetric about zero; with
// truncation the three terms would share the sign of x and the dropped
// products would bias every sum.
__device__ __forceinline__ unsigned short bf16_rne(float x, float &back) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  back = __uint_as_float(u);
  return (unsigned short)(u >> 16);
}

// Three-term split of an FP64 value; the residuals are exact in FP64, each
// term is the bf16 rounding of the FP32 rounding of the current residual.
__device__ __forceinline__ void split3(double x, unsigned short &a, unsigned short &b,
                                       unsigned short &c) {
  float xa, xb, xc;
  a = bf16_rne((float)x, xa);
  const double r1 = x - (double)xa;
  b = bf16_rne((float)r1, xb);
  const double r2 = r1 - (double)xb;
  c = bf16_rne((float)r2, xc);
}

__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }

// value of T[row][k] from a 3-plane fragment-ordered tile (exact in FP64)
__device__ __forceinline__ double tile_value(const unsigned short *t, int row, int k) {
  const int o = frag_off(row, k);
  return ((double)bf2f(t[o + 2 * TILE]) + (double)bf2f(t[o + TILE])) + (double)bf2f(t[o]);
}

__device__ __forceinline__ v8bf ld_frag(const unsigned short *p) {
  return *reinterpret_cast<const v8bf *>(p);
}

template <int NPROD>
__device__ __forceinline__ v4f mfma_split(const v8bf *a, const v8bf *b, v4f acc) {
  if (NPROD == 6) {  // small terms first
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[2], acc, 0, 0, 0);
  }
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0], acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[1], acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0], acc, 0, 0, 0);
  return acc;
}

struct MixedLayout {
  int64_t slot_elems;  // bf16-sized elements per slot
  int64_t off_F, off_P, off_G, off_Z;
};

__host__ __device__ inline MixedLayout mixed_layout(int gga) {
  MixedLayout L;
  L.off_F = 0;
  L.off_P = PLANES;
  L.off_G = 2 * PLANES;                          // FP32 data, 2 bf16 units each
  L.off_Z = 2 * PLANES + (gga ? 3 * TILE * 2 : 0);
  L.slot_elems = L.off_Z + PLANES;
  return L;
}
}  // namespace

int64_t bf16_slot_bytes(int gga) { return mixed_layout(gga).slot_elems * 2; }

// ---------------------------------------------------------------------------
// Collocation: one workgroup per live slot.  Values are computed in FP64,
// split, staged in LDS as [plane][f][p] and written out in both fragment
// orders.
__global__ void __launch_bounds__(256) colloc_tiles_mixed(
    int64_t s0, const int *__restrict__ slot_b, const int *__restrict__ b_fbl,
    const FuncRec *__restrict__ F, const HarmTerm *__restrict__ T,
    const double *__restrict__ px, const double *__restrict__ py,
    const double *__restrict__ pz, int gga, unsigned short *__restrict__ store) {
  const int64_t s = s0 + blockIdx.x;
  const int b = slot_b[s];
  const int fb = b_fbl[s];
  const MixedLayout L = mixed_layout(gga);
  __shared__ FuncRec fr[FB];
  __shared__ __attribute__((aligned(16))) unsigned short sp[3][FB][PT];
  if (threadIdx.x < FB) fr[threadIdx.x] = F[fb * FB + threadIdx.x];
  __syncthreads();
  const int g = threadIdx.x & (PT - 1);
  const int row = threadIdx.x / PT;
  const int64_t gp = (int64_t)b * PT + g;
  const double x = px[gp], y = py[gp], z = pz[gp];
  unsigned short *base = store + (int64_t)blockIdx.x * L.slot_elems;
  float *G = reinterpret_cast<float *>(base + L.off_G);
#pragma unroll 1
  for (int f = row; f < FB; f += 256 / PT) {
    const FuncRec rec = fr[f];
    double v;
    if (gga) {
      float a, bb, c;
      sto_value_grad_mixed(rec, T, x, y, z, v, a, bb, c);
      G[f * PT + g] = a;
      G[TILE + f * PT + g] = bb;
      G[2 * TILE + f * PT + g] = c;
    } else {
      v = sto_value_mixed(rec, T, x, y, z);
    }
    split3(v, sp[0][f][g], sp[1][f][g], sp[2][f][g]);
  }
  __syncthreads();
  // 8 fragments x 64 lanes per plane; one 16-byte chunk per (plane, frag, lane)
  for (int c = threadIdx.x; c < 3 * 8 * 64; c += 256) {
    const int pl = c / 512, fr_ = (c / 64) % 8, l = c % 64;
    const int rt = fr_ >> 1, ks = fr_ & 1;
    const int rowF = 16 * rt + (l & 15), k0 = 32 * ks + 8 * (l >> 4);
    // F order: rows = functions, k = points (contiguous in sp)
    *reinterpret_cast<v8bf *>(base + L.off_F + pl * TILE + c % 512 * 8) =
        *reinterpret_cast<const v8bf *>(&sp[pl][rowF][k0]);
    // Pm order: rows = points, k = functions (gather down a column of sp)
    __attribute__((aligned(16))) unsigned short t[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) t[e] = sp[pl][k0 + e][rowF];
    *reinterpret_cast<v8bf *>(base + L.off_P + pl * TILE + c % 512 * 8) =
        *reinterpret_cast<const v8bf *>(t);
  }
}

// P (FP64, original order) -> three bf16 planes, permuted and padded, in
// fragment order per 64 x 64 block pair: [plane][fm][fk][fragment tile].
__global__ void __launch_bounds__(256) gather_p_split(int nbf, int nfb,
                                                      const int *__restrict__ fperm,
                                                      const double *__restrict__ P,
                                                      unsigned short *__restrict__ Ps) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nbf_pad = nfb * FB;
  const int64_t n = (int64_t)nbf_pad * nbf_pad;
  if (e >= n) return;
  const int i = (int)(e / nbf_pad), j = (int)(e % nbf_pad);
  const double v = (i < nbf && j < nbf) ? P[(int64_t)fperm[i] * nbf + fperm[j]] : 0.0;
  unsigned short a, b, c;
  split3(v, a, b, c);
  const int fm = i / FB, fk = j / FB;
  const int64_t o = ((int64_t)fm * nfb + fk) * TILE + frag_off(i % FB, j % FB);
  Ps[o] = a;
  Ps[n + o] = b;
  Ps[2 * n + o] = c;
}

// ---------------------------------------------------------------------------
// Density contraction, functional and Z.  One workgroup per batch; wave w owns
// rows 16 w .. 16 w + 15 of the current M-block and all 64 points.
template <int NPROD>
__global__ void __launch_bounds__(256) density_fused_mixed(
    int nfb, const unsigned short *__restrict__ Ps, const int *__restrict__ b_off,
    const int *__restrict__ b_fbl, int64_t s0, int64_t b0, int gga,
    unsigned short *__restrict__ store, const double *__restrict__ pw, XcList xl,
    double *__restrict__ coef, double *__restrict__ blockEN) {
  const int b = (int)(b0 + blockIdx.x);
  const int lo = b_off[b], hi = b_off[b + 1];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int l16 = lane & 15, q = lane >> 4;
  const MixedLayout L = mixed_layout(gga);
  const int64_t plane = (int64_t)nfb * nfb * TILE;

  __shared__ double red[4][4][PT];
  __shared__ double colsum[4][PT];
  __shared__ double cf[4][PT];
  for (int e = threadIdx.x; e < 4 * PT; e += 256) (&colsum[0][0])[e] = 0.0;
  __syncthreads();
  const int nq = gga ? 4 : 1;

  for (int mi = lo; mi < hi; ++mi) {
    const int fm = b_fbl[mi];
    double acc64[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc64[j][r] = 0.0;
    for (int ki = lo; ki < hi; ++ki) {
      const int fk = b_fbl[ki];
      const unsigned short *pa = Ps + ((int64_t)fm * nfb + fk) * TILE;
      const unsigned short *pb = store + (ki - s0) * L.slot_elems + L.off_P;
      v4f acc[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        v8bf a[3];
#pragma unroll
        for (int pl = 0; pl < 3; ++pl)
          a[pl] = ld_frag(pa + pl * plane + ((wv * 2 + ks) * 64 + lane) * 8);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          v8bf bv[3];
#pragma unroll
          for (int pl = 0; pl < 3; ++pl)
            bv[pl] = ld_frag(pb + pl * TILE + ((j * 2 + ks) * 64 + lane) * 8);
          acc[j] = mfma_split<NPROD>(a, bv, acc[j]);
        }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc64[j][r] += (double)acc[j][r];
    }
    // Epilogue in FP64: column partial sums of Phi X (and grad Phi X).
    const unsigned short *tb = store + (mi - s0) * L.slot_elems;
    const float *G = reinterpret_cast<const float *>(tb + L.off_G);
    double part[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int g = 16 * j + l16;
      double sm[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int mu = 16 * wv + 4 * q + r;
        const double x = acc64[j][r];
        sm[0] += tile_value(tb + L.off_F, mu, g) * x;
        if (gga) {
          sm[1] += (double)G[mu * PT + g] * x;
          sm[2] += (double)G[TILE + mu * PT + g] * x;
          sm[3] += (double)G[2 * TILE + mu * PT + g] * x;
        }
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        double v = sm[c];
        v += __shfl_xor(v, 16);
        v += __shfl_xor(v, 32);
        part[j][c] = v;
      }
    }
    if (q == 0) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        for (int c = 0; c < 4; ++c) red[wv][c][16 * j + l16] = part[j][c];
    }
    __syncthreads();
    {
      const int c = threadIdx.x / PT, g = threadIdx.x % PT;
      if (c < nq) colsum[c][g] += red[0][c][g] + red[1][c][g] + red[2][c][g] + red[3][c][g];
    }
    __syncthreads();
  }

  if (threadIdx.x < PT) {
    const int g = threadIdx.x;
    const int64_t gp = (int64_t)b * PT + g;
    const double w = pw[gp];
    const double rho = colsum[0][g];
    const double gxr = 2.0 * colsum[1][g], gyr = 2.0 * colsum[2][g], gzr = 2.0 * colsum[3][g];
    const double sigma = gga ? gxr * gxr + gyr * gyr + gzr * gzr : 0.0;
    const XcPoint o = xc_eval(xl.n, xl.ids, xl.thr, rho, sigma);
    const double a = 0.5 * w * o.vrho;
    const double bs = gga ? 2.0 * w * o.vsigma : 0.0;
    cf[0][g] = a;
    cf[1][g] = bs * gxr;
    cf[2][g] = bs * gyr;
    cf[3][g] = bs * gzr;
    coef[4 * gp + 0] = a;
    coef[4 * gp + 1] = bs * gxr;
    coef[4 * gp + 2] = bs * gyr;
    coef[4 * gp + 3] = bs * gzr;
    double e = w * o.e, n = w * rho;
    for (int o2 = 32; o2 > 0; o2 >>= 1) {
      e += __shfl_down(e, o2);
      n += __shfl_down(n, o2);
    }
    if (g == 0) {  // (E, N_alpha, N_beta) per batch, N_beta = 0 here
      blockEN[3 * b] = e;
      blockEN[3 * b + 1] = n;
      blockEN[3 * b + 2] = 0.0;
    }
  }
  __syncthreads();

  // Z = a Phi + b . grad Phi in FP64, split, fragment order (F).
  for (int mi = lo; mi < hi; ++mi) {
    unsigned short *tb = store + (mi - s0) * L.slot_elems;
    const float *G = reinterpret_cast<const float *>(tb + L.off_G);
    for (int c = threadIdx.x; c < 8 * 64; c += 256) {
      const int fr_ = c / 64, l = c % 64;
      const int rt = fr_ >> 1, ks = fr_ & 1;
      const int f = 16 * rt + (l & 15), k0 = 32 * ks + 8 * (l >> 4);
      __attribute__((aligned(16))) unsigned short t0[8], t1[8], t2[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int g = k0 + e;
        double zz = cf[0][g] * tile_value(tb + L.off_F, f, g);
        if (gga)
          zz += cf[1][g] * (double)G[f * PT + g] + cf[2][g] * (double)G[TILE + f * PT + g] +
                cf[3][g] * (double)G[2 * TILE + f * PT + g];
        split3(zz, t0[e], t1[e], t2[e]);
      }
      unsigned short *zt = tb + L.off_Z + c * 8;
      *reinterpret_cast<v8bf *>(zt) = *reinterpret_cast<const v8bf *>(t0);
      *reinterpret_cast<v8bf *>(zt + TILE) = *reinterpret_cast<const v8bf *>(t1);
      *reinterpret_cast<v8bf *>(zt + 2 * TILE) = *reinterpret_cast<const v8bf *>(t2);
    }
  }
}

// ---------------------------------------------------------------------------
// acc(I, J) += Phi_I Z_J^T + Z_I Phi_J^T over batches where I and J are live.
// Wave w owns rows 16 w .. 16 w + 15 of block I and all 64 columns of block J.
template <int NPROD>
__global__ void __launch_bounds__(256) vgemm_mixed(
    int npairs, int nsplit, const int *__restrict__ pair_I, const int *__restrict__ pair_J,
    const int *__restrict__ tile_slot, int64_t nb, int64_t b0, int64_t b1, int64_t s0,
    int gga, const unsigned short *__restrict__ store, double *__restrict__ vpart) {
  const int p = blockIdx.x % npairs;
  const int split = blockIdx.x / npairs;
  const int I = pair_I[p], J = pair_J[p];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int l16 = lane & 15, q = lane >> 4;
  const MixedLayout L = mixed_layout(gga);
  double acc64[4][4];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) acc64[j][r] = 0.0;
  const int *tsI = tile_slot + (int64_t)I * nb;
  const int *tsJ = tile_slot + (int64_t)J * nb;
  for (int64_t b = b0 + split; b < b1; b += nsplit) {
    const int sI = tsI[b], sJ = tsJ[b];
    if (sI < 0 || sJ < 0) continue;
    const unsigned short *tI = store + (sI - s0) * L.slot_elems;
    const unsigned short *tJ = store + (sJ - s0) * L.slot_elems;
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      v8bf aphi[3], az[3];
#pragma unroll
      for (int pl = 0; pl < 3; ++pl) {
        const int o = pl * TILE + ((wv * 2 + ks) * 64 + lane) * 8;
        aphi[pl] = ld_frag(tI + L.off_F + o);
        az[pl] = ld_frag(tI + L.off_Z + o);
      }
      v4f acc[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        acc[j] = v4f{0.f, 0.f, 0.f, 0.f};
        v8bf bphi[3], bz[3];
#pragma unroll
        for (int pl = 0; pl < 3; ++pl) {
          const int o = pl * TILE + ((j * 2 + ks) * 64 + lane) * 8;
          bphi[pl] = ld_frag(tJ + L.off_F + o);
          bz[pl] = ld_frag(tJ + L.off_Z + o);
        }
        acc[j] = mfma_split<NPROD>(aphi, bz, acc[j]);
        acc[j] = mfma_split<NPROD>(az, bphi, acc[j]);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc64[j][r] += (double)acc[j][r];
    }
  }
  double *out = vpart + ((int64_t)split * npairs + p) * (FB * FB);
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int mu = 16 * wv + 4 * q + r, nu = 16 * j + l16;
      out[mu * FB + nu] += acc64[j][r];
    }
}

// ---------------------------------------------------------------------------
void vkb_gather_p(hipStream_t s, int nbf, int nfb, const int *fperm, const double *P,
                  uint16_t *Ps) {
  const int64_t n = (int64_t)nfb * FB * nfb * FB;
  hipLaunchKernelGGL(gather_p_split, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nbf,
                     nfb, fperm, P, Ps);
  NKX_HIP_CHECK(hipGetLastError());
}

void vkb_colloc(hipStream_t s, int64_t s0, int64_t s1, const Plan &pl, int gga,
                uint16_t *store) {
  if (s1 <= s0) return;
  hipLaunchKernelGGL(colloc_tiles_mixed, dim3((unsigned)(s1 - s0)), dim3(256), 0, s, s0,
                     pl.slot_b.p, pl.b_fbl.p, pl.funcs.p, pl.terms.p, pl.px.p, pl.py.p, pl.pz.p,
                     gga, store);
  NKX_HIP_CHECK(hipGetLastError());
}

void vkb_density(hipStream_t s, int nprod, int64_t b0, int64_t b1, int64_t s0, const Plan &pl,
                 const uint16_t *Ps, int gga, uint16_t *store, const XcList &xl, double *coef,
                 double *blockEN) {
  if (b1 <= b0) return;
  if (nprod == 3)
    hipLaunchKernelGGL(density_fused_mixed<3>, dim3((unsigned)(b1 - b0)), dim3(256), 0, s,
                       pl.nfb, Ps, pl.b_off.p, pl.b_fbl.p, s0, b0, gga, store, pl.pw.p, xl, coef,
                       blockEN);
  else
    hipLaunchKernelGGL(density_fused_mixed<6>, dim3((unsigned)(b1 - b0)), dim3(256), 0, s,
                       pl.nfb, Ps, pl.b_off.p, pl.b_fbl.p, s0, b0, gga, store, pl.pw.p, xl, coef,
                       blockEN);
  NKX_HIP_CHECK(hipGetLastError());
}

void vkb_vgemm(hipStream_t s, int nprod, int64_t b0, int64_t b1, int64_t s0, const Plan &pl,
               int nsplit, int gga, const uint16_t *store, double *vpart) {
  if (b1 <= b0 || pl.npairs == 0) return;
  if (nprod == 3)
    hipLaunchKernelGGL(vgemm_mixed<3>, dim3((unsigned)(pl.npairs * nsplit)), dim3(256), 0, s,
                       pl.npairs, nsplit, pl.pair_I.p, pl.pair_J.p, pl.tile_slot.p, pl.nb, b0,
                       b1, s0, gga, store, vpart);
  else
    hipLaunchKernelGGL(vgemm_mixed<6>, dim3((unsigned)(pl.npairs * nsplit)), dim3(256), 0, s,
                       pl.npairs, nsplit, pl.pair_I.p, pl.pair_J.p, pl.tile_slot.p, pl.nb, b0,
                       b1, s0, gga, store, vpart);
  NKX_HIP_CHECK(hipGetLastError());
}

}  // namespace nkx
