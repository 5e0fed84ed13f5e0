// Batched V_xc build on MI355X (gfx950), mixed form.
//
// Collocation values are FP32 (exponential in FP32 on an FP64 range-reduced
// argument, see sto_value_mixed in sto_eval.hpp).  Both contractions run
// on v_mfma_f32_16x16x32_bf16 with every FP32 operand split into three bf16
// terms, x = x0 + x1 + x2 (x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0
// - x1); Henry, Tang & Heinecke, ARITH 2019), and the six products
// a0 b0 + a0 b1 + a1 b0 + a0 b2 + a1 b1 + a2 b0 accumulated in FP32.  The
// dropped terms are O(2^-24) relative.  FP64 is kept where the error budget
// needs it: every FP32 MFMA accumulator covers at most one 64-function block
// (density contraction) or one 32-point stage (V contraction) before it is
// added into FP64 registers, so FP32 accumulation error does not grow with
// the number of functions or points; the functional and the E_xc / N_el / V
// reductions are FP64.
//
// Same data layout and launch structure as the FP64 form (kernels_vxc.hip);
// tiles are FP32.
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
constexpr int TILE = FB * PT;
constexpr int LDR = 72;  // padded LDS row (bf16 elements)

// float -> bf16 with round-to-nearest-even, done on the bits so the rounding
// mode is explicit (finite inputs).  Rounding to nearest makes each residual
// of the split symmetric about zero; with truncation all three terms would
// share the sign of x and the dropped products would bias every sum.
__device__ __forceinline__ __bf16 bf16_rne(float x, float &back) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  back = __uint_as_float(u);
  return __builtin_bit_cast(__bf16, (unsigned short)(u >> 16));
}

__device__ __forceinline__ void split3(float x, __bf16 &a, __bf16 &b, __bf16 &c) {
  float xa, xb, xc;
  a = bf16_rne(x, xa);
  const float r1 = x - xa;  // exact
  b = bf16_rne(r1, xb);
  const float r2 = r1 - xb;  // exact
  c = bf16_rne(r2, xc);
}

template <int NPROD>
__device__ __forceinline__ v4f mfma_split(const v8bf *a, const v8bf *b, v4f acc) {
  // small terms first
  if (NPROD == 6) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[2], acc, 0, 0, 0);
  }
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0], acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[1], acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0], acc, 0, 0, 0);
  return acc;
}
}  // namespace

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) colloc_tiles_f32(
    int64_t s0, const int *__restrict__ slot_b, const int *__restrict__ b_fbl,
    const FuncRec *__restrict__ F, const HarmTerm *__restrict__ T,
    const double *__restrict__ px, const double *__restrict__ py,
    const double *__restrict__ pz, int narr, int gga, float *__restrict__ tiles) {
  const int64_t s = s0 + blockIdx.x;
  const int b = slot_b[s];
  const int fb = b_fbl[s];
  __shared__ FuncRec fr[FB];
  if (threadIdx.x < FB) fr[threadIdx.x] = F[fb * FB + threadIdx.x];
  __syncthreads();
  const int g = threadIdx.x & (PT - 1);
  const int row = threadIdx.x / PT;
  const int64_t gp = (int64_t)b * PT + g;
  const double x = px[gp], y = py[gp], z = pz[gp];
  float *base = tiles + (int64_t)blockIdx.x * narr * TILE;
#pragma unroll 1
  for (int f = row; f < FB; f += 256 / PT) {
    const FuncRec rec = fr[f];
    const int e = f * PT + g;
    if (gga) {
      float v, a, bb, c;
      sto_value_grad_mixed(rec, T, x, y, z, v, a, bb, c);
      base[e] = v;
      base[TILE + e] = a;
      base[2 * TILE + e] = bb;
      base[3 * TILE + e] = c;
    } else {
      base[e] = sto_value_mixed(rec, T, x, y, z);
    }
  }
}

// P (FP64, original order) -> three bf16 planes in the permuted, padded order.
__global__ void __launch_bounds__(256) gather_p_split(int nbf, int nbf_pad,
                                                      const int *__restrict__ fperm,
                                                      const double *__restrict__ P,
                                                      __bf16 *__restrict__ Ps) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)nbf_pad * nbf_pad;
  if (e >= n) return;
  const int i = (int)(e / nbf_pad), j = (int)(e % nbf_pad);
  const double v = (i < nbf && j < nbf) ? P[(int64_t)fperm[i] * nbf + fperm[j]] : 0.0;
  __bf16 a, b, c;
  split3((float)v, a, b, c);
  Ps[e] = a;
  Ps[n + e] = b;
  Ps[2 * n + e] = c;
}

// ---------------------------------------------------------------------------
template <int NPROD>
__global__ void __launch_bounds__(256) density_fused_mixed(
    int nbf_pad, const __bf16 *__restrict__ Ps, const int *__restrict__ b_off,
    const int *__restrict__ b_fbl, int64_t s0, int64_t b0, int narr, int gga,
    float *__restrict__ tiles, const double *__restrict__ pw, XcList xl,
    double *__restrict__ coef, double *__restrict__ blockEN) {
  const int b = (int)(b0 + blockIdx.x);
  const int lo = b_off[b], hi = b_off[b + 1];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int l16 = lane & 15, q = lane >> 4;
  const int64_t plane = (int64_t)nbf_pad * nbf_pad;

  __shared__ __attribute__((aligned(16))) __bf16 Pl[3][FB][LDR];  // [plane][mu][nu]
  __shared__ __attribute__((aligned(16))) __bf16 Fl[3][PT][LDR];  // [plane][point][nu]
  __shared__ double red[4][4][PT];
  __shared__ double colsum[4][PT];
  __shared__ double cf[4][PT];

  for (int e = threadIdx.x; e < 4 * PT; e += 256) (&colsum[0][0])[e] = 0.0;
  __syncthreads();
  const int nq = gga ? 4 : 1;

  for (int mi = lo; mi < hi; ++mi) {
    const int fm = b_fbl[mi];
    double acc64[4][4];
    for (int j = 0; j < 4; ++j)
      for (int r = 0; r < 4; ++r) acc64[j][r] = 0.0;
    for (int ki = lo; ki < hi; ++ki) {
      const int fk = b_fbl[ki];
      v4f acc[4];
      for (int j = 0; j < 4; ++j) acc[j] = v4f{0.f, 0.f, 0.f, 0.f};
      // P planes: rows mu of block fm, 64 contiguous nu of block fk (16 B chunks)
      for (int c = threadIdx.x; c < 3 * FB * 8; c += 256) {
        const int pl = c / (FB * 8), rem = c % (FB * 8);
        const int r = rem >> 3, k8 = (rem & 7) * 8;
        *reinterpret_cast<v8bf *>(&Pl[pl][r][k8]) = *reinterpret_cast<const v8bf *>(
            Ps + pl * plane + (int64_t)(fm * FB + r) * nbf_pad + fk * FB + k8);
      }
      // Phi tile of block fk, transposed to [point][nu] and split
      {
        const float *tsrc = tiles + (ki - s0) * narr * TILE;
        const int p = threadIdx.x & 63, f0 = (threadIdx.x >> 6) * 16;
        __attribute__((aligned(16))) __bf16 h[3][16];
#pragma unroll
        for (int u = 0; u < 16; ++u) split3(tsrc[(f0 + u) * PT + p], h[0][u], h[1][u], h[2][u]);
#pragma unroll
        for (int pl = 0; pl < 3; ++pl) {
          *reinterpret_cast<v8bf *>(&Fl[pl][p][f0]) = *reinterpret_cast<v8bf *>(&h[pl][0]);
          *reinterpret_cast<v8bf *>(&Fl[pl][p][f0 + 8]) = *reinterpret_cast<v8bf *>(&h[pl][8]);
        }
      }
      __syncthreads();
#pragma unroll
      for (int ks = 0; ks < FB; ks += 32) {
        v8bf a[3];
#pragma unroll
        for (int pl = 0; pl < 3; ++pl)
          a[pl] = *reinterpret_cast<const v8bf *>(&Pl[pl][16 * wv + l16][ks + 8 * q]);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          v8bf bv[3];
#pragma unroll
          for (int pl = 0; pl < 3; ++pl)
            bv[pl] = *reinterpret_cast<const v8bf *>(&Fl[pl][16 * j + l16][ks + 8 * q]);
          acc[j] = mfma_split<NPROD>(a, bv, acc[j]);
        }
      }
      // FP32 accumulation spans one 64-function block only
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc64[j][r] += (double)acc[j][r];
      __syncthreads();
    }
    // Epilogue in FP64: column partial sums of Phi X (and grad Phi X).
    const float *tb = tiles + (mi - s0) * narr * TILE;
    double part[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int g = 16 * j + l16;
      double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int mu = 16 * wv + 4 * q + r;
        const double x = acc64[j][r];
        s[0] += (double)tb[mu * PT + g] * x;
        if (gga) {
          s[1] += (double)tb[TILE + mu * PT + g] * x;
          s[2] += (double)tb[2 * TILE + mu * PT + g] * x;
          s[3] += (double)tb[3 * TILE + mu * PT + g] * x;
        }
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        double v = s[c];
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
    if (g == 0) {
      blockEN[2 * b] = e;
      blockEN[2 * b + 1] = n;
    }
  }
  __syncthreads();

  for (int mi = lo; mi < hi; ++mi) {
    float *tb = tiles + (mi - s0) * narr * TILE;
    float *zt = tb + (narr - 1) * TILE;
    for (int e = threadIdx.x; e < TILE; e += 256) {
      const int g = e % PT;
      double z = cf[0][g] * (double)tb[e];
      if (gga)
        z += cf[1][g] * (double)tb[TILE + e] + cf[2][g] * (double)tb[2 * TILE + e] +
             cf[3][g] * (double)tb[3 * TILE + e];
      zt[e] = (float)z;
    }
  }
}

// ---------------------------------------------------------------------------
// acc(I, J) += Phi_I Z_J^T + Z_I Phi_J^T; FP32 within a batch, FP64 across.
constexpr int HPM = 32;        // points per LDS stage (one MFMA k-step)
constexpr int LDM = HPM + 8;   // padded row (bf16)

template <int NPROD>
__global__ void __launch_bounds__(256) vgemm_mixed(
    int npairs, int nsplit, const int *__restrict__ pair_I, const int *__restrict__ pair_J,
    const int *__restrict__ tile_slot, int64_t nb, int64_t b0, int64_t b1, int64_t s0,
    int narr, const float *__restrict__ tiles, double *__restrict__ vpart) {
  const int p = blockIdx.x % npairs;
  const int split = blockIdx.x / npairs;
  const int I = pair_I[p], J = pair_J[p];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int l16 = lane & 15, q = lane >> 4;
  // [operand][plane][function][point]: operands Phi_I, Z_I, Phi_J, Z_J
  __shared__ __attribute__((aligned(16))) __bf16 S[4][3][FB][LDM];
  double acc64[4][4];
  for (int j = 0; j < 4; ++j)
    for (int r = 0; r < 4; ++r) acc64[j][r] = 0.0;
  const int *tsI = tile_slot + (int64_t)I * nb;
  const int *tsJ = tile_slot + (int64_t)J * nb;
  for (int64_t b = b0 + split; b < b1; b += nsplit) {
    const int sI = tsI[b], sJ = tsJ[b];
    if (sI < 0 || sJ < 0) continue;
    const float *src[4] = {tiles + (sI - s0) * narr * TILE,
                           tiles + (sI - s0) * narr * TILE + (narr - 1) * TILE,
                           tiles + (sJ - s0) * narr * TILE,
                           tiles + (sJ - s0) * narr * TILE + (narr - 1) * TILE};
    for (int h = 0; h < PT; h += HPM) {
      v4f acc[4];
      for (int j = 0; j < 4; ++j) acc[j] = v4f{0.f, 0.f, 0.f, 0.f};
      // stage 4 operands x (64 functions x 32 points), split into 3 planes
#pragma unroll
      for (int op = 0; op < 4; ++op) {
        const int r = threadIdx.x >> 2, c8 = (threadIdx.x & 3) * 8;
        const float4 u0 = *reinterpret_cast<const float4 *>(src[op] + r * PT + h + c8);
        const float4 u1 = *reinterpret_cast<const float4 *>(src[op] + r * PT + h + c8 + 4);
        const float vals[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
        __attribute__((aligned(16))) __bf16 t[3][8];
#pragma unroll
        for (int u = 0; u < 8; ++u) split3(vals[u], t[0][u], t[1][u], t[2][u]);
#pragma unroll
        for (int pl = 0; pl < 3; ++pl)
          *reinterpret_cast<v8bf *>(&S[op][pl][r][c8]) = *reinterpret_cast<v8bf *>(&t[pl][0]);
      }
      __syncthreads();
      v8bf aphi[3], az[3];
#pragma unroll
      for (int pl = 0; pl < 3; ++pl) {
        aphi[pl] = *reinterpret_cast<const v8bf *>(&S[0][pl][16 * wv + l16][8 * q]);
        az[pl] = *reinterpret_cast<const v8bf *>(&S[1][pl][16 * wv + l16][8 * q]);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        v8bf bphi[3], bz[3];
#pragma unroll
        for (int pl = 0; pl < 3; ++pl) {
          bphi[pl] = *reinterpret_cast<const v8bf *>(&S[2][pl][16 * j + l16][8 * q]);
          bz[pl] = *reinterpret_cast<const v8bf *>(&S[3][pl][16 * j + l16][8 * q]);
        }
        acc[j] = mfma_split<NPROD>(aphi, bz, acc[j]);
        acc[j] = mfma_split<NPROD>(az, bphi, acc[j]);
      }
      // FP32 accumulation spans one 32-point stage only
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc64[j][r] += (double)acc[j][r];
      __syncthreads();
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
void vkm_gather_p(hipStream_t s, int nbf, int nbf_pad, const int *fperm, const double *P,
                  uint16_t *Ps) {
  const int64_t n = (int64_t)nbf_pad * nbf_pad;
  hipLaunchKernelGGL(gather_p_split, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nbf,
                     nbf_pad, fperm, P, reinterpret_cast<__bf16 *>(Ps));
  NKX_HIP_CHECK(hipGetLastError());
}

void vkm_colloc(hipStream_t s, int64_t s0, int64_t s1, const Plan &pl, int narr, int gga,
                float *tiles) {
  if (s1 <= s0) return;
  hipLaunchKernelGGL(colloc_tiles_f32, dim3((unsigned)(s1 - s0)), dim3(256), 0, s, s0,
                     pl.slot_b.p, pl.b_fbl.p, pl.funcs.p, pl.terms.p, pl.px.p, pl.py.p, pl.pz.p,
                     narr, gga, tiles);
  NKX_HIP_CHECK(hipGetLastError());
}

void vkm_density(hipStream_t s, int nprod, int64_t b0, int64_t b1, int64_t s0, const Plan &pl,
                 const uint16_t *Ps16, int narr, int gga, float *tiles, const XcList &xl,
                 double *coef, double *blockEN) {
  if (b1 <= b0) return;
  const __bf16 *Ps = reinterpret_cast<const __bf16 *>(Ps16);
  if (nprod == 3)
    hipLaunchKernelGGL(density_fused_mixed<3>, dim3((unsigned)(b1 - b0)), dim3(256), 0, s,
                       pl.nbf_pad, Ps, pl.b_off.p, pl.b_fbl.p, s0, b0, narr, gga, tiles, pl.pw.p,
                       xl, coef, blockEN);
  else
    hipLaunchKernelGGL(density_fused_mixed<6>, dim3((unsigned)(b1 - b0)), dim3(256), 0, s,
                       pl.nbf_pad, Ps, pl.b_off.p, pl.b_fbl.p, s0, b0, narr, gga, tiles, pl.pw.p,
                       xl, coef, blockEN);
  NKX_HIP_CHECK(hipGetLastError());
}

void vkm_vgemm(hipStream_t s, int nprod, int64_t b0, int64_t b1, int64_t s0, const Plan &pl,
               int nsplit, int narr, const float *tiles, double *vpart) {
  if (b1 <= b0 || pl.npairs == 0) return;
  if (nprod == 3)
    hipLaunchKernelGGL(vgemm_mixed<3>, dim3((unsigned)(pl.npairs * nsplit)), dim3(256), 0, s,
                       pl.npairs, nsplit, pl.pair_I.p, pl.pair_J.p, pl.tile_slot.p, pl.nb, b0,
                       b1, s0, narr, tiles, vpart);
  else
    hipLaunchKernelGGL(vgemm_mixed<6>, dim3((unsigned)(pl.npairs * nsplit)), dim3(256), 0, s,
                       pl.npairs, nsplit, pl.pair_I.p, pl.pair_J.p, pl.tile_slot.p, pl.nb, b0,
                       b1, s0, narr, tiles, vpart);
  NKX_HIP_CHECK(hipGetLastError());
}

}  // namespace nkx
