// Batched V_xc build on MI355X (gfx950), FP64 form.
//
// Data layout in HBM (one chunk of point batches at a time):
//   tile store   [slot][array][64 functions][PT points], array = Phi, dPhi/dx,
//                dPhi/dy, dPhi/dz, Z (GGA) or Phi, Z (LDA).  A slot is one
//                (function block, point batch) pair that survived screening.
//   Pd           permuted, zero-padded density matrix (nbf_pad x nbf_pad).
//   vpart        split-K partial V tiles [split][pair][64][64].
//
// Kernels per chunk:
//   colloc_tiles     Phi and grad Phi for every live slot (HBM-write bound).
//   density_fused    one workgroup per batch: X = P Phi on MFMA over the
//                    batch's live function blocks, rho / grad rho epilogue in
//                    FP64, the functional, then Z = w (v_rho/2 Phi + 2 v_sigma
//                    grad rho . grad Phi) written back to the tile store.
//   vgemm            one workgroup per (function-block pair I <= J, split):
//                    acc += Phi_I Z_J^T + Z_I Phi_J^T over the batches where
//                    both blocks are live (MFMA-bound).
// and once per build: vreduce (split sum + un-permute + mirror) and the FP64
// energy / electron-count reduction.
#include "nkx_internal.hpp"
#include "sto_eval.hpp"
#include "vxc_kernels.hpp"
#include "xc_functionals.hpp"

namespace nkx {

typedef double v4d __attribute__((ext_vector_type(4)));

// v_mfma_f64_16x16x4f64 lane layout (checked by tests/gpu_probe): lane l
// supplies A[l%16][l/16] and B[l/16][l%16]; result register r of lane l
// holds D[4r + l/16][l%16].

constexpr int FB = 64;   // functions per block
constexpr int PT = 64;   // points per batch
constexpr int TILE = FB * PT;

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) colloc_tiles_f64(
    int64_t s0, const int *__restrict__ slot_b, const int *__restrict__ b_fbl,
    const FuncRec *__restrict__ F, const HarmTerm *__restrict__ T,
    const double *__restrict__ px, const double *__restrict__ py,
    const double *__restrict__ pz, int narr, int gga, double *__restrict__ tiles) {
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
  double *base = tiles + (int64_t)blockIdx.x * narr * TILE;
#pragma unroll 1
  for (int f = row; f < FB; f += 256 / PT) {
    const FuncRec rec = fr[f];
    const int e = f * PT + g;
    if (gga) {
      double v, a, bb, c;
      sto_value_grad<double>(rec, T, x, y, z, v, a, bb, c);
      base[e] = v;
      base[TILE + e] = a;
      base[2 * TILE + e] = bb;
      base[3 * TILE + e] = c;
    } else {
      base[e] = sto_value<double>(rec, T, x, y, z);
    }
  }
}

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) density_fused_f64(
    int nbf_pad, const double *__restrict__ Pd, const int *__restrict__ b_off,
    const int *__restrict__ b_fbl, int64_t s0, int64_t b0, int narr, int gga,
    double *__restrict__ tiles, const double *__restrict__ pw, XcList xl,
    double *__restrict__ coef, double *__restrict__ blockEN) {
  const int b = (int)(b0 + blockIdx.x);
  const int lo = b_off[b], hi = b_off[b + 1];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int l16 = lane & 15, q = lane >> 4;

  __shared__ double Pk[FB][FB];       // Pk[k][i] = P(mu = i, nu = k) of the block pair
  __shared__ double Ph[FB][PT];       // Phi tile of the K block
  __shared__ double red[4][4][PT];    // per-wave column partials
  __shared__ double colsum[4][PT];    // rho, d/dx, d/dy, d/dz column sums
  __shared__ double cf[4][PT];        // Z coefficients

  for (int e = threadIdx.x; e < 4 * PT; e += 256) (&colsum[0][0])[e] = 0.0;
  __syncthreads();

  const int nq = gga ? 4 : 1;
  for (int mi = lo; mi < hi; ++mi) {
    const int fm = b_fbl[mi];
    v4d acc[4];
    for (int j = 0; j < 4; ++j) acc[j] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int ki = lo; ki < hi; ++ki) {
      const int fk = b_fbl[ki];
      {
        // P block: rows nu in block fk, cols mu in block fm (P symmetric), and
        // the Phi tile of block fk; consecutive lanes move consecutive double2.
        const double *psrc = Pd + (int64_t)(fk * FB) * nbf_pad + fm * FB;
        const double *tsrc = tiles + (ki - s0) * narr * TILE;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int e = 2 * (u * 256 + (int)threadIdx.x);
          const int r = e >> 6, c = e & 63;
          *reinterpret_cast<double2 *>(&Pk[r][c]) =
              *reinterpret_cast<const double2 *>(psrc + (int64_t)r * nbf_pad + c);
          *reinterpret_cast<double2 *>(&Ph[r][c]) =
              *reinterpret_cast<const double2 *>(tsrc + e);
        }
      }
      __syncthreads();
#pragma unroll 4
      for (int k0 = 0; k0 < FB; k0 += 4) {
        const double a = Pk[k0 + q][16 * wv + l16];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const double bv = Ph[k0 + q][16 * j + l16];
          acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc[j], 0, 0, 0);
        }
      }
      __syncthreads();
    }
    // Epilogue: column partial sums sum_mu Phi(mu,g) X(mu,g) (and gradients).
    const double *tb = tiles + (mi - s0) * narr * TILE;
    double part[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int g = 16 * j + l16;
      double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int mu = 16 * wv + 4 * r + q;
        const double x = acc[j][r];
        s[0] += tb[mu * PT + g] * x;
        if (gga) {
          s[1] += tb[TILE + mu * PT + g] * x;
          s[2] += tb[2 * TILE + mu * PT + g] * x;
          s[3] += tb[3 * TILE + mu * PT + g] * x;
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

  // Functional and Z coefficients, one thread per point (wave 0).
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

  // Z tiles.
  for (int mi = lo; mi < hi; ++mi) {
    double *tb = tiles + (mi - s0) * narr * TILE;
    double *zt = tb + (narr - 1) * TILE;
    for (int e = threadIdx.x; e < TILE; e += 256) {
      const int g = e % PT;
      double z = cf[0][g] * tb[e];
      if (gga) z += cf[1][g] * tb[TILE + e] + cf[2][g] * tb[2 * TILE + e] + cf[3][g] * tb[3 * TILE + e];
      zt[e] = z;
    }
  }
}

// ---------------------------------------------------------------------------
// acc(I, J) += Phi_I Z_J^T + Z_I Phi_J^T over batches where I and J are live.
constexpr int HP = 32;      // points per LDS stage
constexpr int LDP = HP + 1; // padded row

__global__ void __launch_bounds__(256) vgemm_f64(
    int npairs, int nsplit, const int *__restrict__ pair_I, const int *__restrict__ pair_J,
    const int *__restrict__ tile_slot, int64_t nb, int64_t b0, int64_t b1, int64_t s0,
    int narr, const double *__restrict__ tiles, double *__restrict__ vpart) {
  const int p = blockIdx.x % npairs;
  const int split = blockIdx.x / npairs;
  const int I = pair_I[p], J = pair_J[p];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int l16 = lane & 15, q = lane >> 4;
  __shared__ double A0[FB][LDP], A1[FB][LDP], B0[FB][LDP], B1[FB][LDP];
  v4d acc[4];
  for (int j = 0; j < 4; ++j) acc[j] = v4d{0.0, 0.0, 0.0, 0.0};
  const int *tsI = tile_slot + (int64_t)I * nb;
  const int *tsJ = tile_slot + (int64_t)J * nb;
  for (int64_t b = b0 + split; b < b1; b += nsplit) {
    const int sI = tsI[b], sJ = tsJ[b];
    if (sI < 0 || sJ < 0) continue;
    const double *tI = tiles + (sI - s0) * narr * TILE;
    const double *tJ = tiles + (sJ - s0) * narr * TILE;
    for (int h = 0; h < PT; h += HP) {
      {
        const double *zI = tI + (narr - 1) * TILE, *zJ = tJ + (narr - 1) * TILE;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int e = 2 * (u * 256 + (int)threadIdx.x);
          const int r = e / HP, c = e % HP;
          const int off = r * PT + h + c;
          const double2 a0 = *reinterpret_cast<const double2 *>(tI + off);
          const double2 a1 = *reinterpret_cast<const double2 *>(zI + off);
          const double2 b0v = *reinterpret_cast<const double2 *>(tJ + off);
          const double2 b1v = *reinterpret_cast<const double2 *>(zJ + off);
          A0[r][c] = a0.x;  A0[r][c + 1] = a0.y;
          A1[r][c] = a1.x;  A1[r][c + 1] = a1.y;
          B0[r][c] = b0v.x; B0[r][c + 1] = b0v.y;
          B1[r][c] = b1v.x; B1[r][c + 1] = b1v.y;
        }
      }
      __syncthreads();
#pragma unroll 2
      for (int k0 = 0; k0 < HP; k0 += 4) {
        const double aphi = A0[16 * wv + l16][k0 + q];
        const double az = A1[16 * wv + l16][k0 + q];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const double bz = B1[16 * j + l16][k0 + q];
          const double bphi = B0[16 * j + l16][k0 + q];
          acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(aphi, bz, acc[j], 0, 0, 0);
          acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(az, bphi, acc[j], 0, 0, 0);
        }
      }
      __syncthreads();
    }
  }
  double *out = vpart + ((int64_t)split * npairs + p) * (FB * FB);
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int mu = 16 * wv + 4 * r + q, nu = 16 * j + l16;
      out[mu * FB + nu] += acc[j][r];
    }
}

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) vreduce_kernel(int npairs, int nsplit, int nbf,
                                                      const int *__restrict__ pair_I,
                                                      const int *__restrict__ pair_J,
                                                      const int *__restrict__ fperm,
                                                      const double *__restrict__ vpart,
                                                      double *__restrict__ V) {
  const int p = blockIdx.x;
  const int I = pair_I[p], J = pair_J[p];
  for (int e = threadIdx.x; e < FB * FB; e += 256) {
    const int mu = e / FB, nu = e % FB;
    if (I == J && mu > nu) continue;
    const int i = I * FB + mu, j = J * FB + nu;
    if (i >= nbf || j >= nbf) continue;
    double v = 0.0;
    for (int s = 0; s < nsplit; ++s) v += vpart[((int64_t)s * npairs + p) * (FB * FB) + e];
    const int oi = fperm[i], oj = fperm[j];
    V[(int64_t)oi * nbf + oj] = v;
    V[(int64_t)oj * nbf + oi] = v;
  }
}

__global__ void __launch_bounds__(256) en_reduce_kernel(int64_t nb, const double *__restrict__ en,
                                                        double *__restrict__ out) {
  __shared__ double se[256], sn[256];
  double e = 0.0, n = 0.0;
  for (int64_t b = threadIdx.x; b < nb; b += 256) {
    e += en[2 * b];
    n += en[2 * b + 1];
  }
  se[threadIdx.x] = e;
  sn[threadIdx.x] = n;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      se[threadIdx.x] += se[threadIdx.x + s];
      sn[threadIdx.x] += sn[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = se[0];
    out[1] = sn[0];
  }
}

__global__ void __launch_bounds__(256) gather_p_f64(int nbf, int nbf_pad,
                                                    const int *__restrict__ fperm,
                                                    const double *__restrict__ P,
                                                    double *__restrict__ Pd) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nbf_pad * nbf_pad) return;
  const int i = (int)(e / nbf_pad), j = (int)(e % nbf_pad);
  Pd[e] = (i < nbf && j < nbf) ? P[(int64_t)fperm[i] * nbf + fperm[j]] : 0.0;
}

// Z = (w/2) Phi for the overlap build (S = Phi W Phi^T through vgemm).
__global__ void __launch_bounds__(256) zoverlap_f64(int64_t s0, const int *__restrict__ slot_b,
                                                    const double *__restrict__ pw, int narr,
                                                    double *__restrict__ tiles) {
  const int64_t s = s0 + blockIdx.x;
  const int b = slot_b[s];
  double *tb = tiles + (int64_t)blockIdx.x * narr * TILE;
  for (int e = threadIdx.x; e < TILE; e += 256)
    tb[(narr - 1) * TILE + e] = 0.5 * pw[(int64_t)b * PT + e % PT] * tb[e];
}

// ---------------------------------------------------------------------------
void vk_gather_p(hipStream_t s, int nbf, int nbf_pad, const int *fperm, const double *P,
                 double *Pd) {
  const int64_t n = (int64_t)nbf_pad * nbf_pad;
  hipLaunchKernelGGL(gather_p_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nbf,
                     nbf_pad, fperm, P, Pd);
  NKX_HIP_CHECK(hipGetLastError());
}

void vk_colloc(hipStream_t s, int64_t s0, int64_t s1, const Plan &pl, int narr, int gga,
               double *tiles) {
  if (s1 <= s0) return;
  hipLaunchKernelGGL(colloc_tiles_f64, dim3((unsigned)(s1 - s0)), dim3(256), 0, s, s0,
                     pl.slot_b.p, pl.b_fbl.p, pl.funcs.p, pl.terms.p, pl.px.p, pl.py.p, pl.pz.p,
                     narr, gga, tiles);
  NKX_HIP_CHECK(hipGetLastError());
}

void vk_density(hipStream_t s, int64_t b0, int64_t b1, int64_t s0, const Plan &pl,
                const double *Pd, int narr, int gga, double *tiles, const XcList &xl,
                double *coef, double *blockEN) {
  if (b1 <= b0) return;
  hipLaunchKernelGGL(density_fused_f64, dim3((unsigned)(b1 - b0)), dim3(256), 0, s, pl.nbf_pad,
                     Pd, pl.b_off.p, pl.b_fbl.p, s0, b0, narr, gga, tiles, pl.pw.p, xl, coef,
                     blockEN);
  NKX_HIP_CHECK(hipGetLastError());
}

void vk_zoverlap(hipStream_t s, int64_t s0, int64_t s1, const Plan &pl, int narr,
                 double *tiles) {
  if (s1 <= s0) return;
  hipLaunchKernelGGL(zoverlap_f64, dim3((unsigned)(s1 - s0)), dim3(256), 0, s, s0, pl.slot_b.p,
                     pl.pw.p, narr, tiles);
  NKX_HIP_CHECK(hipGetLastError());
}

void vk_vgemm(hipStream_t s, int64_t b0, int64_t b1, int64_t s0, const Plan &pl, int nsplit,
              int narr, const double *tiles, double *vpart) {
  if (b1 <= b0 || pl.npairs == 0) return;
  hipLaunchKernelGGL(vgemm_f64, dim3((unsigned)(pl.npairs * nsplit)), dim3(256), 0, s,
                     pl.npairs, nsplit, pl.pair_I.p, pl.pair_J.p, pl.tile_slot.p, pl.nb, b0, b1,
                     s0, narr, tiles, vpart);
  NKX_HIP_CHECK(hipGetLastError());
}

void vk_vreduce(hipStream_t s, const Plan &pl, int nsplit, const double *vpart, double *V) {
  if (pl.npairs == 0) return;
  hipLaunchKernelGGL(vreduce_kernel, dim3(pl.npairs), dim3(256), 0, s, pl.npairs, nsplit, pl.nbf,
                     pl.pair_I.p, pl.pair_J.p, pl.fperm_d.p, vpart, V);
  NKX_HIP_CHECK(hipGetLastError());
}

void vk_en_reduce(hipStream_t s, int64_t nb, const double *en, double *out) {
  hipLaunchKernelGGL(en_reduce_kernel, dim3(1), dim3(256), 0, s, nb, en, out);
  NKX_HIP_CHECK(hipGetLastError());
}

}  // namespace nkx
