// Batched V_xc build on MI355X (gfx950), FP64 form.
//
// Data layout in HBM (one chunk of point batches at a time):
//   tile store   [slot][array][64 functions][PT points], array = Phi, dPhi/dx,
//                dPhi/dy, dPhi/dz, Z (GGA) or Phi, Z (LDA); spin-polarised
//                builds carry Z_alpha, Z_beta in place of Z.  A slot is one
//                (function block, point batch) pair that survived screening.
//   Pd           permuted, zero-padded density matrix (nbf_pad x nbf_pad).
//   vpart        split-K partial V tiles [split][pair][64][64].
//
// Kernels per chunk:
//   colloc_tiles     Phi and grad Phi for every live slot (HBM-write bound).
//   density_fused    one workgroup per 4 batches: X = P Phi on MFMA over the
//                    batch's live function blocks, rho / grad rho epilogue in
//                    FP64, the functional, then Z = w (v_rho/2 Phi + 2 v_sigma
//                    grad rho . grad Phi) written back to the tile store.
//   vgemm            one workgroup per (function-block pair I <= J, split):
//                    acc += Phi_I Z_J^T + Z_I Phi_J^T over the batches where
//                    both blocks are live (MFMA-bound).
// and once per build: vreduce (split sum + un-permute + mirror) and the FP64
// energy / electron-count reduction.
#include <algorithm>

#include "nkx_internal.hpp"
#include "sto_eval.hpp"
#include "vxc_kernels.hpp"
#include "xc_functionals.hpp"
#include "xc_polarized.hpp"

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
// One workgroup per group of GBT consecutive (Morton-adjacent) batches: every
// 64 x 64 block of P is loaded into LDS once and applied to all batches of
// the group in which both of its blocks are live, so P is streamed from the
// cache hierarchy once per GBT batches (1, 2 or 4, chosen at launch).  For
// each batch the K blocks are
// visited in ascending order, as with one batch per workgroup, so every sum
// is formed in the same order.
//
// NSPIN = 1: P is the total density matrix.  NSPIN = 2: Pd holds P_alpha then
// P_beta (each nbf_pad^2); the functional is evaluated spin-polarised and the
// tile store carries Z_alpha, Z_beta as its last two arrays.  blockEN holds
// (E, N_alpha, N_beta) per batch (N_beta = 0 for NSPIN = 1, N_alpha = N).
constexpr int MAXNFB = 1024;   // function blocks the union table holds

template <int NSPIN, int GBT>
__global__ void __launch_bounds__(256) density_fused_f64(
    MolTab mt, int64_t pspin, int nfb, const double *__restrict__ Pd, const int *__restrict__ b_off,
    const int *__restrict__ b_fbl, int64_t s0, int64_t b0, int64_t b1, int narr, int gga,
    double *__restrict__ tiles, const double *__restrict__ pw, XcList xl,
    double *__restrict__ coef, double *__restrict__ blockEN, double *__restrict__ epoint,
    double *__restrict__ rhopt) {
  const int64_t bg = b0 + (int64_t)blockIdx.x * GBT;
  const int nsub = (int)(b1 - bg < GBT ? b1 - bg : GBT);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int l16 = lane & 15, q = lane >> 4;

  __shared__ double Pk[FB][FB];   // Pk[k][i] = P(mu = i, nu = k) of the block pair
  __shared__ double Ph[FB][PT];   // Phi tile of the K block
  __shared__ double red[4][4][PT];
  __shared__ double colsum[NSPIN][GBT][4][PT];  // rho, d/dx, d/dy, d/dz column sums
  __shared__ double cf[NSPIN][GBT][4][PT];      // Z coefficients
  __shared__ int tab[GBT][MAXNFB];              // slot of (sub-batch, block) or -1
  __shared__ int ublk[MAXNFB];                 // union of live blocks, ascending
  __shared__ int nu_s;

  for (int e = threadIdx.x; e < GBT * nfb; e += 256) tab[e / nfb][e % nfb] = -1;
  for (int e = threadIdx.x; e < NSPIN * GBT * 4 * PT; e += 256) (&colsum[0][0][0][0])[e] = 0.0;
  __syncthreads();
  for (int sb = 0; sb < nsub; ++sb) {
    const int lo = b_off[bg + sb], hi = b_off[bg + sb + 1];
    for (int k = lo + (int)threadIdx.x; k < hi; k += 256) tab[sb][b_fbl[k]] = k;
  }
  __syncthreads();
  if (wv == 0) {  // ordered compaction of the union
    int cnt = 0;
    for (int base = 0; base < nfb; base += 64) {
      const int f = base + lane;
      bool live = false;
      if (f < nfb)
        for (int sb = 0; sb < GBT; ++sb) live = live || tab[sb][f] >= 0;
      const unsigned long long m = __ballot(live);
      if (live) ublk[cnt + __popcll(m & ((1ull << lane) - 1ull))] = f;
      cnt += __popcll(m);
    }
    if (lane == 0) nu_s = cnt;
  }
  __syncthreads();
  const int nu = nu_s;

  const int nq = gga ? 4 : 1;
  for (int sp = 0; sp < NSPIN; ++sp) {
    const double *Psp = Pd + (int64_t)sp * pspin;
    for (int mi = 0; mi < nu; ++mi) {
      const int fm = ublk[mi];
      const int mol = mt.blk_mol[fm], ld = mt.ld[mol];
      const double *Pmol = Psp + mt.P_off[mol] + mt.blk_local[fm] * FB;
      v4d acc[GBT][4];
#pragma unroll
      for (int sb = 0; sb < GBT; ++sb)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[sb][j] = v4d{0.0, 0.0, 0.0, 0.0};
      for (int ki = 0; ki < nu; ++ki) {
        const int fk = ublk[ki];
        bool any = false;
        for (int sb = 0; sb < nsub; ++sb) any = any || (tab[sb][fk] >= 0 && tab[sb][fm] >= 0);
        if (!any) continue;
        {
          // rows nu of block fk, columns mu of block fm (P symmetric); both
          // blocks belong to molecule mol
          const double *psrc = Pmol + (int64_t)(mt.blk_local[fk] * FB) * ld;
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int e = 2 * (u * 256 + (int)threadIdx.x);
            const int r = e >> 6, c = e & 63;
            *reinterpret_cast<double2 *>(&Pk[r][c]) =
                *reinterpret_cast<const double2 *>(psrc + (int64_t)r * ld + c);
          }
        }
#pragma unroll
        for (int sb = 0; sb < GBT; ++sb) {
          if (sb >= nsub) break;
          const int slk = tab[sb][fk];
          if (slk < 0 || tab[sb][fm] < 0) continue;
          const double *tsrc = tiles + (slk - s0) * narr * TILE;
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int e = 2 * (u * 256 + (int)threadIdx.x);
            const int r = e >> 6, c = e & 63;
            *reinterpret_cast<double2 *>(&Ph[r][c]) = *reinterpret_cast<const double2 *>(tsrc + e);
          }
          __syncthreads();
#pragma unroll 4
          for (int k0 = 0; k0 < FB; k0 += 4) {
            const double a = Pk[k0 + q][16 * wv + l16];
#pragma unroll
            for (int j = 0; j < 4; ++j)
              acc[sb][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Ph[k0 + q][16 * j + l16],
                                                               acc[sb][j], 0, 0, 0);
          }
          __syncthreads();
        }
      }
      // Epilogue: column partial sums sum_mu Phi(mu,g) X(mu,g) (and gradients).
#pragma unroll
      for (int sb = 0; sb < GBT; ++sb) {
        if (sb >= nsub) break;
        const int slm = tab[sb][fm];
        if (slm < 0) continue;
        const double *tb = tiles + (slm - s0) * narr * TILE;
        double part[4][4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int g = 16 * j + l16;
          double sacc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int mu = 16 * wv + 4 * r + q;
            const double x = acc[sb][j][r];
            sacc[0] += tb[mu * PT + g] * x;
            if (gga) {
              sacc[1] += tb[TILE + mu * PT + g] * x;
              sacc[2] += tb[2 * TILE + mu * PT + g] * x;
              sacc[3] += tb[3 * TILE + mu * PT + g] * x;
            }
          }
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            double v = sacc[c];
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
          if (c < nq)
            colsum[sp][sb][c][g] += red[0][c][g] + red[1][c][g] + red[2][c][g] + red[3][c][g];
        }
        __syncthreads();
      }
    }
  }

  // Functional and Z coefficients: wave sb handles the 64 points of batch sb.
  if (wv < nsub) {
    const int sb = wv, g = lane;
    const int64_t b = bg + sb;
    const int64_t gp = b * PT + g;
    const double w = pw[gp];
    double e, na, nb = 0.0;
    if (NSPIN == 1) {
      const double rho = colsum[0][sb][0][g];
      const double gxr = 2.0 * colsum[0][sb][1][g], gyr = 2.0 * colsum[0][sb][2][g],
                   gzr = 2.0 * colsum[0][sb][3][g];
      const double sigma = gga ? gxr * gxr + gyr * gyr + gzr * gzr : 0.0;
      const XcPoint o = xc_eval(xl.n, xl.ids, xl.thr, rho, sigma);
      const double a = 0.5 * w * o.vrho;
      const double bs = gga ? 2.0 * w * o.vsigma : 0.0;
      cf[0][sb][0][g] = a;
      cf[0][sb][1][g] = bs * gxr;
      cf[0][sb][2][g] = bs * gyr;
      cf[0][sb][3][g] = bs * gzr;
      e = w * o.e;
      na = w * rho;
      if (rhopt) {  // rho and grad rho of this point
        rhopt[4 * gp] = rho;
        rhopt[4 * gp + 1] = gxr;
        rhopt[4 * gp + 2] = gyr;
        rhopt[4 * gp + 3] = gzr;
      }
    } else {
      const double r2[2] = {colsum[0][sb][0][g], colsum[NSPIN - 1][sb][0][g]};
      double gr[2][3];
      for (int s2 = 0; s2 < 2; ++s2)
        for (int c = 0; c < 3; ++c) gr[s2][c] = 2.0 * colsum[s2 & (NSPIN - 1)][sb][c + 1][g];
      double sg[3] = {0.0, 0.0, 0.0};
      if (gga) {
        sg[0] = gr[0][0] * gr[0][0] + gr[0][1] * gr[0][1] + gr[0][2] * gr[0][2];
        sg[1] = gr[0][0] * gr[1][0] + gr[0][1] * gr[1][1] + gr[0][2] * gr[1][2];
        sg[2] = gr[1][0] * gr[1][0] + gr[1][1] * gr[1][1] + gr[1][2] * gr[1][2];
      }
      double exc, vr[2], vs[3];
      xc_eval_polarized(xl.n, xl.ids, xl.thr, r2, sg, exc, vr, vs);
      for (int s2 = 0; s2 < 2; ++s2) {
        const double vss = gga ? 2.0 * w * vs[2 * s2] : 0.0, vab = gga ? w * vs[1] : 0.0;
        cf[s2 & (NSPIN - 1)][sb][0][g] = 0.5 * w * vr[s2];
        for (int c = 0; c < 3; ++c)
          cf[s2 & (NSPIN - 1)][sb][c + 1][g] = vss * gr[s2][c] + vab * gr[1 - s2][c];
      }
      e = w * (r2[0] + r2[1]) * exc;
      na = w * r2[0];
      nb = w * r2[1];
    }
    for (int s2 = 0; s2 < NSPIN; ++s2)
      for (int c = 0; c < 4; ++c) coef[4 * NSPIN * gp + 4 * s2 + c] = cf[s2][sb][c][g];
    if (epoint) epoint[gp] = e;  // w e_xc of this point
    for (int o2 = 32; o2 > 0; o2 >>= 1) {
      e += __shfl_down(e, o2);
      na += __shfl_down(na, o2);
      nb += __shfl_down(nb, o2);
    }
    if (g == 0) {
      blockEN[3 * b] = e;
      blockEN[3 * b + 1] = na;
      blockEN[3 * b + 2] = nb;
    }
  }
  __syncthreads();

  // Z tiles, Z_s in array narr - NSPIN + s.
  for (int mi = 0; mi < nu; ++mi) {
    const int fm = ublk[mi];
    for (int sb = 0; sb < nsub; ++sb) {
      const int slm = tab[sb][fm];
      if (slm < 0) continue;
      double *tb = tiles + (slm - s0) * narr * TILE;
      for (int sp = 0; sp < NSPIN; ++sp) {
        double *zt = tb + (narr - NSPIN + sp) * TILE;
        for (int e = threadIdx.x; e < TILE; e += 256) {
          const int g = e % PT;
          double z = cf[sp][sb][0][g] * tb[e];
          if (gga)
            z += cf[sp][sb][1][g] * tb[TILE + e] + cf[sp][sb][2][g] * tb[2 * TILE + e] +
                 cf[sp][sb][3][g] * tb[3 * TILE + e];
          zt[e] = z;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// acc(I, J) += Phi_I Z_J^T + Z_I Phi_J^T over batches where I and J are live.
constexpr int HP = 32;      // points per LDS stage
constexpr int LDP = HP + 1; // padded row

// blockIdx.x = (zs * nsplit + split) * npairs + pair; Z is tile array
// zbase + zs, and the partials go to vpart + zs * nsplit * npairs * FB * FB.
__global__ void __launch_bounds__(256) vgemm_f64(
    int npairs, int nsplit, const int *__restrict__ pair_I, const int *__restrict__ pair_J,
    const int *__restrict__ tile_slot, MolTab mt, int64_t b0, int64_t b1, int64_t s0,
    int narr, int zbase, const double *__restrict__ tiles, double *__restrict__ vpart) {
  const int p = blockIdx.x % npairs;
  const int split = (blockIdx.x / npairs) % nsplit;
  const int zs = blockIdx.x / (npairs * nsplit);
  const int zarr = zbase + zs;
  vpart += (int64_t)zs * nsplit * npairs * (FB * FB);
  const int I = pair_I[p], J = pair_J[p];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int l16 = lane & 15, q = lane >> 4;
  __shared__ double A0[FB][LDP], A1[FB][LDP], B0[FB][LDP], B1[FB][LDP];
  v4d acc[4];
  for (int j = 0; j < 4; ++j) acc[j] = v4d{0.0, 0.0, 0.0, 0.0};
  // batches of the pair's molecule inside [b0, b1)
  const int mol = mt.blk_mol[I];
  const int64_t mb0 = mt.b0[mol], mnb = mt.nb[mol];
  const int *tsI = tile_slot + mt.ts_off[mol] + (int64_t)mt.blk_local[I] * mnb - mb0;
  const int *tsJ = tile_slot + mt.ts_off[mol] + (int64_t)mt.blk_local[J] * mnb - mb0;
  const int64_t lo = b0 > mb0 ? b0 : mb0, hi = b1 < mb0 + mnb ? b1 : mb0 + mnb;
  for (int64_t b = lo + split; b < hi; b += nsplit) {
    const int sI = tsI[b], sJ = tsJ[b];
    if (sI < 0 || sJ < 0) continue;
    const double *tI = tiles + (sI - s0) * narr * TILE;
    const double *tJ = tiles + (sJ - s0) * narr * TILE;
    for (int h = 0; h < PT; h += HP) {
      {
        const double *zI = tI + zarr * TILE, *zJ = tJ + zarr * TILE;
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
__global__ void __launch_bounds__(256) vreduce_kernel(int npairs, int nsplit,
                                                      const int *__restrict__ pair_I,
                                                      const int *__restrict__ pair_J,
                                                      const int *__restrict__ fperm, MolTab mt,
                                                      const int *__restrict__ mol_nbf,
                                                      const double *__restrict__ vpart,
                                                      double *const *__restrict__ Vmol) {
  const int p = blockIdx.x;
  const int I = pair_I[p], J = pair_J[p];
  const int mol = mt.blk_mol[I];
  const int nbf = mol_nbf[mol];
  double *V = Vmol[mol];
  const int lI = mt.blk_local[I], lJ = mt.blk_local[J];
  for (int e = threadIdx.x; e < FB * FB; e += 256) {
    const int mu = e / FB, nu = e % FB;
    if (I == J && mu > nu) continue;
    const int i = lI * FB + mu, j = lJ * FB + nu;
    if (i >= nbf || j >= nbf) continue;
    double v = 0.0;
    for (int s = 0; s < nsplit; ++s) v += vpart[((int64_t)s * npairs + p) * (FB * FB) + e];
    const int oi = fperm[I * FB + mu], oj = fperm[J * FB + nu];
    V[(int64_t)oi * nbf + oj] = v;
    V[(int64_t)oj * nbf + oi] = v;
  }
}

// Fixed-order sum of the per-batch (E, N_alpha, N_beta) triples, one
// workgroup per molecule over its batches.
__global__ void __launch_bounds__(256) en_reduce_kernel(MolTab mt, const double *__restrict__ en,
                                                        double *__restrict__ out) {
  __shared__ double sv[3][256];
  const int mol = blockIdx.x;
  const int64_t b0 = mt.b0[mol], b1 = b0 + mt.nb[mol];
  double v[3] = {0.0, 0.0, 0.0};
  for (int64_t b = b0 + threadIdx.x; b < b1; b += 256)
    for (int c = 0; c < 3; ++c) v[c] += en[3 * b + c];
  for (int c = 0; c < 3; ++c) sv[c][threadIdx.x] = v[c];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int c = 0; c < 3; ++c) sv[c][threadIdx.x] += sv[c][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; ++c) out[3 * mol + c] = sv[c][0];
}

// Zero fill (stream-ordered kernel, also valid inside a captured graph).
__global__ void __launch_bounds__(256) fill0_kernel(int64_t n, double *__restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = 0.0;
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

template <int NSPIN, int GBT>
static void launch_density(hipStream_t s, int64_t b0, int64_t b1, int64_t s0, const Plan &pl,
                           const double *Pd, int narr, int gga, double *tiles, const XcList &xl,
                           double *coef, double *blockEN, double *epoint, double *rhopt) {
  const unsigned nwg = (unsigned)((b1 - b0 + GBT - 1) / GBT);
  hipLaunchKernelGGL((density_fused_f64<NSPIN, GBT>), dim3(nwg), dim3(256), 0, s, pl.tab(),
                     pl.P_total, pl.nfb, Pd, pl.b_off.p, pl.b_fbl.p, s0, b0, b1, narr, gga, tiles,
                     pl.pw.p, xl, coef, blockEN, epoint, rhopt);
}

void vk_density(hipStream_t s, int nspin, int group, int64_t b0, int64_t b1, int64_t s0,
                const Plan &pl, const double *Pd, int narr, int gga, double *tiles,
                const XcList &xl, double *coef, double *blockEN, double *epoint,
                double *rhopt) {
  if (b1 <= b0) return;
  NKX_REQUIRE(pl.nfb <= MAXNFB, NKX_ERR_UNSUPPORTED,
              "more than 1024 function blocks (65536 functions)");
  const int g0 = group == 66 ? 2 : (group >= 65 ? 1 : group);  // the int8 form's codes
  const int g = g0 == 1 || g0 == 2 ? g0 : 4;
#define NKX_DENS(NS, G) launch_density<NS, G>(s, b0, b1, s0, pl, Pd, narr, gga, tiles, xl, coef, blockEN, epoint, rhopt)
  if (nspin == 2) {
    if (g == 1) NKX_DENS(2, 1);
    else if (g == 2) NKX_DENS(2, 2);
    else NKX_DENS(2, 4);
  } else {
    if (g == 1) NKX_DENS(1, 1);
    else if (g == 2) NKX_DENS(1, 2);
    else NKX_DENS(1, 4);
  }
#undef NKX_DENS
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
              int narr, int nz, const double *tiles, double *vpart) {
  if (b1 <= b0 || pl.npairs == 0) return;
  hipLaunchKernelGGL(vgemm_f64, dim3((unsigned)(pl.npairs * nsplit * nz)), dim3(256), 0, s,
                     pl.npairs, nsplit, pl.pair_I.p, pl.pair_J.p, pl.tile_slot.p, pl.tab(), b0, b1,
                     s0, narr, narr - nz, tiles, vpart);
  NKX_HIP_CHECK(hipGetLastError());
}

void vk_vreduce(hipStream_t s, const Plan &pl, int nsplit, const double *vpart,
                double *const *Vmol_dev, const int *mol_nbf_dev) {
  if (pl.npairs == 0) return;
  hipLaunchKernelGGL(vreduce_kernel, dim3(pl.npairs), dim3(256), 0, s, pl.npairs, nsplit,
                     pl.pair_I.p, pl.pair_J.p, pl.fperm_d.p, pl.tab(), mol_nbf_dev, vpart,
                     Vmol_dev);
  NKX_HIP_CHECK(hipGetLastError());
}

void vk_en_reduce(hipStream_t s, const Plan &pl, const double *en, double *out) {
  hipLaunchKernelGGL(en_reduce_kernel, dim3(pl.nmol), dim3(256), 0, s, pl.tab(), en, out);
  NKX_HIP_CHECK(hipGetLastError());
}

void vk_fill0(hipStream_t s, double *x, int64_t n) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(fill0_kernel, dim3((unsigned)blocks), dim3(256), 0, s, n, x);
  NKX_HIP_CHECK(hipGetLastError());
}

}  // namespace nkx
