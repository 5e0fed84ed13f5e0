// Nuclear gradient of E_xc at fixed P, mixed form: the two contractions of the
// basis-centre term, X = P Phi and W = P Y, run on int8 slices with exact
// integer MFMA accumulation (the slice form of the V_xc build); the
// functional, the STO derivatives of the epilogue and every sum are FP64.
// The terms and their derivation are those of the FP64 form
// (kernels_grad.hip):
//   c_mu(g) = -2 [ d phi_mu(g) W_mu(g) + (H_mu(g) b_g) X_mu(g) ],
//   Y_nu(g) = w v_rho phi_nu + b_g . grad phi_nu = 2 a_g phi_nu + b_g . grad phi_nu,
// with (a, b) the per-point coefficients of the density stage.  The
// point-motion and Becke weight terms reuse the FP64 kernels unchanged.
//
// Per live slot the store holds the V build's sections; the gradient uses Pm
// (Phi, rows = points, k = functions) and writes Y in the same orientation
// into the Z section, so both are B operands of P[mu][nu] with per-point
// exponents.
#include "nkx_internal.hpp"
#include "slices_i8.hpp"
#include "sto_eval.hpp"
#include "vxc_kernels.hpp"

namespace nkx {

namespace {
constexpr int GC = FB * 3;  // per-slot partial: 64 functions x 3 components
}  // namespace

// Y slices of one live slot, written at yout + (slot - s0) ystride (the
// slot's Z section, or a separate array for the second spin); one workgroup
// per slot.  Thread t owns fragment item t: point 16 (t / 64) + t % 16,
// functions 16 ((t % 64) / 16) .. +15.
// Phi comes from the Pm slices (per-point exponent), grad Phi from the FP32
// planes; the row maximum over the four threads of a point by shuffles.
__global__ void __launch_bounds__(256) y_slots_i8(int64_t s0, const int *__restrict__ slot_b,
                                                  int gga, const double *__restrict__ coef,
                                                  signed char *__restrict__ store,
                                                  signed char *__restrict__ yout,
                                                  int64_t ystride) {
  const int64_t sl = s0 + blockIdx.x;
  const int64_t b = slot_b[sl];
  signed char *tb = store + (int64_t)blockIdx.x * slot_bytes_i8(gga);
  const int t = threadIdx.x, l = t & 63;
  const int g = 16 * (t >> 6) + (l & 15), k0 = 16 * (l >> 4);
  const int64_t gp = b * PT + g;
  const double a2 = 2.0 * coef[4 * gp], bx = coef[4 * gp + 1], by = coef[4 * gp + 2],
               bz = coef[4 * gp + 3];
  v4i fp[4];
#pragma unroll
  for (int p = 0; p < 4; ++p) fp[p] = ld16(tb + OFF_PM + p * PLANE + t * 16);
  const int eP = tb[OFF_PM + 4 * PLANE + g];
  // every load is issued before the first use (a branch between them would
  // make each wait for the one before)
  const float *G = reinterpret_cast<const float *>(tb + OFF_G);
  float gv[3][16];
  if (gga) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int e = 0; e < 16; ++e) gv[c][e] = G[c * TILE + (k0 + e) * PT + g];
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int e = 0; e < 16; ++e) gv[c][e] = 0.0f;
  }
  const signed char *c0 = reinterpret_cast<const signed char *>(&fp[0]);
  const signed char *c1 = reinterpret_cast<const signed char *>(&fp[1]);
  const signed char *c2 = reinterpret_cast<const signed char *>(&fp[2]);
  const signed char *c3 = reinterpret_cast<const signed char *>(&fp[3]);
  // 2^(eP - 6) scales the slices exactly; 0 for a point stored as zero
  const double sP = pow2_or_zero(eP == -127 ? -127 : eP - 6);
  double y[16];
  double m = 0.0;
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const double v = (double)c0[e] + ldexp((double)c1[e], -7) + ldexp((double)c2[e], -14) +
                     ldexp((double)c3[e], -21);
    const double ph = v * sP;
    const double yy = gga ? z_value(ph, (double)gv[0][e], (double)gv[1][e], (double)gv[2][e],
                                    a2, bx, by, bz)
                          : a2 * ph;
    y[e] = yy;
    m = fmax(m, fabs(yy));
  }
  m = fmax(m, __shfl_xor(m, 16));
  m = fmax(m, __shfl_xor(m, 32));
  const int ey = row_exp(m);
  __attribute__((aligned(16))) signed char out[4][16];
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    signed char sc[4];
    slice4(y[e], ey, sc);
#pragma unroll
    for (int p = 0; p < 4; ++p) out[p][e] = sc[p];
  }
  signed char *ys = yout + (int64_t)blockIdx.x * ystride;
#pragma unroll
  for (int p = 0; p < 4; ++p)
    *reinterpret_cast<v4i *>(ys + p * PLANE + t * 16) = *reinterpret_cast<const v4i *>(out[p]);
  if (l < 16) ys[4 * PLANE + g] = (signed char)ey;
}

// One workgroup (four waves) per batch.  For each live block I of the batch,
// X and W accumulate over the live blocks K: wave w takes rows 16 w .. +15 of
// I (its P fragment of pair (I, K) in registers), every point of the batch
// (four 16-point fragments of the K slot's Pm and Y sections), the ten slice
// products per tile and an FP64 combine per K.  The steps (I, K) run in that
// order; the K slot's Pm and Y sections are copied into LDS once per step for
// the four waves (two buffers: the next step's copies are issued while the
// current step computes; a second register set for the P fragments would
// cost the occupancy).  The epilogue, after the last K of
// an I, is the FP64 form's: STO gradient (and Hessian for GGA) in FP64 per
// (mu, g), per-function sums over the batch's points into cpart, per-point
// sums over functions into tpt (with the sign of the point-motion term).
__global__ void __launch_bounds__(256) grad_basis_i8(
    MolTab mt, const signed char *__restrict__ Ps, const int *__restrict__ b_off,
    const int *__restrict__ b_fbl, int64_t s0, int64_t b0, int gga,
    const signed char *__restrict__ store, const double *__restrict__ coef,
    const FuncRec *__restrict__ F, const HarmTerm *__restrict__ T,
    const double *__restrict__ px, const double *__restrict__ py,
    const double *__restrict__ pz, const signed char *__restrict__ ybase, int64_t ystride,
    int accum, double *__restrict__ cpart, double *__restrict__ tpt) {
  const int b = (int)(b0 + blockIdx.x);
  const int lo = b_off[b], hi = b_off[b + 1];
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int l16 = lane & 15, q = lane >> 4;
  const int64_t sb_bytes = slot_bytes_i8(gga);

  // Pm and Y sections of a step's K slot, two buffers, declared apart so that
  // a copy into one is known not to touch the other
  __shared__ __attribute__((aligned(16))) signed char bufA[2][SEC];
  __shared__ __attribute__((aligned(16))) signed char bufB[2][SEC];
  __shared__ double cfs[4][PT];
  __shared__ double red[4][3][PT];
  __shared__ double tsum[3][PT];
  __shared__ int lbl[MAXNFB];  // local block index of the batch's live blocks
  for (int e = threadIdx.x; e < 4 * PT; e += 256) {
    const int c = e / PT, g = e % PT;
    cfs[c][g] = coef[4 * ((int64_t)b * PT + g) + c];
  }
  for (int e = threadIdx.x; e < 3 * PT; e += 256) (&tsum[0][0])[e] = 0.0;
  const int n = hi - lo;
  for (int k = threadIdx.x; k < n; k += 256) lbl[k] = mt.blk_local[b_fbl[lo + k]];
  __syncthreads();
  if (n > 0) {
    // the batch's blocks are blocks of one molecule
    const int mol = mt.blk_mol[b_fbl[lo]];
    const signed char *pbase = Ps + mt.Ps_off[mol];
    const int nfbm = mt.nfb[mol];
    const int nsteps = n * n;
    // the K slot's sections of a step -> LDS buffer dst (direct copies)
    auto issue = [&](int st, signed char (*dst)[SEC]) {
      const int64_t sk = lo + st % n - s0;
      dma_section<4>(store + sk * sb_bytes + OFF_PM, &dst[0][0], wv, lane);
      dma_section<4>(ybase + sk * ystride, &dst[1][0], (wv + 2) & 3, lane);
    };
    double X[4][4], W[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) X[j][r] = W[j][r] = 0.0;
    // Epilogue of block I = b_fbl[mi]: c_mu(g) for the 16 (mu, g) pairs of
    // this lane; result register r of lane l holds row 4 (l / 16) + r, point
    // l % 16 of a tile.
    auto epilogue = [&](int mi) {
      const int fm = b_fbl[mi];
      double cm[4][3], tp[4][3];
#pragma unroll
      for (int r = 0; r < 4; ++r)
        for (int k = 0; k < 3; ++k) cm[r][k] = 0.0;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 3; ++k) tp[j][k] = 0.0;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int mu = 16 * wv + 4 * q + r;
        const FuncRec rec = F[fm * FB + mu];
        if (rec.norm == 0.0) continue;  // padding function
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int g = 16 * j + l16;
          const int64_t gp = (int64_t)b * PT + g;
          double c[3];
          if (gga) {
            double v, gr[3], h[6];
            sto_value_grad_hess(rec, T, px[gp], py[gp], pz[gp], v, gr, h);
            const double bx = cfs[1][g], by = cfs[2][g], bz = cfs[3][g];
            const double x = X[j][r], w = W[j][r];
            c[0] = -2.0 * (gr[0] * w + (h[0] * bx + h[1] * by + h[2] * bz) * x);
            c[1] = -2.0 * (gr[1] * w + (h[1] * bx + h[3] * by + h[4] * bz) * x);
            c[2] = -2.0 * (gr[2] * w + (h[2] * bx + h[4] * by + h[5] * bz) * x);
          } else {
            double v, gx, gy, gz;
            sto_value_grad<double>(rec, T, px[gp], py[gp], pz[gp], v, gx, gy, gz);
            const double w = W[j][r];
            c[0] = -2.0 * gx * w;
            c[1] = -2.0 * gy * w;
            c[2] = -2.0 * gz * w;
          }
          for (int k = 0; k < 3; ++k) {
            cm[r][k] += c[k];
            tp[j][k] += c[k];
          }
        }
      }
      // per function: sum over the batch's points (j in registers, l16 lanes)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        for (int k = 0; k < 3; ++k) {
          double v = cm[r][k];
          v += __shfl_xor(v, 1);
          v += __shfl_xor(v, 2);
          v += __shfl_xor(v, 4);
          v += __shfl_xor(v, 8);
          if (l16 == 0) cpart[(mi - s0) * GC + (16 * wv + 4 * q + r) * 3 + k] = v;
        }
      // per point: sum over this block's functions (q lanes, then waves)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 3; ++k) {
          double v = tp[j][k];
          v += __shfl_xor(v, 16);
          v += __shfl_xor(v, 32);
          if (q == 0) red[wv][k][16 * j + l16] = v;
        }
      __syncthreads();
      if (threadIdx.x < 3 * PT) {
        const int k = threadIdx.x / PT, g = threadIdx.x % PT;
        tsum[k][g] += red[0][k][g] + red[1][k][g] + red[2][k][g] + red[3][k][g];
      }
      __syncthreads();
    };
    // one step (I, K): its P fragment, wait for its sections, issue the next
    // step's copies, then the tile products; after the last K of an I, its
    // epilogue
    auto step = [&](int st, signed char (*cur)[SEC], signed char (*nxt)[SEC]) {
      const int ii = st / n, kk = st - ii * n;
      const signed char *pa =
          pbase + ((int64_t)__builtin_amdgcn_readfirstlane(lbl[ii]) * nfbm +
                   __builtin_amdgcn_readfirstlane(lbl[kk])) * SEC;
      v4i A[4];
#pragma unroll
      for (int p = 0; p < 4; ++p) A[p] = ld16(pa + p * PLANE + (wv * 64 + lane) * 16);
      const int eA = *reinterpret_cast<const int *>(pa + 4 * PLANE + 16 * wv + 4 * q);
      wait_vm0();
      __syncthreads();  // buffer cur complete, nxt free
      if (st + 1 < nsteps) issue(st + 1, nxt);
      double sA[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) sA[r] = pow2_or_zero((int)(signed char)(eA >> (8 * r)));
#pragma unroll
      for (int j = 0; j < 4; ++j) {
#pragma unroll
        for (int op = 0; op < 2; ++op) {
          const signed char *sec = &cur[op][0];
          v4i B[4];
#pragma unroll
          for (int p = 0; p < 4; ++p) B[p] = ld16(sec + p * PLANE + (j * 64 + lane) * 16);
          const int eB = sec[4 * PLANE + 16 * j + l16];
          const double sB = eB == -127 ? 0.0 : pow2(eB - 19);
          int hi4[4], lo4[4];
          tile_product(A, B, hi4, lo4);
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const double v = fma((double)lo4[r], 0x1p-14, (double)hi4[r]) * (sA[r] * sB);
            if (op)
              W[j][r] += v;
            else
              X[j][r] += v;
          }
        }
      }
      if (kk == n - 1) {  // uniform: the last K of block I
        epilogue(lo + ii);
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) X[j][r] = W[j][r] = 0.0;
      }
    };
    issue(0, bufA);
    for (int st = 0; st < nsteps; st += 2) {
      step(st, bufA, bufB);
      if (st + 1 < nsteps) step(st + 1, bufB, bufA);
    }
  }
  if (threadIdx.x < 3 * PT) {
    const int k = threadIdx.x / PT, g = threadIdx.x % PT;
    const int64_t e = 3 * ((int64_t)b * PT + g) + k;
    tpt[e] = (accum ? tpt[e] : 0.0) - tsum[k][g];
  }
}

// ---------------------------------------------------------------------------
void vkgm_y_slots(hipStream_t s, int64_t s0, int64_t s1, const Plan &pl, int gga,
                  const double *coef, signed char *store, signed char *ystore_b) {
  if (s1 <= s0) return;
  signed char *yout = ystore_b ? ystore_b : store + OFF_Z;
  const int64_t ystride = ystore_b ? (int64_t)SEC : slot_bytes_i8(gga);
  hipLaunchKernelGGL(y_slots_i8, dim3((unsigned)(s1 - s0)), dim3(256), 0, s, s0, pl.slot_b.p,
                     gga, coef, store, yout, ystride);
  NKX_HIP_CHECK(hipGetLastError());
}

void vkgm_basis(hipStream_t s, int64_t b0, int64_t b1, int64_t s0, const Plan &pl,
                const signed char *Ps, int gga, const signed char *store, const double *coef,
                const signed char *ystore_b, bool accum, double *cpart, double *tpt) {
  if (b1 <= b0) return;
  const signed char *ybase = ystore_b ? ystore_b : store + OFF_Z;
  const int64_t ystride = ystore_b ? (int64_t)SEC : slot_bytes_i8(gga);
  hipLaunchKernelGGL(grad_basis_i8, dim3((unsigned)(b1 - b0)), dim3(256), 0, s, pl.tab(), Ps,
                     pl.b_off.p, pl.b_fbl.p, s0, b0, gga, store, coef, pl.funcs.p, pl.terms.p,
                     pl.px.p, pl.py.p, pl.pz.p, ybase, ystride, accum ? 1 : 0, cpart, tpt);
  NKX_HIP_CHECK(hipGetLastError());
}

}  // namespace nkx
