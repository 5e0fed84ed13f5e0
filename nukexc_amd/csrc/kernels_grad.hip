// Nuclear gradient of E_xc at fixed density matrix P (FP64).  The reference
// has no gradient; the terms follow from E = sum_g w_g f(rho_g, sigma_g) with
// basis functions centred on atoms, grid points moving with their owner atom
// and Becke weights w_g = w0_g P_B(r_g) / sum_i P_i(r_g) (B = owner):
//
//  basis-centre term, atom A, function mu on A, per point g:
//    c_mu(g) = -2 [ d phi_mu(g) (P Y)_mu(g) + (H_mu(g) b_g) (P Phi)_mu(g) ]
//    Y_nu(g) = w v_rho phi_nu + b_g . grad phi_nu,   b_g = 2 w v_sigma grad rho
//    (H_mu = Hessian of phi_mu; GGA only)
//  point-motion term: rigid translation of every atom together with the
//    point leaves rho(r_g) unchanged, so the owner receives -sum_mu c_mu(g);
//  weight term: f_g dw_g/dR_A = (w f)_g sum_i ([i == B]/P_B - 1/Z) dP_i/dR_A
//    for A != B, and the owner receives minus their sum (w is invariant under
//    rigid translation of atoms and point).
// Becke derivatives: P_i = prod_j s(mu_ij), mu_ij = (r_i - r_j)/R~_ij,
//   R~ = min(R_ij, 5) (capped) or R_ij, s = (1 - p(p(p(mu))))/2,
//   dmu_ij/dR_i = -u_i/R~ - [R~ = R_ij] mu (R_i - R_j)/R_ij^2,
//   dmu_ij/dR_j = +u_j/R~ + [R~ = R_ij] mu (R_i - R_j)/R_ij^2,
//   u_i = (r - R_i)/r_i; s' = 0 where mu is clamped at +-1.
#include "nkx_internal.hpp"
#include "sto_eval.hpp"
#include "vxc_kernels.hpp"

namespace nkx {

typedef double v4d __attribute__((ext_vector_type(4)));

namespace {
constexpr int FB = 64;
constexpr int PT = 64;
constexpr int TILE = FB * PT;
constexpr int GC = FB * 3;  // per-slot partial: 64 functions x 3 components
}  // namespace

// ---------------------------------------------------------------------------
// One workgroup per batch.  Tile store arrays: Phi, dPhi/dx, dPhi/dy, dPhi/dz
// (narr >= 4).  coef per point (stride cstride doubles): (w v_rho / 2, b_x,
// b_y, b_z) of the spin whose P is Pd; accum adds the point sums to tpt (the
// second spin of an open shell) instead of writing them.
__global__ void __launch_bounds__(256) grad_basis_f64(
    int nbf_pad, const double *__restrict__ Pd, const int *__restrict__ b_off,
    const int *__restrict__ b_fbl, int64_t s0, int64_t b0, int narr, int gga,
    const double *__restrict__ tiles, const double *__restrict__ coef,
    const FuncRec *__restrict__ F, const HarmTerm *__restrict__ T,
    const double *__restrict__ px, const double *__restrict__ py,
    const double *__restrict__ pz, int cstride, int accum, double *__restrict__ cpart,
    double *__restrict__ tpt) {
  const int b = (int)(b0 + blockIdx.x);
  const int lo = b_off[b], hi = b_off[b + 1];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int l16 = lane & 15, q = lane >> 4;

  __shared__ double Pk[FB][FB];   // Pk[k][i] = P(mu = i, nu = k)
  __shared__ double Ph[FB][PT];   // Phi of the K block
  __shared__ double Yh[FB][PT];   // Y of the K block
  __shared__ double cfs[4][PT];
  __shared__ double red[4][3][PT];
  __shared__ double tsum[3][PT];

  for (int e = threadIdx.x; e < 4 * PT; e += 256) {
    const int c = e / PT, g = e % PT;
    cfs[c][g] = coef[cstride * ((int64_t)b * PT + g) + c];
  }
  for (int e = threadIdx.x; e < 3 * PT; e += 256) (&tsum[0][0])[e] = 0.0;
  __syncthreads();

  for (int mi = lo; mi < hi; ++mi) {
    const int fm = b_fbl[mi];
    v4d accX[4], accW[4];
    for (int j = 0; j < 4; ++j) accX[j] = accW[j] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int ki = lo; ki < hi; ++ki) {
      const int fk = b_fbl[ki];
      const double *psrc = Pd + (int64_t)(fk * FB) * nbf_pad + fm * FB;
      const double *tsrc = tiles + (ki - s0) * narr * TILE;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = 2 * (u * 256 + (int)threadIdx.x);
        const int r = e >> 6, c = e & 63;
        *reinterpret_cast<double2 *>(&Pk[r][c]) =
            *reinterpret_cast<const double2 *>(psrc + (int64_t)r * nbf_pad + c);
        const double2 ph = *reinterpret_cast<const double2 *>(tsrc + e);
        *reinterpret_cast<double2 *>(&Ph[r][c]) = ph;
        double2 y;
        y.x = 2.0 * cfs[0][c] * ph.x;
        y.y = 2.0 * cfs[0][c + 1] * ph.y;
        if (gga) {
#pragma unroll
          for (int d = 1; d <= 3; ++d) {
            const double2 gd = *reinterpret_cast<const double2 *>(tsrc + d * TILE + e);
            y.x += cfs[d][c] * gd.x;
            y.y += cfs[d][c + 1] * gd.y;
          }
        }
        *reinterpret_cast<double2 *>(&Yh[r][c]) = y;
      }
      __syncthreads();
#pragma unroll 4
      for (int k0 = 0; k0 < FB; k0 += 4) {
        const double a = Pk[k0 + q][16 * wv + l16];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          accX[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Ph[k0 + q][16 * j + l16], accX[j], 0,
                                                          0, 0);
          accW[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Yh[k0 + q][16 * j + l16], accW[j], 0,
                                                          0, 0);
        }
      }
      __syncthreads();
    }
    // Epilogue: c_mu(g) for the 16 (mu, g) pairs of this lane.
    const double *tb = tiles + (mi - s0) * narr * TILE;
    double cm[4][3], tp[4][3];
#pragma unroll
    for (int r = 0; r < 4; ++r)
      for (int k = 0; k < 3; ++k) cm[r][k] = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 3; ++k) tp[j][k] = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int mu = 16 * wv + 4 * r + q;
      const FuncRec rec = F[fm * FB + mu];
      if (rec.norm == 0.0) continue;  // padding function
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int g = 16 * j + l16;
        const double X = accX[j][r], W = accW[j][r];
        double c[3];
        for (int k = 0; k < 3; ++k) c[k] = -2.0 * tb[(1 + k) * TILE + mu * PT + g] * W;
        if (gga) {
          const int64_t gp = (int64_t)b * PT + g;
          double v, gr[3], h[6];
          sto_value_grad_hess(rec, T, px[gp], py[gp], pz[gp], v, gr, h);
          const double bx = cfs[1][g], by = cfs[2][g], bz = cfs[3][g];
          c[0] -= 2.0 * (h[0] * bx + h[1] * by + h[2] * bz) * X;
          c[1] -= 2.0 * (h[1] * bx + h[3] * by + h[4] * bz) * X;
          c[2] -= 2.0 * (h[2] * bx + h[4] * by + h[5] * bz) * X;
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
        if (l16 == 0) cpart[(mi - s0) * GC + (16 * wv + 4 * r + q) * 3 + k] = v;
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
  }
  if (threadIdx.x < 3 * PT) {
    const int k = threadIdx.x / PT, g = threadIdx.x % PT;
    const int64_t e = 3 * ((int64_t)b * PT + g) + k;
    tpt[e] = (accum ? tpt[e] : 0.0) - tsum[k][g];
  }
}

// gacc[seg][I * 192 + t] += sum over batches of segment seg of cpart of block I
__global__ void __launch_bounds__(256) grad_fold_kernel(int64_t nb, int64_t b0, int64_t b1,
                                                        int64_t s0, int nseg, int nbf_pad,
                                                        const int *__restrict__ tile_slot,
                                                        const double *__restrict__ cpart,
                                                        double *__restrict__ gacc) {
  const int I = blockIdx.x, seg = blockIdx.y;
  const int t = threadIdx.x;
  if (t >= GC) return;
  const int64_t n = b1 - b0;
  const int64_t lo = b0 + n * seg / nseg, hi = b0 + n * (seg + 1) / nseg;
  const int *ts = tile_slot + (int64_t)I * nb;
  double acc = 0.0;
  for (int64_t b = lo; b < hi; ++b) {
    const int sl = ts[b];
    if (sl >= 0) acc += cpart[(sl - s0) * GC + t];
  }
  gacc[(int64_t)seg * nbf_pad * 3 + I * GC + t] += acc;
}

__global__ void __launch_bounds__(256) owners_kernel(int64_t n, const int *__restrict__ pidx,
                                                     const int *__restrict__ owner,
                                                     int *__restrict__ pown) {
  const int64_t sp = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sp >= n) return;
  const int src = pidx[sp];
  pown[sp] = src >= 0 ? owner[src] : -1;
}

__device__ __forceinline__ double bp(double x) { return 1.5 * x - 0.5 * x * x * x; }
__device__ __forceinline__ double bdp(double x) { return 1.5 * (1.0 - x * x); }

__device__ __forceinline__ double gdist(double x, double y, double z, const double *c) {
  const double dx = x - c[0], dy = y - c[1], dz = z - c[2];
  return sqrt(dx * dx + dy * dy + dz * dz);
}

// Per atom pair (i, j): 1/R~_ij and the factor of the mu (R_i - R_j) term of
// dmu_ij/dR, [R~ = R_ij] / R_ij^2 (zero where the cap applies), with
// R~ = min(R_ij, 5) (capped partition) or R_ij.  One thread per pair.
__global__ void __launch_bounds__(256) pair_table_kernel(int natoms, const double *__restrict__ C,
                                                         int cap, double2 *__restrict__ tab) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= natoms * natoms) return;
  const int i = e / natoms, j = e % natoms;
  const double dx = C[3 * i] - C[3 * j], dy = C[3 * i + 1] - C[3 * j + 1],
               dz = C[3 * i + 2] - C[3 * j + 2];
  const double R = sqrt(dx * dx + dy * dy + dz * dz);
  const bool capped = cap && R >= 5.0;
  tab[e] = i == j ? make_double2(0.0, 0.0)
                  : make_double2(1.0 / (capped ? 5.0 : R), capped ? 0.0 : 1.0 / (R * R));
}

// Becke cell function P_i at (x, y, z); same screening and clamping as the
// partition kernel, with mu_ij = (r_i - r_j) / R~_ij taken as a product with
// the tabulated 1/R~_ij (tab = row i of the pair table).
__device__ double becke_cell(int i, int natoms, const double *C, const double2 *tab, double x,
                             double y, double z, double ri, bool cap) {
  double Pi = 1.0;
  for (int j = 0; j < natoms; ++j) {
    if (j == i) continue;
    const double rj = gdist(x, y, z, C + 3 * j);
    if (cap && ri - rj <= -5.0) continue;  // mu <= -1: factor 1
    double mu = (ri - rj) * tab[j].x;
    mu = mu > 1.0 ? 1.0 : (mu < -1.0 ? -1.0 : mu);
    Pi *= 0.5 * (1.0 - bp(bp(bp(mu))));
    if (Pi == 0.0) break;
  }
  return Pi;
}

// DPP move of a double (both halves) from the lane that CTRL selects inside the
// same row of 16 lanes; every lane must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
  const unsigned hi =
      (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ double lane_value(double v, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum over the 64 lanes of a wave; every lane must be active, and every lane
// receives the same value.  Inside each row of 16 lanes by DPP moves (quad
// permutations for lane ^ 1 and lane ^ 2, then the half-row and row mirrors):
// each step adds the same two partial sums in every lane (a + b and b + a are
// the same double), so all lanes of a row hold the same row sum.  The four row
// sums are read out of lanes 0, 16, 32 and 48 and added in a fixed order.
// No LDS instruction is involved (lane shuffles through ds_bpermute would
// put twelve dependent LDS round trips in every sum).
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_move<0xB1>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp_move<0x4E>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_move<0x141>(v);  // row_half_mirror: lane 7 - i of each 8
  v += dpp_move<0x140>(v);  // row_mirror: lane 15 - i of each 16
  return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}

// One thread per sorted point, one wave per workgroup; per-atom sums in LDS
// (3 natoms doubles).  Every loop over atoms runs in step across the wave
// (lanes without a contribution add zero), so each per-atom term is summed
// over the wave (wave_sum: every lane active) and added to LDS by lane 0: all
// lanes of a wave touch the same atoms, and one LDS atomic per lane would
// serialise on them.  With one writer per workgroup (no atomics) the sums, and
// so the gradient, come out the same on every run.  The points of a wave are
// 64 consecutive sorted points (npts_pad and the grid stride are multiples of
// 64), so every lane runs the same trip count.  ptab: the atom-pair table of
// pair_table_kernel.
__global__ void __launch_bounds__(64) grad_weights_kernel(
    int64_t npts_pad, int natoms, const double *__restrict__ C, const double *__restrict__ px,
    const double *__restrict__ py, const double *__restrict__ pz, const int *__restrict__ pown,
    const double *__restrict__ epoint, const double *__restrict__ tpt, int partition,
    const double2 *__restrict__ ptab, double *__restrict__ wpart) {
  extern __shared__ double acc[];
  for (int e = threadIdx.x; e < 3 * natoms; e += blockDim.x) acc[e] = 0.0;
  __syncthreads();
  const bool cap = partition == 1;
  const int lane = threadIdx.x & 63;
  for (int64_t sp = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; sp < npts_pad;
       sp += (int64_t)gridDim.x * blockDim.x) {
    const int B = pown[sp];
    const bool act = B >= 0 && B < natoms;  // not padding
    double S[3] = {0.0, 0.0, 0.0};          // the owner receives -S
    double x = 0.0, y = 0.0, z = 0.0, wf = 0.0, rmin = 1e300, Z = 0.0, PB = 0.0;
    if (act) {
      for (int k = 0; k < 3; ++k) S[k] = -tpt[3 * sp + k];
      wf = epoint[sp];
      x = px[sp];
      y = py[sp];
      z = pz[sp];
    }
    const bool wact = act && partition != 0 && wf != 0.0;
    if (wact) {
      if (cap)
        for (int i = 0; i < natoms; ++i) rmin = fmin(rmin, gdist(x, y, z, C + 3 * i));
      for (int i = 0; i < natoms; ++i) {
        const double ri = gdist(x, y, z, C + 3 * i);
        if (cap && ri - rmin >= 5.0) continue;
        const double Pi = becke_cell(i, natoms, C, ptab + i * natoms, x, y, z, ri, cap);
        Z += Pi;
        if (i == B) PB = Pi;
      }
    }
    const bool gact = wact && PB > 0.0 && Z > 0.0;
    if (__ballot(gact)) {
      for (int i = 0; i < natoms; ++i) {
        bool li = gact;
        double ri = 1.0, Pi = 0.0;
        if (li) {
          ri = gdist(x, y, z, C + 3 * i);
          li = !(cap && ri - rmin >= 5.0);
        }
        if (li) {
          Pi = becke_cell(i, natoms, C, ptab + i * natoms, x, y, z, ri, cap);
          li = Pi != 0.0;
        }
        if (!__ballot(li)) continue;
        const double coefi = li ? wf * ((i == B ? 1.0 / PB : 0.0) - 1.0 / Z) * Pi : 0.0;
        const double *Ci = C + 3 * i;
        const double2 *ti = ptab + i * natoms;
        const double iri = 1.0 / ri;
        const double ui[3] = {(x - Ci[0]) * iri, (y - Ci[1]) * iri, (z - Ci[2]) * iri};
        double gi[3] = {0.0, 0.0, 0.0};
        for (int j = 0; j < natoms; ++j) {
          if (j == i) continue;
          const double *Cj = C + 3 * j;
          double gj[3] = {0.0, 0.0, 0.0};
          bool lj = li;
          double rj = 1.0;
          if (lj) {
            rj = gdist(x, y, z, Cj);
            lj = !(cap && ri - rj <= -5.0);
          }
          if (lj) {
            const double d[3] = {Ci[0] - Cj[0], Ci[1] - Cj[1], Ci[2] - Cj[2]};
            const double2 tij = ti[j];  // 1/R~, [not capped] / R^2
            const double mu = (ri - rj) * tij.x;
            if (mu > -1.0 && mu < 1.0) {  // clamped: s' = 0
              const double p1 = bp(mu), p2 = bp(p1), p3 = bp(p2);
              const double sv = 0.5 * (1.0 - p3);
              const double ds = -0.5 * bdp(p2) * bdp(p1) * bdp(mu);
              const double f = coefi * ds / sv;  // s > 0 because P_i > 0
              const double ct = mu * tij.y;
              const double irj = 1.0 / rj;
              const double uj[3] = {(x - Cj[0]) * irj, (y - Cj[1]) * irj, (z - Cj[2]) * irj};
              for (int k = 0; k < 3; ++k) {
                gi[k] += f * (-ui[k] * tij.x - ct * d[k]);
                if (j != B) gj[k] = f * (uj[k] * tij.x + ct * d[k]);
              }
            } else {
              lj = false;
            }
          }
          if (!__ballot(lj && j != B)) continue;
          for (int k = 0; k < 3; ++k) {
            S[k] += gj[k];
            const double v = wave_sum(gj[k]);
            if (lane == 0) acc[3 * j + k] += v;
          }
        }
        for (int k = 0; k < 3; ++k) {
          const double gv = (li && i != B) ? gi[k] : 0.0;
          S[k] += gv;
          const double v = wave_sum(gv);
          if (lane == 0) acc[3 * i + k] += v;
        }
      }
    }
    // owner terms, one wave sum per distinct owner among the wave's points
    unsigned long long pend = __ballot(act);
    while (pend) {
      const int src = __ffsll((long long)pend) - 1;
      const int Bs = __shfl(B, src);
      const bool mine = act && B == Bs;
      for (int k = 0; k < 3; ++k) {
        const double v = wave_sum(mine ? -S[k] : 0.0);
        if (lane == 0) acc[3 * Bs + k] += v;
      }
      pend &= ~__ballot(mine);
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 3 * natoms; e += blockDim.x)
    wpart[(int64_t)blockIdx.x * 3 * natoms + e] = acc[e];
}

// out[k] = sum over the rows of in[r][k]; one workgroup per column.  Thread t
// sums the rows r = t (mod 256) in ascending order, then a fixed tree in LDS
// adds the 256 partial sums, so the order of the additions is the same on
// every run.
__global__ void __launch_bounds__(256) colsum_kernel(int nrow, int ncol,
                                                     const double *__restrict__ in,
                                                     double *__restrict__ out) {
  __shared__ double part[256];
  const int k = blockIdx.x, t = threadIdx.x;
  double v = 0.0;
  for (int r = t; r < nrow; r += 256) v += in[(int64_t)r * ncol + k];
  part[t] = v;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) part[t] += part[t + h];
    __syncthreads();
  }
  if (t == 0) out[k] = part[0];
}

// ---------------------------------------------------------------------------
void vkg_basis(hipStream_t s, int64_t b0, int64_t b1, int64_t s0, const Plan &pl,
               const double *Pd, int narr, int gga, const double *tiles, const double *coef,
               int cstride, bool accum, double *cpart, double *tpt) {
  if (b1 <= b0) return;
  hipLaunchKernelGGL(grad_basis_f64, dim3((unsigned)(b1 - b0)), dim3(256), 0, s, pl.nbf_pad, Pd,
                     pl.b_off.p, pl.b_fbl.p, s0, b0, narr, gga, tiles, coef, pl.funcs.p,
                     pl.terms.p, pl.px.p, pl.py.p, pl.pz.p, cstride, accum ? 1 : 0, cpart, tpt);
  NKX_HIP_CHECK(hipGetLastError());
}

void vkg_fold(hipStream_t s, int64_t b0, int64_t b1, int64_t s0, const Plan &pl, int nseg,
              const double *cpart, double *gacc) {
  if (b1 <= b0) return;
  hipLaunchKernelGGL(grad_fold_kernel, dim3((unsigned)pl.nfb, (unsigned)nseg), dim3(256), 0, s,
                     pl.nb, b0, b1, s0, nseg, pl.nbf_pad, pl.tile_slot.p, cpart, gacc);
  NKX_HIP_CHECK(hipGetLastError());
}

void vkg_owners(hipStream_t s, const Plan &pl, const int *owner, int *pown) {
  const int64_t n = pl.npts_pad;
  hipLaunchKernelGGL(owners_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n,
                     pl.pidx.p, owner, pown);
  NKX_HIP_CHECK(hipGetLastError());
}

void vkg_weights(hipStream_t s, const Plan &pl, int natoms, const double *centers,
                 const int *pown, const double *epoint, const double *tpt, int partition,
                 int nblk, double *ptab, double *wpart) {
  double2 *tab = reinterpret_cast<double2 *>(ptab);
  const int npair = natoms * natoms;
  hipLaunchKernelGGL(pair_table_kernel, dim3((unsigned)((npair + 255) / 256)), dim3(256), 0, s,
                     natoms, centers, partition == 1 ? 1 : 0, tab);
  NKX_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(grad_weights_kernel, dim3((unsigned)nblk), dim3(64),
                     (size_t)3 * natoms * sizeof(double), s, pl.npts_pad, natoms, centers,
                     pl.px.p, pl.py.p, pl.pz.p, pown, epoint, tpt, partition, tab, wpart);
  NKX_HIP_CHECK(hipGetLastError());
}

void vkg_colsum(hipStream_t s, int nrow, int ncol, const double *in, double *out) {
  if (ncol <= 0) return;
  hipLaunchKernelGGL(colsum_kernel, dim3((unsigned)ncol), dim3(256), 0, s, nrow, ncol, in, out);
  NKX_HIP_CHECK(hipGetLastError());
}

}  // namespace nkx
