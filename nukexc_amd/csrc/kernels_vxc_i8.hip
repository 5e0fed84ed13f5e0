// Batched V_xc build on MI355X (gfx950), mixed form: int8 slices with exact
// integer MFMA accumulation (slice form and tile layout in slices_i8.hpp):
// collocation into slices, P gather, the density contraction forms, the
// per-batch functional, Z slices and the V contraction.
#include <type_traits>

#include "nkx_internal.hpp"
#include "sto_eval.hpp"
#include "vxc_kernels.hpp"
#include "slices_i8.hpp"
#include "xc_functionals.hpp"
#include "xc_polarized.hpp"

namespace nkx {

int64_t mixed_slot_bytes(int gga) { return slot_bytes_i8(gga); }

// ---------------------------------------------------------------------------
// Collocation: one workgroup per live slot; FP64 values in LDS, row maxima
// along functions (F) and along points (Pm), slices in both orders.  Thread
// t evaluates the 16 consecutive functions 16 (t / 64) .. +15 of the block
// at point t % 64 (one function at a time for the whole wave, a point per
// lane).  Functions are sorted by centre, so consecutive functions that
// share the centre share dx, dy, dz, r and 1/r, and those that also share
// zeta, k and the normalisation share the radial factor.
__global__ void __launch_bounds__(256) colloc_tiles_i8(
    int64_t s0, const int *__restrict__ slot_b, const int *__restrict__ b_fbl,
    const FuncRec *__restrict__ F, const HarmTerm *__restrict__ T,
    const double *__restrict__ px, const double *__restrict__ py,
    const double *__restrict__ pz, int gga, signed char *__restrict__ store) {
  const int64_t s = s0 + blockIdx.x;
  const int b = slot_b[s];
  const int fb = b_fbl[s];
  __shared__ double v[FB][65];
  __shared__ int eF[FB], eP[PT];
  const int g = threadIdx.x & (PT - 1);
  // the wave's functions are wave-uniform: their records are scalar loads and
  // the branches on them scalar branches
  const int f0 = 16 * __builtin_amdgcn_readfirstlane(threadIdx.x / PT);
  const FuncRec *fr = F + (int64_t)fb * FB;
  const int64_t gp = (int64_t)b * PT + g;
  const double x = px[gp], y = py[gp], z = pz[gp];
  signed char *base = store + (int64_t)blockIdx.x * slot_bytes_i8(gga);
  float *G = reinterpret_cast<float *>(base + OFF_G);
  double dx = 0.0, dy = 0.0, dz = 0.0, r = 1.0, rinv = 1.0, R = 0.0, common = 0.0;
#pragma unroll 1
  for (int i = 0; i < 16; ++i) {
    const int f = f0 + i;
    const FuncRec rec = fr[f];
    const bool newc =
        i == 0 || rec.cx != fr[f - 1].cx || rec.cy != fr[f - 1].cy || rec.cz != fr[f - 1].cz;
    if (newc) {
      dx = x - rec.cx;
      dy = y - rec.cy;
      dz = z - rec.cz;
      r = sqrt(dx * dx + dy * dy + dz * dz);
      rinv = 1.0 / r;
    }
    if (newc || rec.zeta != fr[f - 1].zeta || rec.k != fr[f - 1].k ||
        rec.norm != fr[f - 1].norm) {
      R = rec.norm * ipow(r, rec.k) * exp_neg_mixed(rec.zeta * r);
      common = ((double)rec.k * rinv - rec.zeta) * R * rinv;
    }
    double S = 0.0, Sx = 0.0, Sy = 0.0, Sz = 0.0;
    for (int t = 0; t < rec.term_cnt; ++t) {
      const HarmTerm h = T[rec.term_off + t];
      const double xa1 = ipow(dx, h.a > 0 ? h.a - 1 : 0);
      const double yb1 = ipow(dy, h.b > 0 ? h.b - 1 : 0);
      const double zc1 = ipow(dz, h.c > 0 ? h.c - 1 : 0);
      const double xa = h.a > 0 ? xa1 * dx : 1.0;
      const double yb = h.b > 0 ? yb1 * dy : 1.0;
      const double zc = h.c > 0 ? zc1 * dz : 1.0;
      S += h.coef * xa * yb * zc;
      if (gga) {
        if (h.a > 0) Sx += h.coef * h.a * xa1 * yb * zc;
        if (h.b > 0) Sy += h.coef * h.b * xa * yb1 * zc;
        if (h.c > 0) Sz += h.coef * h.c * xa * yb * zc1;
      }
    }
    v[f][g] = R * S;
    if (gga) {
      G[f * PT + g] = (float)(R * Sx + S * (dx * common));
      G[TILE + f * PT + g] = (float)(R * Sy + S * (dy * common));
      G[2 * TILE + f * PT + g] = (float)(R * Sz + S * (dz * common));
    }
  }
  __syncthreads();
  {
    // row maxima along functions (F) and along points (Pm): four threads per
    // row, 16 elements each, then the four partial maxima by shuffles
    const int rr = threadIdx.x >> 2, qt = threadIdx.x & 3;
    double m = 0.0, mp = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      m = fmax(m, fabs(v[rr][16 * qt + k]));
      mp = fmax(mp, fabs(v[16 * qt + k][rr]));
    }
    m = fmax(m, __shfl_xor(m, 1));
    m = fmax(m, __shfl_xor(m, 2));
    mp = fmax(mp, __shfl_xor(mp, 1));
    mp = fmax(mp, __shfl_xor(mp, 2));
    if (qt == 0) {
      eF[rr] = row_exp(m);
      eP[rr] = row_exp(mp);
    }
  }
  __syncthreads();
  store_section<false>(v, eF, base + OFF_F);
  store_section<true>(v, eP, base + OFF_PM);
}

// P (FP64, original order) -> one section per 64 x 64 block pair (fm, fk),
// permuted and zero-padded; one workgroup per block pair.  The row exponents
// are shared by the KG blocks fk of a group (fk / KG): each row's maximum is
// taken over the group's columns, so that the contraction can sum the exact
// int32 slice products of a group's steps before one FP64 combine.
__global__ void __launch_bounds__(256) gather_p_i8(int nbf, int nfb,
                                                   const int *__restrict__ fperm,
                                                   const double *__restrict__ P,
                                                   signed char *__restrict__ Ps) {
  const int fm = blockIdx.x / nfb, fk = blockIdx.x % nfb;
  __shared__ double v[FB][65];
  __shared__ double rmax[4][FB];
  __shared__ int ex[FB];
  for (int e = threadIdx.x; e < TILE; e += 256) {
    const int r = e / FB, c = e % FB;
    const int i = fm * FB + r, j = fk * FB + c;
    v[r][c] = (i < nbf && j < nbf) ? P[(int64_t)fperm[i] * nbf + fperm[j]] : 0.0;
  }
  {
    // row maxima over the columns of the group's blocks: thread (row, part)
    const int r = threadIdx.x & (FB - 1), part = threadIdx.x >> 6;
    const int i = fm * FB + r;
    const int c0 = (fk / KG) * KG * FB;
    const int c1 = min(c0 + KG * FB, nbf);
    double m = 0.0;
    if (i < nbf)
      for (int j = c0 + part; j < c1; j += 4)
        m = fmax(m, fabs(P[(int64_t)fperm[i] * nbf + fperm[j]]));
    rmax[part][r] = m;
  }
  __syncthreads();
  if (threadIdx.x < FB) {
    const double m = fmax(fmax(rmax[0][threadIdx.x], rmax[1][threadIdx.x]),
                          fmax(rmax[2][threadIdx.x], rmax[3][threadIdx.x]));
    ex[threadIdx.x] = row_exp(m);
  }
  __syncthreads();
  store_section<false>(v, ex, Ps + (int64_t)blockIdx.x * SEC);
}


// Per batch, the Pm sections of the blocks of one exponent group (local block
// index / KG) are re-sliced to a common exponent per point (the group's
// largest): one workgroup of 256 threads per batch; thread t takes point row
// t % 64 and the 16 functions 16 (t / 64) .. +15 of every section of the
// group (16 contiguous bytes per plane).  An element's four digits give the
// exact integer V = d0 2^21 + d1 2^14 + d2 2^7 + d3; V 2^-s is rounded (half
// to even, so that the sums over points see no bias) and split again into
// balanced base-128 digits.  A batch's slots come in ascending block order,
// so a group is a run of consecutive slots.
__global__ void __launch_bounds__(256) pm_group_i8(int64_t s0, int64_t b0,
                                                   const int *__restrict__ b_off,
                                                   const int *__restrict__ b_fbl, MolTab mt,
                                                   int gga, signed char *__restrict__ store) {
  const int64_t b = b0 + blockIdx.x;
  const int lo = b_off[b], hi = b_off[b + 1];
  const int pr = threadIdx.x & 63, ch = threadIdx.x >> 6;
  const int off = frag8(pr, 16 * ch);
  const int64_t sb = slot_bytes_i8(gga);
  int k = lo;
  while (k < hi) {
    const int g = mt.blk_local[b_fbl[k]] / KG;
    int ke = k, em = -127;
    for (; ke < hi && mt.blk_local[b_fbl[ke]] / KG == g; ++ke)
      em = max(em, (int)store[(ke - s0) * sb + OFF_PM + 4 * PLANE + pr]);
    for (int kk = k; kk < ke; ++kk) {
      signed char *sec = store + (kk - s0) * sb + OFF_PM;
      const int sh = em - (int)sec[4 * PLANE + pr];
      if (sh <= 0) continue;
      v4i w[4];
#pragma unroll
      for (int pl = 0; pl < 4; ++pl) w[pl] = ld16(sec + pl * PLANE + off);
      const signed char *c0 = reinterpret_cast<const signed char *>(&w[0]);
      const signed char *c1 = reinterpret_cast<const signed char *>(&w[1]);
      const signed char *c2 = reinterpret_cast<const signed char *>(&w[2]);
      const signed char *c3 = reinterpret_cast<const signed char *>(&w[3]);
      __attribute__((aligned(16))) signed char out[4][16];
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        int V = (int)c0[e] * (1 << 21) + (int)c1[e] * (1 << 14) + (int)c2[e] * (1 << 7) +
                (int)c3[e];
        if (sh > 28) {
          V = 0;
        } else {
          const int qv = V >> sh, rv = V - qv * (1 << sh), half = 1 << (sh - 1);
          V = qv + ((rv > half || (rv == half && (qv & 1))) ? 1 : 0);
        }
#pragma unroll
        for (int pl = 3; pl > 0; --pl) {
          const int d = ((V + 64) & 127) - 64;
          out[pl][e] = (signed char)d;
          V = (V - d) >> 7;
        }
        out[0][e] = (signed char)V;
      }
#pragma unroll
      for (int pl = 0; pl < 4; ++pl)
        *reinterpret_cast<v4i *>(sec + pl * PLANE + off) = *reinterpret_cast<const v4i *>(out[pl]);
    }
    __syncthreads();  // every thread has read the old exponents of the group
    if (ch == 0)
      for (int kk = k; kk < ke; ++kk)
        store[(kk - s0) * sb + OFF_PM + 4 * PLANE + pr] = (signed char)em;
    k = ke;
  }
}

// ---------------------------------------------------------------------------
// Density stage, three kernels: the contraction (density_x2_i8), per batch
// the reduction and the functional (density_f_i8), the Z slices (z_slots_i8).
//
// density_x2_i8: X = P Phi for a pair of function blocks of a batch group.
// One workgroup of four waves per (batch group of GBT batches, pair of
// union-list blocks m_0 < m_1); wave w holds rows 32 (w % 2) .. +31 of block
// m_(w / 2), so each B fragment read from LDS feeds two tile products.  The
// active k of the union list and their masks are listed in LDS once; the
// steps run over them in ascending order, each staging the Pm sections of
// (k, batch) in LDS (two buffers: the next step's copies run while the
// current one computes, one barrier per step) and each wave's P fragments in
// registers one step ahead.  Workgroup w takes batch group 8 (w / 8 / nmg) +
// w % 8 and pair (w / 8) % nmg, so the pairs of one batch group run on one
// XCD (workgroups are dealt to the eight XCDs in turn) and share its L2 for
// the Pm sections.  The epilogue writes, per live slot (m, batch), the column
// sums of Phi X (and grad Phi X) over the block's 64 functions, summed over
// the block's functions in order, to part; Phi comes from the slot's Pm
// section.
// ABS (one batch per workgroup): the pair is an absolute pair of the
// molecule's blocks (2 p, 2 p + 1; a workgroup whose two blocks are both dead
// in its batch exits), and the workgroups are dealt in tiles of tpn pairs x
// tbn consecutive batches, one whole tile to an XCD at a time, so that the
// tile's P sections and Pm sections are read from that XCD's L2 by all its
// batches and pairs rather than once per workgroup from the memory side.
template <int GBT, bool ABS>
__global__ void __launch_bounds__(256) density_x2_i8(
    MolTab mt, int nfb, const signed char *__restrict__ Ps, const int *__restrict__ b_off,
    const int *__restrict__ b_fbl, int64_t s0, int64_t b0, int64_t b1, int nmg, int tpn,
    int tbn, int gga, const signed char *__restrict__ store, double *__restrict__ part) {
  static_assert(!ABS || GBT == 1, "absolute pairs take one batch per workgroup");
  static_assert(GBT <= 3, "step masks take bits 0 .. 2 GBT - 1, bit 7 ends a group");
  const int xcd = blockIdx.x & 7;
  const int rr = blockIdx.x >> 3;
  int mgi = 0;
  int64_t bg;
  if (ABS) {  // nmg = pairs of the largest molecule; tile = (pair group, batch group)
    const int tsz = tpn * tbn;
    const int tile = (rr / tsz) * 8 + xcd, li = rr % tsz;
    const int ntp = (nmg + tpn - 1) / tpn;
    mgi = (tile % ntp) * tpn + li % tpn;
    bg = b0 + (int64_t)(tile / ntp) * tbn + li / tpn;
    if (mgi >= nmg) return;
  } else {
    mgi = rr % nmg;
    bg = b0 + ((int64_t)(rr / nmg) * 8 + xcd) * GBT;
  }
  if (bg >= b1) return;
  // one batch: its union list is its own block list, so a pair index past
  // it has nothing to do (common in molecule sets, where nmg follows the
  // largest molecule)
  if (!ABS && GBT == 1 && 2 * mgi >= b_off[bg + 1] - b_off[bg]) return;
  const int nsub = (int)(b1 - bg < GBT ? b1 - bg : GBT);
  // wave-uniform values (the wave index, and the step lists and slot table
  // read from LDS) go through readfirstlane, so that the branches on them
  // are scalar branches
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int mb = wv >> 1, rh = wv & 1;
  const int l16 = lane & 15, q = lane >> 4;
  const int64_t sb_bytes = slot_bytes_i8(gga);

  // Pm sections of a step, two buffers, declared apart so that a copy into
  // one is known not to touch the other
  __shared__ __attribute__((aligned(16))) signed char bufA[GBT][SEC];
  __shared__ __attribute__((aligned(16))) signed char bufB[GBT][SEC];
  __shared__ double red[4][4][PT];
  // slot table and union list of the batch group (one batch: not needed)
  __shared__ int tab[GBT == 1 ? 1 : GBT][GBT == 1 ? 1 : MAXNFB];
  __shared__ int ublk[GBT == 1 ? 1 : MAXNFB];
  __shared__ int pos_s[2];
  // active steps, one record each: k | mask << 16 (mask bit m * GBT + sb:
  // block m and batch sb take part; bit 7: last step of an exponent group),
  // then the slot of (k, sb) per batch (-1 where k is dead in sb)
  constexpr int SR = GBT == 1 ? 2 : 4;
  typedef int Rec __attribute__((ext_vector_type(SR)));
  __shared__ __attribute__((aligned(16))) int srec[MAXNFB][SR];
  __shared__ int nu_s, nst_s;

  int fm0, fm1;
  int sl0 = -1, sl1 = -1;  // one batch: the slots of fm0, fm1
  if constexpr (GBT == 1) {
    // one batch: the steps are the batch's own blocks (ascending), its slots
    // lo .. hi - 1, and both blocks of the pair take part in every step
    const int lo = b_off[bg], hi = b_off[bg + 1], n = hi - lo;
    if (ABS) {  // blocks 2 mgi, 2 mgi + 1 of the batch's molecule, if live
      if (n == 0) return;  // uniform: empty batch
      const int f0 = b_fbl[lo];
      const int blk0 = f0 - mt.blk_local[f0], nfm = mt.nfb[mt.blk_mol[f0]];
      const int a = 2 * mgi < nfm ? blk0 + 2 * mgi : -1;
      const int c = 2 * mgi + 1 < nfm ? blk0 + 2 * mgi + 1 : -1;
      if (threadIdx.x < 2) pos_s[threadIdx.x] = -1;
      __syncthreads();
      for (int i = threadIdx.x; i < n; i += 256) {
        const int f = b_fbl[lo + i];
        if (a >= 0 && f == a) pos_s[0] = i;
        if (c >= 0 && f == c) pos_s[1] = i;
      }
      __syncthreads();
      const int pa = __builtin_amdgcn_readfirstlane(pos_s[0]);
      const int pc = __builtin_amdgcn_readfirstlane(pos_s[1]);
      if (pa < 0 && pc < 0) return;  // uniform: both blocks dead here
      fm0 = pa >= 0 ? a : c;
      sl0 = lo + (pa >= 0 ? pa : pc);
      fm1 = pa >= 0 && pc >= 0 ? c : -1;
      sl1 = fm1 >= 0 ? lo + pc : -1;
    } else {
      fm0 = __builtin_amdgcn_readfirstlane(b_fbl[lo + 2 * mgi]);
      sl0 = lo + 2 * mgi;
      fm1 = 2 * mgi + 1 < n ? __builtin_amdgcn_readfirstlane(b_fbl[lo + 2 * mgi + 1]) : -1;
      sl1 = fm1 >= 0 ? sl0 + 1 : -1;
    }
    // records; bit 7: the last step of its exponent group (blocks with the
    // same local index / KG of one molecule, a run of steps)
    const unsigned pmask = 1u | (fm1 >= 0 ? 2u : 0u);
    for (int i = threadIdx.x; i < n; i += 256) {
      const int f = b_fbl[lo + i];
      const int key = f - mt.blk_local[f] % KG;
      bool last = i + 1 == n;
      if (!last) {
        const int f2 = b_fbl[lo + i + 1];
        last = f2 - mt.blk_local[f2] % KG != key;
      }
      srec[i][0] = f | (int)((pmask | (last ? 0x80u : 0u)) << 16);
      srec[i][1] = lo + i;
    }
    if (threadIdx.x == 0) nst_s = n;
    __syncthreads();
  } else {
    for (int e = threadIdx.x; e < GBT * nfb; e += 256) tab[e / nfb][e % nfb] = -1;
    __syncthreads();
    for (int sb = 0; sb < nsub; ++sb) {
      const int lo = b_off[bg + sb], hi = b_off[bg + sb + 1];
      for (int k = lo + (int)threadIdx.x; k < hi; k += 256) tab[sb][b_fbl[k]] = k;
    }
    __syncthreads();
    if (wv == 0) {
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
    const int nu = __builtin_amdgcn_readfirstlane(nu_s);
    if (ABS) {  // blocks 2 mgi, 2 mgi + 1 of the batch's molecule, if live
      const int lo = b_off[bg];
      if (lo == b_off[bg + 1]) return;  // uniform: empty batch
      const int f0 = b_fbl[lo];
      const int blk0 = f0 - mt.blk_local[f0], nfm = mt.nfb[mt.blk_mol[f0]];
      int a = 2 * mgi < nfm ? blk0 + 2 * mgi : -1;
      int c = 2 * mgi + 1 < nfm ? blk0 + 2 * mgi + 1 : -1;
      if (a >= 0 && tab[0][a] < 0) a = -1;
      if (c >= 0 && tab[0][c] < 0) c = -1;
      if (a < 0) {
        a = c;
        c = -1;
      }
      if (a < 0) return;  // uniform: both blocks dead here
      fm0 = a;
      fm1 = c;
    } else {
      if (mgi * 2 >= nu) return;  // uniform: no block of this pair
      fm0 = __builtin_amdgcn_readfirstlane(ublk[mgi * 2]);
      fm1 = mgi * 2 + 1 < nu ? __builtin_amdgcn_readfirstlane(ublk[mgi * 2 + 1]) : -1;
    }
    if (wv == 0) {
      int cnt = 0;
      for (int base = 0; base < nu; base += 64) {
        const int ki = base + lane;
        unsigned pm = 0;
        if (ki < nu) {
          const int fk = ublk[ki];
  #pragma unroll
          for (int sb = 0; sb < GBT; ++sb) {
            if (sb >= nsub || tab[sb][fk] < 0) continue;
            if (tab[sb][fm0] >= 0) pm |= 1u << sb;
            if (fm1 >= 0 && tab[sb][fm1] >= 0) pm |= 1u << (GBT + sb);
          }
        }
        const unsigned long long m = __ballot(pm != 0);
        if (pm) {
          const int i = cnt + __popcll(m & ((1ull << lane) - 1ull));
          const int fk = ublk[ki];
          srec[i][0] = fk | (int)(pm << 16);
  #pragma unroll
          for (int sb = 0; sb < SR - 1; ++sb)
            srec[i][1 + sb] = sb < nsub ? tab[sb < GBT ? sb : 0][fk] : -1;
        }
        cnt += __popcll(m);
      }
      // bit 7 of a step's mask: the last step of its exponent group (blocks
      // with the same local index / KG of one molecule; the steps are in
      // ascending block order, so a group is a run of steps)
      for (int i = lane; i < cnt; i += 64) {
        const int f = srec[i][0] & 0xFFFF;
        const int key = f - mt.blk_local[f] % KG;
        bool last = i + 1 == cnt;
        if (!last) {
          const int f2 = srec[i + 1][0] & 0xFFFF;
          last = f2 - mt.blk_local[f2] % KG != key;
        }
        if (last) srec[i][0] |= 0x80 << 16;
      }
      if (lane == 0) nst_s = cnt;
    }
    __syncthreads();
  }
  const int nst = __builtin_amdgcn_readfirstlane(nst_s);
  const int nq = gga ? 4 : 1;
  const int fm = mb ? fm1 : fm0;
  constexpr unsigned ROW = (1u << GBT) - 1u;
  // this wave's P row: sections (fm, k) of its molecule, whose blocks are
  // numbered consecutively from mbase (so no table load per step: such a load
  // would make the wave wait for the section copies issued before it)
  const signed char *prow = nullptr;
  int mbase = 0;
  if (fm >= 0) {
    const int mol = mt.blk_mol[fm];
    mbase = fm - mt.blk_local[fm];
    prow = Ps + mt.Ps_off[mol] + (int64_t)(fm - mbase) * mt.nfb[mol] * SEC;
  }

  // P fragments and row exponents of a step (two register sets, used in turn)
  struct AFrag {
    v4i v[2][4];
    int e[2];
  };
  // a step's record, read one step before its copies are issued (so that no
  // LDS look-up sits between a barrier and the copies); past the end: any
  // record in bounds, never used
  auto rec = [&](int i) -> Rec {
    return *reinterpret_cast<const Rec *>(&srec[i < nst ? i : 0][0]);
  };
  auto issue = [&](const Rec &rn, signed char (*dst)[SEC], AFrag &an) {
    const int r0 = __builtin_amdgcn_readfirstlane(rn[0]);
    const int fk = r0 & 0xFFFF;
    const unsigned pm = ((unsigned)r0 >> 16) & 0xFFu;
#pragma unroll
    for (int sb = 0; sb < GBT; ++sb) {
      if (!(((pm >> sb) | (pm >> (GBT + sb))) & 1u)) continue;
      dma_section<4>(store + (__builtin_amdgcn_readfirstlane(rn[1 + sb]) - s0) * sb_bytes + OFF_PM,
                     &dst[sb][0],
                     (wv + sb) & 3, lane);
    }
    if ((pm >> (mb * GBT)) & ROW) {
      const signed char *pa = prow + (int64_t)(fk - mbase) * SEC;
#pragma unroll
      for (int f = 0; f < 2; ++f) {
#pragma unroll
        for (int p = 0; p < 4; ++p)
          an.v[f][p] = ld16(pa + p * PLANE + ((2 * rh + f) * 64 + lane) * 16);
        an.e[f] = *reinterpret_cast<const int *>(pa + 4 * PLANE + 16 * (2 * rh + f) + 4 * q);
      }
    }
  };

  // x64: FP64 sums over the groups done; hacc, lacc: exact int32 sums of the
  // current group's tile products (hi = 128 d0 + d1, lo = 128 d2 + d3), whose
  // row and point exponents (eAg from the P fragments, eBg from the Pm
  // sections) are the same for every step of the group
  double x64[GBT][2][4][4];
  v4i hacc[GBT][2][4], lacc[GBT][2][4];
  int eAg[2] = {0, 0};
  int eBg[GBT][4];
#pragma unroll
  for (int sb = 0; sb < GBT; ++sb)
#pragma unroll
    for (int f = 0; f < 2; ++f)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        hacc[sb][f][j] = v4i{0, 0, 0, 0};
        lacc[sb][f][j] = v4i{0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < 4; ++r) x64[sb][f][j][r] = 0.0;
      }
#pragma unroll
  for (int sb = 0; sb < GBT; ++sb)
#pragma unroll
    for (int j = 0; j < 4; ++j) eBg[sb][j] = 0;
  // One step: wait for its operands, start the next step's copies into the
  // other buffer and register set, then the tile products into the group's
  // int32 sums; after a group's last step, the sums are scaled into x64 once.
  // Rows or points stored as zero have zero slices, so their scale does not
  // matter.
  auto step = [&](int i, const Rec &rc, const Rec &rn, signed char (*cur)[SEC],
                  signed char (*nxt)[SEC], const AFrag &ac, AFrag &an) {
    wait_vm0();
    __syncthreads();  // buffer cur complete, nxt free
    const unsigned pmc = ((unsigned)__builtin_amdgcn_readfirstlane(rc[0]) >> 16) & 0xFFu;
    if (i + 1 < nst) issue(rn, nxt, an);
    if ((pmc >> (mb * GBT)) & ROW) {
      eAg[0] = ac.e[0];
      eAg[1] = ac.e[1];
    }
#pragma unroll
    for (int sb = 0; sb < GBT; ++sb) {
      if (!((pmc >> (mb * GBT + sb)) & 1u)) continue;
      const signed char *bb = &cur[sb][0];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        v4i B[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) B[p] = ld16(bb + p * PLANE + (j * 64 + lane) * 16);
        eBg[sb][j] = bb[4 * PLANE + 16 * j + l16];
#pragma unroll
        for (int f = 0; f < 2; ++f) tile_product_acc(ac.v[f], B, hacc[sb][f][j], lacc[sb][f][j]);
      }
    }
    if (pmc & 0x80u) {
#pragma unroll
      for (int sb = 0; sb < GBT; ++sb)
#pragma unroll
        for (int f = 0; f < 2; ++f)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
              x64[sb][f][j][r] =
                  fma(fma((double)lacc[sb][f][j][r], 0x1p-14, (double)hacc[sb][f][j][r]),
                      pow2((int)(signed char)(eAg[f] >> (8 * r)) + eBg[sb][j] - 19),
                      x64[sb][f][j][r]);
            hacc[sb][f][j] = v4i{0, 0, 0, 0};
            lacc[sb][f][j] = v4i{0, 0, 0, 0};
          }
    }
  };
  AFrag a0, a1;
  Rec rA = rec(0), rB = rec(1);
  if (nst > 0) issue(rA, bufA, a0);
  for (int i = 0; i < nst; i += 2) {
    step(i, rA, rB, bufA, bufB, a0, a1);
    rA = rec(i + 2);
    if (i + 1 < nst) {
      step(i + 1, rB, rA, bufB, bufA, a1, a0);
      rB = rec(i + 3);
    }
  }

  // Epilogue: per live slot (m, sb), column sums over the block's functions.
  // The two waves of block m put their X rows into LDS (over the step
  // buffers); then every thread takes one point g and 16 consecutive
  // functions, so that the loads of grad Phi (rows of 64 points) and of the
  // Pm slices (16 functions of a point: one 16-byte load per plane) are whole
  // rows, and the four partial sums of a point are added in order.
  // rows 0..31 of the X tile over bufA, rows 32..63 over bufB
  double (*XsA)[65] = reinterpret_cast<double (*)[65]>(&bufA[0][0]);
  double (*XsB)[65] = reinterpret_cast<double (*)[65]>(&bufB[0][0]);
  static_assert(sizeof(bufA) >= sizeof(double) * 32 * 65, "X tile must fit the step buffers");
  auto Xs = [&](int mu) -> double * { return mu < 32 ? XsA[mu] : XsB[mu - 32]; };
  __syncthreads();  // the steps are done with the buffers
  for (int sb = 0; sb < nsub; ++sb) {
#pragma unroll 1
    for (int m = 0; m < 2; ++m) {
      const int fmx = m ? fm1 : fm0;
      const int sl = GBT == 1 ? (m ? sl1 : sl0)
                              : (fmx >= 0 ? __builtin_amdgcn_readfirstlane(tab[sb][fmx]) : -1);
      if (sl < 0) continue;  // uniform
      if (mb == m) {
#pragma unroll
        for (int f = 0; f < 2; ++f)
#pragma unroll
          for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              double x = 0.0;
#pragma unroll
              for (int s2 = 0; s2 < GBT; ++s2)
                if (s2 == sb) x = x64[s2][f][j][r];
              Xs(32 * rh + 16 * f + 4 * q + r)[16 * j + l16] = x;
            }
      }
      __syncthreads();
      {
        // every global load of the slot is issued before the first use
        // (a branch between them would make each wait for the one before)
        const int g = threadIdx.x & 63, qq = threadIdx.x >> 6;
        const signed char *tb = store + (sl - s0) * sb_bytes;
        const float *G = reinterpret_cast<const float *>(tb + OFF_G) + 16 * qq * PT + g;
        const int eP = tb[OFF_PM + 4 * PLANE + g];
        const int off = frag8(g, 16 * qq);
        v4i pw4[4];
#pragma unroll
        for (int pl = 0; pl < 4; ++pl) pw4[pl] = ld16(tb + OFF_PM + pl * PLANE + off);
        float gv[3][16];
        if (gga) {
#pragma unroll
          for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int e = 0; e < 16; ++e) gv[c][e] = G[c * TILE + e * PT];
        } else {
#pragma unroll
          for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int e = 0; e < 16; ++e) gv[c][e] = 0.0f;
        }
        const signed char *c0 = reinterpret_cast<const signed char *>(&pw4[0]);
        const signed char *c1 = reinterpret_cast<const signed char *>(&pw4[1]);
        const signed char *c2 = reinterpret_cast<const signed char *>(&pw4[2]);
        const signed char *c3 = reinterpret_cast<const signed char *>(&pw4[3]);
        // 2^(eP - 6) scales the slices exactly; 0 for a point stored as zero
        const double sP = pow2_or_zero(eP == -127 ? -127 : eP - 6);
        double sm[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const double x = Xs(16 * qq + e)[g];
          const double v = (double)c0[e] + ldexp((double)c1[e], -7) +
                           ldexp((double)c2[e], -14) + ldexp((double)c3[e], -21);
          sm[0] += (v * sP) * x;
          sm[1] += (double)gv[0][e] * x;
          sm[2] += (double)gv[1][e] * x;
          sm[3] += (double)gv[2][e] * x;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) red[qq][c][g] = sm[c];
      }
      __syncthreads();
      {
        const int c = threadIdx.x >> 6, g = threadIdx.x & 63;
        const double acc =
            c < nq ? ((red[0][c][g] + red[1][c][g]) + red[2][c][g]) + red[3][c][g] : 0.0;
        part[(sl - s0) * (4 * PT) + c * PT + g] = acc;
      }
      __syncthreads();  // red and the X tile free for the next slot
    }
  }
}

// Per batch: column sums of the live slots in ascending block order,
// functional, Z coefficients, (E, N) partial sums.
__global__ void __launch_bounds__(256) density_f_i8(
    int64_t s0, int64_t b0, const int *__restrict__ b_off, int gga,
    const double *__restrict__ part, const double *__restrict__ pw, XcList xl,
    double *__restrict__ coef, double *__restrict__ blockEN, double *__restrict__ epoint) {
  const int64_t b = b0 + blockIdx.x;
  __shared__ double cs[4][PT];
  const int c = threadIdx.x / PT, g = threadIdx.x % PT;
  const int lo = b_off[b], hi = b_off[b + 1];
  double acc = 0.0;
  for (int k = lo; k < hi; ++k) acc += part[(k - s0) * (4 * PT) + c * PT + g];
  cs[c][g] = acc;
  __syncthreads();
  if (threadIdx.x < PT) {
    const int64_t gp = b * PT + g;
    const double w = pw[gp];
    const double rho = cs[0][g];
    const double gxr = 2.0 * cs[1][g], gyr = 2.0 * cs[2][g], gzr = 2.0 * cs[3][g];
    const double sigma = gga ? gxr * gxr + gyr * gyr + gzr * gzr : 0.0;
    const XcPoint o = xc_eval(xl.n, xl.ids, xl.thr, rho, sigma);
    const double a = 0.5 * w * o.vrho;
    const double bs = gga ? 2.0 * w * o.vsigma : 0.0;
    coef[4 * gp + 0] = a;
    coef[4 * gp + 1] = bs * gxr;
    coef[4 * gp + 2] = bs * gyr;
    coef[4 * gp + 3] = bs * gzr;
    double e = w * o.e, n = w * rho;
    if (epoint) epoint[gp] = e;  // w e_xc of this point
    for (int o2 = 32; o2 > 0; o2 >>= 1) {
      e += __shfl_down(e, o2);
      n += __shfl_down(n, o2);
    }
    if (g == 0) {
      blockEN[3 * b] = e;
      blockEN[3 * b + 1] = n;
      blockEN[3 * b + 2] = 0.0;
    }
  }
}

// Spin-polarised form of density_f_i8: the slot sums of the two spin
// contractions (part_a, part_b) give rho_s and grad rho_s per point; the
// polarised functional gives the Z coefficients of each spin,
//   a_s = w v_rho_s / 2,  b_s = w (2 v_sigma_ss grad rho_s + v_sigma_ab grad rho_s'),
// (xc_integrals.hpp:655-664) into coef_a, coef_b ([point][4] each), and
// (E, N_alpha, N_beta) per batch.
__global__ void __launch_bounds__(256) density_fp_i8(
    int64_t s0, int64_t b0, const int *__restrict__ b_off, int gga,
    const double *__restrict__ part_a, const double *__restrict__ part_b,
    const double *__restrict__ pw, XcList xl, double *__restrict__ coef_a,
    double *__restrict__ coef_b, double *__restrict__ blockEN, double *__restrict__ epoint) {
  const int64_t b = b0 + blockIdx.x;
  __shared__ double cs[2][4][PT];
  const int c = threadIdx.x / PT, g = threadIdx.x % PT;
  const int lo = b_off[b], hi = b_off[b + 1];
  double acc_a = 0.0, acc_b = 0.0;
  for (int k = lo; k < hi; ++k) {
    acc_a += part_a[(k - s0) * (4 * PT) + c * PT + g];
    acc_b += part_b[(k - s0) * (4 * PT) + c * PT + g];
  }
  cs[0][c][g] = acc_a;
  cs[1][c][g] = acc_b;
  __syncthreads();
  if (threadIdx.x < PT) {
    const int64_t gp = b * PT + g;
    const double w = pw[gp];
    const double r2[2] = {cs[0][0][g], cs[1][0][g]};
    double gr[2][3];
    for (int s2 = 0; s2 < 2; ++s2)
      for (int k = 0; k < 3; ++k) gr[s2][k] = 2.0 * cs[s2][k + 1][g];
    double sg[3] = {0.0, 0.0, 0.0};
    if (gga) {
      sg[0] = gr[0][0] * gr[0][0] + gr[0][1] * gr[0][1] + gr[0][2] * gr[0][2];
      sg[1] = gr[0][0] * gr[1][0] + gr[0][1] * gr[1][1] + gr[0][2] * gr[1][2];
      sg[2] = gr[1][0] * gr[1][0] + gr[1][1] * gr[1][1] + gr[1][2] * gr[1][2];
    }
    double exc, vr[2], vs[3];
    xc_eval_polarized(xl.n, xl.ids, xl.thr, r2, sg, exc, vr, vs);
    double *cf[2] = {coef_a, coef_b};
    for (int s2 = 0; s2 < 2; ++s2) {
      const double vss = gga ? 2.0 * w * vs[2 * s2] : 0.0, vab = gga ? w * vs[1] : 0.0;
      cf[s2][4 * gp] = 0.5 * w * vr[s2];
      for (int k = 0; k < 3; ++k) cf[s2][4 * gp + 1 + k] = vss * gr[s2][k] + vab * gr[1 - s2][k];
    }
    double e = w * (r2[0] + r2[1]) * exc, na = w * r2[0], nb = w * r2[1];
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
}

// Z = a Phi + b . grad Phi of one live slot in FP64, per-function exponents,
// slices (F order) written at zout + (slot - s0) zstride (the slot's own Z
// section, or a separate array for the second spin); one workgroup per slot.  Thread t owns the 16 entries of
// fragment item t (row 16 (t / 64) + t % 16, points 16 ((t % 64) / 16) ..
// +15): 16-byte loads of the four F planes, 64-byte loads of each gradient
// row piece, the row maximum over the four threads of a row by shuffles,
// 16-byte stores of the Z planes.
__global__ void __launch_bounds__(256) z_slots_i8(int64_t s0, const int *__restrict__ slot_b,
                                                  int gga, const double *__restrict__ coef,
                                                  signed char *__restrict__ store,
                                                  signed char *__restrict__ zout,
                                                  int64_t zstride) {
  const int64_t sl = s0 + blockIdx.x;
  const int64_t b = slot_b[sl];
  __shared__ double cf[PT][4];
  cf[threadIdx.x >> 2][threadIdx.x & 3] = coef[4 * b * PT + threadIdx.x];
  signed char *tb = store + (int64_t)blockIdx.x * slot_bytes_i8(gga);
  const int t = threadIdx.x, l = t & 63;
  const int row = 16 * (t >> 6) + (l & 15), k0 = 16 * (l >> 4);
  v4i fp[4];
#pragma unroll
  for (int p = 0; p < 4; ++p) fp[p] = ld16(tb + OFF_F + p * PLANE + t * 16);
  const int eF = tb[OFF_F + 4 * PLANE + row];
  const float *G = reinterpret_cast<const float *>(tb + OFF_G) + row * PT + k0;
  // every load is issued before the first use (a branch between them would
  // make each wait for the one before)
  float4 gq[3][4];
  if (gga) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q4 = 0; q4 < 4; ++q4)
        gq[c][q4] = *reinterpret_cast<const float4 *>(G + c * TILE + 4 * q4);
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q4 = 0; q4 < 4; ++q4) gq[c][q4] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  __syncthreads();
  const signed char *c0 = reinterpret_cast<const signed char *>(&fp[0]);
  const signed char *c1 = reinterpret_cast<const signed char *>(&fp[1]);
  const signed char *c2 = reinterpret_cast<const signed char *>(&fp[2]);
  const signed char *c3 = reinterpret_cast<const signed char *>(&fp[3]);
  // 2^(eF - 6) scales the slices exactly; 0 for a row stored as zero
  const double sF = pow2_or_zero(eF == -127 ? -127 : eF - 6);
  double z[16];
  double m = 0.0;
#pragma unroll
  for (int q4 = 0; q4 < 4; ++q4) {
    const float gxa[4] = {gq[0][q4].x, gq[0][q4].y, gq[0][q4].z, gq[0][q4].w},
                gya[4] = {gq[1][q4].x, gq[1][q4].y, gq[1][q4].z, gq[1][q4].w},
                gza[4] = {gq[2][q4].x, gq[2][q4].y, gq[2][q4].z, gq[2][q4].w};
#pragma unroll
    for (int e4 = 0; e4 < 4; ++e4) {
      const int e = 4 * q4 + e4, g = k0 + e;
      const double v = (double)c0[e] + ldexp((double)c1[e], -7) + ldexp((double)c2[e], -14) +
                       ldexp((double)c3[e], -21);
      const double ph = v * sF;
      const double zz = gga ? z_value(ph, (double)gxa[e4], (double)gya[e4], (double)gza[e4],
                                      cf[g][0], cf[g][1], cf[g][2], cf[g][3])
                            : cf[g][0] * ph;
      z[e] = zz;
      m = fmax(m, fabs(zz));
    }
  }
  m = fmax(m, __shfl_xor(m, 16));
  m = fmax(m, __shfl_xor(m, 32));
  const int ez = row_exp(m);
  __attribute__((aligned(16))) signed char out[4][16];
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    signed char sc[4];
    slice4(z[e], ez, sc);
#pragma unroll
    for (int p = 0; p < 4; ++p) out[p][e] = sc[p];
  }
  signed char *zs = zout + (int64_t)blockIdx.x * zstride;
#pragma unroll
  for (int p = 0; p < 4; ++p)
    *reinterpret_cast<v4i *>(zs + p * PLANE + t * 16) = *reinterpret_cast<const v4i *>(out[p]);
  if (l < 16) zs[4 * PLANE + row] = (signed char)ez;
}

// ---------------------------------------------------------------------------
// V contraction on 2 x 2 block super-tiles: acc(I, J) += Phi_I Z_J^T +
// Z_I Phi_J^T over the batches where I and J are live, for the four tiles
// (I_a, J_b) of a super-pair (blocks I_0 < I_1, J_0 < J_1 of one molecule;
// a missing second block is -1; on the diagonal I_1 x J_0 is the transpose
// of I_0 x J_1 and has no pair index).  One workgroup of eight waves per
// (super-pair, split); its batches b = lo + split + nsplit i are taken in
// windows of 512 candidates, compacted in ascending order into an LDS list
// of the four slots.  The Z and F sections of the J slots (shared by all
// waves) are staged in LDS, two buffers: the next batch's are copied in
// while the current one computes.  Wave w owns rows 32 h .. 32 h + 31 of
// block I_a and the 64 columns of block J_b (a = w / 4, h = (w / 2) % 2,
// b = w % 2), so that every J fragment read from LDS feeds two 16-row tile
// products.  A batch runs in two phases, F_I Z_J^T and then Z_I F_J^T; the
// wave's I fragments live in registers, one set each: Z_I is loaded when the
// batch starts (used in the second phase) and F_I of the next batch between
// the phases (F_I is dead by then).  Each element sums its batches in
// ascending order, F_I Z_J^T before Z_I F_J^T within a batch, as in the
// one-pair kernel.  Each tile is added into vpart at its pair index.
__global__ void __launch_bounds__(512) vgemm_i8(
    int nsp, int nsplit, const int4 *__restrict__ sp_blk, const int4 *__restrict__ sp_pidx,
    int npairs, const int *__restrict__ tile_slot, MolTab mt, int64_t b0, int64_t b1,
    int64_t s0, int gga, const signed char *__restrict__ store,
    const signed char *__restrict__ zbase, int64_t zstride, double *__restrict__ vpart) {
  // split-major: the super-pairs of one split run together and walk the same
  // batches, sharing each batch's slots in the caches
  const int sp = blockIdx.x % nsp;
  const int split = blockIdx.x / nsp;
  const int4 blk = sp_blk[sp];
  // wave-uniform values through readfirstlane, so that the branches on them
  // are scalar branches
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wa = wv >> 2, wh = (wv >> 1) & 1, wb = wv & 1;
  const int l16 = lane & 15, q = lane >> 4;
  const int64_t sb_bytes = slot_bytes_i8(gga);
  // J sections, two buffers (Z_J0 F_J0 Z_J1 F_J1 each), declared apart so
  // that a copy into one is known not to touch the other
  __shared__ __attribute__((aligned(16))) signed char bufA[4][SEC];
  __shared__ __attribute__((aligned(16))) signed char bufB[4][SEC];
  __shared__ int ls[4][512];  // slots of I_0, I_1, J_0, J_1 per listed batch
  __shared__ int wcnt[8];
  // this wave's tile and its pair index (-1: no work for this wave)
  const int4 pix = sp_pidx[sp];
  const int pp = wa ? (wb ? pix.w : pix.z) : (wb ? pix.y : pix.x);
  double v64[2][4][4];
#pragma unroll
  for (int g = 0; g < 2; ++g)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) v64[g][j][r] = 0.0;
  // batches of the super-pair's molecule inside [b0, b1)
  const int mol = mt.blk_mol[blk.x];
  const int64_t mb0 = mt.b0[mol], mnb = mt.nb[mol];
  const int *ts0 = tile_slot + mt.ts_off[mol] - mb0;
  const int bl[4] = {blk.x, blk.y, blk.z, blk.w};
  int64_t tso[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) tso[u] = bl[u] >= 0 ? (int64_t)mt.blk_local[bl[u]] * mnb : -1;
  const int64_t lo = b0 > mb0 ? b0 : mb0, hi = b1 < mb0 + mnb ? b1 : mb0 + mnb;

  // this wave's I fragments (two 16-row groups) and row exponents
  v4i Fr[2][4], Zr[2][4];
  int eF[2] = {0, 0}, eZ[2] = {0, 0};
  auto f_section = [&](int s) { return store + (s - s0) * sb_bytes + OFF_F; };
  auto z_section = [&](int s) { return zbase + (s - s0) * zstride; };
  auto load_I = [&](const signed char *sec, v4i (*dst)[4], int *ex) {
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const int rg = 2 * wh + g;
#pragma unroll
      for (int pl = 0; pl < 4; ++pl) dst[g][pl] = ld16(sec + pl * PLANE + (rg * 64 + lane) * 16);
      ex[g] = *reinterpret_cast<const int *>(sec + 4 * PLANE + 16 * rg + 4 * q);
    }
  };
  // whether the wave's tile has work in listed batch i
  auto active = [&](int i) {
    return pp >= 0 && __builtin_amdgcn_readfirstlane(ls[wa][i]) >= 0 &&
           __builtin_amdgcn_readfirstlane(ls[2 + wb][i]) >= 0;
  };
  // F_I of batch i; where the wave's tile is idle in batch i (or past the
  // list) the loads read slot s0 instead, whose values go unused, so that
  // every path issues the same loads and the waits after them stay exact
  auto next_F = [&](int i, int n) {
    const int sI = i < n && active(i) ? __builtin_amdgcn_readfirstlane(ls[wa][i]) : (int)s0;
    load_I(f_section(sI), Fr, eF);
  };
  // J sections of listed batch i -> LDS buffer dst (direct copies)
  auto issue_J = [&](int i, signed char (*dst)[SEC]) {
#pragma unroll
    for (int jb = 0; jb < 2; ++jb) {
      const int sJ = __builtin_amdgcn_readfirstlane(ls[2 + jb][i]);
      if (sJ < 0) continue;
      dma_section<8>(z_section(sJ), &dst[2 * jb][0], (wv + 4 * jb) & 7, lane);
      dma_section<8>(f_section(sJ), &dst[2 * jb + 1][0], (wv + 4 * jb + 2) & 7, lane);
    }
  };
  // the first batch of a window: J sections through VGPRs and LDS stores
  // (with a direct copy into bufA ahead of the batch loop as well, the
  // compiler's wait insertion stops telling the copies into bufB apart from
  // the reads of bufA and waits for them before every batch's first LDS
  // read), and F_I
  auto fill = [&](int n, signed char (*dst)[SEC]) {
#pragma unroll
    for (int jb = 0; jb < 2; ++jb) {
      const int sJ = __builtin_amdgcn_readfirstlane(ls[2 + jb][0]);
      if (sJ < 0) continue;
      const signed char *src[2] = {z_section(sJ), f_section(sJ)};
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        v4i t[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int o = 16 * threadIdx.x + k * 16 * 512;
          if (o < SEC) t[k] = ld16(src[u] + o);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int o = 16 * threadIdx.x + k * 16 * 512;
          if (o < SEC) *reinterpret_cast<v4i *>(&dst[2 * jb + u][o]) = t[k];
        }
      }
    }
    next_F(0, n);
  };
  // one phase of a batch: A = F_I (p = 0) or Z_I (p = 1) against the Z or F
  // section of J_b; the B fragment of column fragment j + 1 is read from LDS
  // while the products of j run (two B register sets)
  auto phase = [&](int p, signed char (*cur)[SEC]) {
    const signed char *sec = &cur[2 * wb + p][0];
    // rows and columns stored as zero have zero slices: no zero test on the
    // scales
    double sc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) sc[j] = pow2((int)sec[4 * PLANE + 16 * j + l16] - 19);
    v4i B[2][4];
#pragma unroll
    for (int pl = 0; pl < 4; ++pl) B[0][pl] = ld16(sec + pl * PLANE + lane * 16);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (j < 3) {
#pragma unroll
        for (int pl = 0; pl < 4; ++pl)
          B[(j + 1) & 1][pl] = ld16(sec + pl * PLANE + ((j + 1) * 64 + lane) * 16);
      }
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        int h[4], o[4];
        tile_product(p ? Zr[g] : Fr[g], B[j & 1], h, o);
        const int er = p ? eZ[g] : eF[g];
#pragma unroll
        for (int r = 0; r < 4; ++r)
          v64[g][j][r] = fma(fma((double)o[r], 0x1p-14, (double)h[r]),
                             pow2((int)(signed char)(er >> (8 * r))) * sc[j], v64[g][j][r]);
      }
    }
  };
  // one listed batch: wait for its operands, load Z_I, start the next
  // batch's J copies, then the two phases with F_I of the next batch loaded
  // between them
  auto step = [&](int i, int n, signed char (*cur)[SEC], signed char (*nxt)[SEC]) {
    wait_vm0();
    __syncthreads();  // buffer cur complete, nxt free
    const bool act = active(i);
    if (act) load_I(z_section(__builtin_amdgcn_readfirstlane(ls[wa][i])), Zr, eZ);
    if (i + 1 < n) issue_J(i + 1, nxt);
    if (act) phase(0, cur);
    next_F(i + 1, n);
    if (act) phase(1, cur);
  };

  for (int64_t c0 = lo + split; c0 < hi; c0 += 512 * (int64_t)nsplit) {
    const int64_t b = c0 + (int64_t)nsplit * threadIdx.x;
    int sl[4] = {-1, -1, -1, -1};
    if (b < hi) {
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (tso[u] >= 0) sl[u] = ts0[tso[u] + b];
    }
    const bool act = (sl[0] >= 0 || sl[1] >= 0) && (sl[2] >= 0 || sl[3] >= 0);
    const unsigned long long m = __ballot(act);
    __syncthreads();  // the previous window is done with the list and buffers
    if (lane == 0) wcnt[wv] = __popcll(m);
    __syncthreads();
    int off = 0, n = 0;
    for (int w = 0; w < 8; ++w) {
      if (w < wv) off += wcnt[w];
      n += wcnt[w];
    }
    n = __builtin_amdgcn_readfirstlane(n);
    if (act) {
      const int i = off + __popcll(m & ((1ull << lane) - 1ull));
#pragma unroll
      for (int u = 0; u < 4; ++u) ls[u][i] = sl[u];
    }
    __syncthreads();
    if (n == 0) continue;
    fill(n, bufA);
    for (int i = 0; i < n; i += 2) {
      step(i, n, bufA, bufB);
      if (i + 1 < n) step(i + 1, n, bufB, bufA);
    }
  }
  if (pp < 0) return;
  double *out = vpart + ((int64_t)split * npairs + pp) * (FB * FB);
#pragma unroll
  for (int g = 0; g < 2; ++g)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int mu = 32 * wh + 16 * g + 4 * q + r, nu = 16 * j + l16;
        out[mu * FB + nu] += v64[g][j][r];
      }
}

// ---------------------------------------------------------------------------
void vkm_gather_p(hipStream_t s, int nbf, int nfb, const int *fperm, const double *P,
                  signed char *Ps) {
  hipLaunchKernelGGL(gather_p_i8, dim3((unsigned)(nfb * nfb)), dim3(256), 0, s, nbf, nfb, fperm,
                     P, Ps);
  NKX_HIP_CHECK(hipGetLastError());
}

int64_t mixed_p_bytes(int nfb) { return (int64_t)nfb * nfb * SEC; }

void vkm_pm_group(hipStream_t s, int64_t b0, int64_t b1, int64_t s0, const Plan &pl, int gga,
                  signed char *store) {
  if (b1 <= b0) return;
  hipLaunchKernelGGL(pm_group_i8, dim3((unsigned)(b1 - b0)), dim3(256), 0, s, s0, b0, pl.b_off.p,
                     pl.b_fbl.p, pl.tab(), gga, store);
  NKX_HIP_CHECK(hipGetLastError());
}

void vkm_colloc(hipStream_t s, int64_t s0, int64_t s1, const Plan &pl, int gga,
                signed char *store) {
  if (s1 <= s0) return;
  hipLaunchKernelGGL(colloc_tiles_i8, dim3((unsigned)(s1 - s0)), dim3(256), 0, s, s0,
                     pl.slot_b.p, pl.b_fbl.p, pl.funcs.p, pl.terms.p, pl.px.p, pl.py.p, pl.pz.p,
                     gga, store);
  NKX_HIP_CHECK(hipGetLastError());
}

template <int GBT>
static void launch_density_x2_i8(hipStream_t s, int64_t b0, int64_t b1, int64_t s0,
                                 const Plan &pl, const signed char *Ps, int gga,
                                 const signed char *store, double *part) {
  const int64_t ngroups = (b1 - b0 + GBT - 1) / GBT;
  const int64_t ngq = (ngroups + 7) / 8;
  const int umax = std::min(pl.nfb, GBT * pl.max_mol_nfb);
  const int nmg = (umax + 1) / 2;
  const int64_t nwg = ngq * 8 * nmg;
  NKX_REQUIRE(nwg < (int64_t)1 << 31, NKX_ERR_UNSUPPORTED, "density grid too large");
  hipLaunchKernelGGL((density_x2_i8<GBT, false>), dim3((unsigned)nwg), dim3(256), 0, s,
                     pl.tab(), pl.nfb, Ps, pl.b_off.p, pl.b_fbl.p, s0, b0, b1, nmg, 0, 0, gga,
                     store, part);
}

// absolute block pairs in XCD tiles of tpn pairs x tbn batches
static void launch_density_x2a_i8(hipStream_t s, int64_t b0, int64_t b1, int64_t s0,
                                  const Plan &pl, const signed char *Ps, int gga,
                                  const signed char *store, double *part, int tpn, int tbn) {
  const int npair = (pl.max_mol_nfb + 1) / 2;
  const int64_t ntiles = (int64_t)((npair + tpn - 1) / tpn) * ((b1 - b0 + tbn - 1) / tbn);
  const int64_t nwg = (ntiles + 7) / 8 * 8 * tpn * tbn;
  NKX_REQUIRE(nwg < (int64_t)1 << 31, NKX_ERR_UNSUPPORTED, "density grid too large");
  hipLaunchKernelGGL((density_x2_i8<1, true>), dim3((unsigned)nwg), dim3(256), 0, s, pl.tab(),
                     pl.nfb, Ps, pl.b_off.p, pl.b_fbl.p, s0, b0, b1, npair, tpn, tbn, gga,
                     store, part);
}

// shape: 65 / 66 one / two batches per workgroup, 71 / 72 one batch per
// workgroup on absolute block pairs dealt to the XCDs in tiles of 4 x 16 /
// 8 x 8 (pairs x batches); other values take 65
bool mixed_density_shape_ok(int shape) {
  return shape == 65 || shape == 66 || shape == 71 || shape == 72;
}

void vkm_density(hipStream_t s, int shape, int64_t b0, int64_t b1, int64_t s0, const Plan &pl,
                 const signed char *Ps, int gga, signed char *store, const XcList &xl,
                 double *coef, double *blockEN, double *part) {
  if (b1 <= b0) return;
  NKX_REQUIRE(pl.nfb <= MAXNFB, NKX_ERR_UNSUPPORTED,
              "more than 1024 function blocks (65536 functions)");
  if (!mixed_density_shape_ok(shape)) shape = 65;
  NKX_REQUIRE(part != nullptr, NKX_ERR_STATE, "density contraction without a partial buffer");
  switch (shape) {
    case 66: launch_density_x2_i8<2>(s, b0, b1, s0, pl, Ps, gga, store, part); break;
    case 71: launch_density_x2a_i8(s, b0, b1, s0, pl, Ps, gga, store, part, 4, 16); break;
    case 72: launch_density_x2a_i8(s, b0, b1, s0, pl, Ps, gga, store, part, 8, 8); break;
    default: launch_density_x2_i8<1>(s, b0, b1, s0, pl, Ps, gga, store, part); break;
  }
  NKX_HIP_CHECK(hipGetLastError());
}

// Second half of the density stage: per-batch functional from the slot
// partial sums, then the Z slices of every slot.
void vkm_density_finish(hipStream_t s, int shape, int64_t b0, int64_t b1, int64_t s0,
                        const Plan &pl, int gga, signed char *store, const XcList &xl,
                        double *coef, double *blockEN, const double *part) {
  if (b1 <= b0) return;
  (void)shape;
  hipLaunchKernelGGL(density_f_i8, dim3((unsigned)(b1 - b0)), dim3(256), 0, s, s0, b0,
                     pl.b_off.p, gga, part, pl.pw.p, xl, coef, blockEN, nullptr);
  NKX_HIP_CHECK(hipGetLastError());
  const int64_t s1 = pl.b_off_h[b1];
  if (s1 > s0)
    hipLaunchKernelGGL(z_slots_i8, dim3((unsigned)(s1 - s0)), dim3(256), 0, s, s0,
                       pl.slot_b.p, gga, coef, store, store + OFF_Z, slot_bytes_i8(gga));
  NKX_HIP_CHECK(hipGetLastError());
}

void vkm_density_finish_pol(hipStream_t s, int64_t b0, int64_t b1, int64_t s0, const Plan &pl,
                            int gga, signed char *store, signed char *zstore_b,
                            const XcList &xl, double *coef_a, double *coef_b,
                            double *blockEN, const double *part_a, const double *part_b,
                            double *epoint) {
  if (b1 <= b0) return;
  hipLaunchKernelGGL(density_fp_i8, dim3((unsigned)(b1 - b0)), dim3(256), 0, s, s0, b0,
                     pl.b_off.p, gga, part_a, part_b, pl.pw.p, xl, coef_a, coef_b, blockEN,
                     epoint);
  NKX_HIP_CHECK(hipGetLastError());
  if (!store) return;  // functional only (the gradient forms its own slices)
  const int64_t s1 = pl.b_off_h[b1];
  if (s1 > s0) {
    hipLaunchKernelGGL(z_slots_i8, dim3((unsigned)(s1 - s0)), dim3(256), 0, s, s0,
                       pl.slot_b.p, gga, coef_a, store, store + OFF_Z, slot_bytes_i8(gga));
    NKX_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(z_slots_i8, dim3((unsigned)(s1 - s0)), dim3(256), 0, s, s0,
                       pl.slot_b.p, gga, coef_b, store, zstore_b, (int64_t)SEC);
  }
  NKX_HIP_CHECK(hipGetLastError());
}

int64_t mixed_zsection_bytes() { return SEC; }

void vkm_density_functional(hipStream_t s, int64_t b0, int64_t b1, int64_t s0, const Plan &pl,
                            int gga, const XcList &xl, double *coef, double *blockEN,
                            const double *part, double *epoint) {
  if (b1 <= b0) return;
  hipLaunchKernelGGL(density_f_i8, dim3((unsigned)(b1 - b0)), dim3(256), 0, s, s0, b0,
                     pl.b_off.p, gga, part, pl.pw.p, xl, coef, blockEN, epoint);
  NKX_HIP_CHECK(hipGetLastError());
}

void vkm_vgemm(hipStream_t s, int64_t b0, int64_t b1, int64_t s0, const Plan &pl, int nsplit,
               int gga, const signed char *store, double *vpart, const signed char *zstore_b) {
  if (b1 <= b0 || pl.nspairs == 0) return;
  // Z sections: each slot's own, or (second spin) SEC-strided in zstore_b
  const signed char *zbase = zstore_b ? zstore_b : store + OFF_Z;
  const int64_t zstride = zstore_b ? (int64_t)SEC : slot_bytes_i8(gga);
  hipLaunchKernelGGL(vgemm_i8, dim3((unsigned)(pl.nspairs * nsplit)), dim3(512), 0, s,
                     pl.nspairs, nsplit, reinterpret_cast<const int4 *>(pl.sp_blk.p),
                     reinterpret_cast<const int4 *>(pl.sp_pidx.p), pl.npairs, pl.tile_slot.p,
                     pl.tab(), b0, b1, s0, gga, store, zbase, zstride, vpart);
  NKX_HIP_CHECK(hipGetLastError());
}

}  // namespace nkx
