// Plan construction and orchestration of the batched V_xc build.
//
// Plan (once per geometry / basis / screening setting):
//   * points sorted along a Morton curve and cut into batches of 64, the last
//     one padded with zero-weight copies of its final point;
//   * functions sorted by (screening-radius class, Morton code of the centre)
//     and cut into blocks of 64, padded with zero functions, so that a block
//     holds functions of similar extent at nearby centres;
//   * a (block, batch) pair is live when some function of the block reaches
//     the batch's bounding box within its screening radius (the radius where
//     N r^{n-1} e^{-zeta r} drops to the screening tolerance, the
//     compute_cutoff rule of stobasis.hpp:90-173);
//   * live pairs are numbered batch-major (CSR), which is also their slot
//     number in the tile store.
// Build (once per SCF iteration): see kernels_vxc.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <numeric>
#include <string>

#include "nkx_internal.hpp"
#include "vxc_kernels.hpp"

namespace nkx {

namespace {

uint64_t spread21(uint64_t v) {
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffull;
  v = (v | v << 16) & 0x1f0000ff0000ffull;
  v = (v | v << 8) & 0x100f00f00f00f00full;
  v = (v | v << 4) & 0x10c30c30c30c30c3ull;
  v = (v | v << 2) & 0x1249249249249249ull;
  return v;
}

struct Box {
  double lo[3], hi[3];
};

uint64_t morton(const double *p, const Box &bb) {
  uint64_t code = 0;
  for (int d = 0; d < 3; ++d) {
    const double ext = std::max(bb.hi[d] - bb.lo[d], 1e-12);
    double t = (p[d] - bb.lo[d]) / ext;
    t = std::min(std::max(t, 0.0), 1.0);
    code |= spread21((uint64_t)(t * 2097151.0)) << d;
  }
  return code;
}

double box_dist2(const Box &b, double x, double y, double z) {
  const double c[3] = {x, y, z};
  double d2 = 0.0;
  for (int d = 0; d < 3; ++d) {
    double e = 0.0;
    if (c[d] < b.lo[d]) e = b.lo[d] - c[d];
    else if (c[d] > b.hi[d]) e = c[d] - b.hi[d];
    d2 += e * e;
  }
  return d2;
}

constexpr int FB = 64;
constexpr int PT = 64;

}  // namespace

namespace {

// Plan of one molecule in local numbering (blocks, batches, slots from 0).
struct PlanPart {
  int nbf = 0, nfb = 0;
  int64_t npts = 0, nb = 0;
  std::vector<int> fperm;  // sorted -> original function
  std::vector<FuncRec> fr;
  std::vector<HarmTerm> terms;
  std::vector<double> px, py, pz, pw;
  std::vector<int> pidx;
  std::vector<int> b_off, b_fbl, slot_b, tile_slot;
  std::vector<int> pair_I, pair_J;
  double nb2g = 0, nbg = 0;
};

// Points P (3 npts) and weights W of one molecule, its basis B.
PlanPart make_part(const std::vector<double> &P, const std::vector<double> &W,
                   const BasisHost &B, double tol) {
  PlanPart pt;
  const int64_t npts = (int64_t)W.size();
  pt.npts = npts;
  // ---- points ------------------------------------------------------------
  Box bb{{1e300, 1e300, 1e300}, {-1e300, -1e300, -1e300}};
  for (int64_t g = 0; g < npts; ++g)
    for (int d = 0; d < 3; ++d) {
      bb.lo[d] = std::min(bb.lo[d], P[3 * g + d]);
      bb.hi[d] = std::max(bb.hi[d], P[3 * g + d]);
    }
  std::vector<std::pair<uint64_t, int64_t>> key(npts);
  for (int64_t g = 0; g < npts; ++g) key[g] = {morton(&P[3 * g], bb), g};
  std::sort(key.begin(), key.end());
  const int64_t nb = (npts + PT - 1) / PT;
  const int64_t npad = nb * PT;
  pt.nb = nb;
  pt.px.resize(npad);
  pt.py.resize(npad);
  pt.pz.resize(npad);
  pt.pw.resize(npad);
  pt.pidx.resize(npad);
  std::vector<Box> boxes(nb);
  for (int64_t b = 0; b < nb; ++b) {
    Box &x = boxes[b];
    for (int d = 0; d < 3; ++d) {
      x.lo[d] = 1e300;
      x.hi[d] = -1e300;
    }
    for (int i = 0; i < PT; ++i) {
      const int64_t sp = b * PT + i;
      const int64_t src = sp < npts ? key[sp].second : key[npts - 1].second;
      pt.px[sp] = P[3 * src];
      pt.py[sp] = P[3 * src + 1];
      pt.pz[sp] = P[3 * src + 2];
      pt.pw[sp] = sp < npts ? W[src] : 0.0;
      pt.pidx[sp] = sp < npts ? (int)src : -1;
      for (int d = 0; d < 3; ++d) {
        x.lo[d] = std::min(x.lo[d], P[3 * src + d]);
        x.hi[d] = std::max(x.hi[d], P[3 * src + d]);
      }
    }
  }

  // ---- functions ---------------------------------------------------------
  const int nbf = B.nbf;
  const int nfb = (nbf + FB - 1) / FB;
  const int nbf_pad = nfb * FB;
  pt.nbf = nbf;
  pt.nfb = nfb;
  std::vector<double> rad(nbf);
  Box cb{{1e300, 1e300, 1e300}, {-1e300, -1e300, -1e300}};
  for (int i = 0; i < nbf; ++i) {
    double nrm, cut;
    sto_norm_cutoff(B.n[i], B.zeta[i], tol, &nrm, &cut);
    // Same radial bound with the stored normalisation (callers may scale it).
    if (B.norm[i] != nrm && B.norm[i] > 0) {
      double r = cut;
      for (int it = 0; it < 60; ++it) {
        const double v = std::log(B.norm[i]) + (B.n[i] - 1) * std::log(r) - B.zeta[i] * r -
                         std::log(tol);
        if (v <= 0) break;
        r *= 1.25;
      }
      cut = r;
    }
    rad[i] = cut;
    const double ctr[3] = {B.cx[i], B.cy[i], B.cz[i]};
    for (int d = 0; d < 3; ++d) {
      cb.lo[d] = std::min(cb.lo[d], ctr[d]);
      cb.hi[d] = std::max(cb.hi[d], ctr[d]);
    }
  }
  std::vector<std::tuple<int, uint64_t, int>> fkey(nbf);
  for (int i = 0; i < nbf; ++i) {
    const double ctr[3] = {B.cx[i], B.cy[i], B.cz[i]};
    const int cls = (int)std::floor(2.0 * std::log2(std::max(rad[i], 0.25)));
    fkey[i] = std::make_tuple(cls, morton(ctr, cb), i);
  }
  std::sort(fkey.begin(), fkey.end());
  pt.fperm.resize(nbf);
  for (int i = 0; i < nbf; ++i) pt.fperm[i] = std::get<2>(fkey[i]);

  pt.fr.resize(nbf_pad);
  for (int sp = 0; sp < nbf_pad; ++sp) {
    FuncRec r{};
    if (sp < nbf) {
      const int i = pt.fperm[sp];
      const auto &t = solid_harmonic_terms(B.l[i], B.m[i]);
      r.cx = B.cx[i];
      r.cy = B.cy[i];
      r.cz = B.cz[i];
      r.zeta = B.zeta[i];
      r.norm = B.norm[i];
      r.k = B.n[i] - B.l[i] - 1;
      r.term_off = (int)pt.terms.size();
      r.term_cnt = (int)t.size();
      pt.terms.insert(pt.terms.end(), t.begin(), t.end());
    } else {
      r.zeta = 1.0;
      r.norm = 0.0;
      r.term_off = 0;
      r.term_cnt = 0;
    }
    pt.fr[sp] = r;
  }

  // ---- screening: live (block, batch) pairs -------------------------------
  pt.b_off.assign(nb + 1, 0);
  pt.b_fbl.reserve((size_t)nb * 4);
  for (int64_t b = 0; b < nb; ++b) {
    pt.b_off[b] = (int)pt.b_fbl.size();
    for (int f = 0; f < nfb; ++f) {
      bool live = false;
      for (int k = f * FB; k < std::min(nbf, (f + 1) * FB) && !live; ++k) {
        const int i = pt.fperm[k];
        live = box_dist2(boxes[b], B.cx[i], B.cy[i], B.cz[i]) < rad[i] * rad[i];
      }
      if (live) {
        pt.b_fbl.push_back(f);
        pt.slot_b.push_back((int)b);
      }
    }
    NKX_REQUIRE(pt.b_fbl.size() < (size_t)INT32_MAX, NKX_ERR_UNSUPPORTED,
                "too many live (block, batch) pairs for 32-bit slot numbers");
  }
  pt.b_off[nb] = (int)pt.b_fbl.size();
  // algorithmic work: true functions of the live blocks, true points
  for (int64_t b = 0; b < nb; ++b) {
    double nfun = 0;
    for (int k = pt.b_off[b]; k < pt.b_off[b + 1]; ++k)
      nfun += std::min(nbf, (pt.b_fbl[k] + 1) * FB) - pt.b_fbl[k] * FB;
    const double g = (double)std::min<int64_t>(PT, npts - b * PT);
    pt.nb2g += nfun * nfun * g;
    pt.nbg += nfun * g;
  }
  pt.tile_slot.assign((size_t)nfb * nb, -1);
  std::vector<char> pair_live((size_t)nfb * nfb, 0);
  for (int64_t b = 0; b < nb; ++b) {
    for (int k = pt.b_off[b]; k < pt.b_off[b + 1]; ++k) {
      pt.tile_slot[(size_t)pt.b_fbl[k] * nb + b] = k;
      for (int k2 = k; k2 < pt.b_off[b + 1]; ++k2)
        pair_live[(size_t)pt.b_fbl[k] * nfb + pt.b_fbl[k2]] = 1;
    }
  }
  for (int I = 0; I < nfb; ++I)
    for (int J = I; J < nfb; ++J)
      if (pair_live[(size_t)I * nfb + J]) {
        pt.pair_I.push_back(I);
        pt.pair_J.push_back(J);
      }
  return pt;
}

// Concatenate the parts into global numbering and upload.
void upload_plan(Plan &pl, const std::vector<PlanPart> &parts, hipStream_t s) {
  const int nmol = (int)parts.size();
  pl.nmol = nmol;
  pl.mol_nbf_h.resize(nmol);
  pl.mol_ld_h.resize(nmol);
  pl.mol_nfb_h.resize(nmol);
  pl.mol_blk0_h.resize(nmol);
  pl.mol_b0_h.resize(nmol);
  pl.mol_nb_h.resize(nmol);
  pl.mol_npts_h.resize(nmol);
  pl.mol_P_off_h.resize(nmol);
  pl.mol_Ps_off_h.resize(nmol);
  std::vector<int64_t> ts_off(nmol);
  int nfb = 0;
  int64_t nb = 0, nslots = 0, ts = 0, poff = 0, psoff = 0, npts = 0;
  double nb2g = 0, nbg = 0;
  for (int m = 0; m < nmol; ++m) {
    const PlanPart &pt = parts[m];
    pl.mol_nbf_h[m] = pt.nbf;
    pl.mol_ld_h[m] = pt.nfb * FB;
    pl.mol_nfb_h[m] = pt.nfb;
    pl.mol_blk0_h[m] = nfb;
    pl.mol_b0_h[m] = nb;
    pl.mol_nb_h[m] = pt.nb;
    pl.mol_npts_h[m] = pt.npts;
    pl.mol_P_off_h[m] = poff;
    pl.mol_Ps_off_h[m] = psoff;
    ts_off[m] = ts;
    nfb += pt.nfb;
    nb += pt.nb;
    nslots += pt.b_off[pt.nb];
    ts += (int64_t)pt.nfb * pt.nb;
    poff += (int64_t)pt.nfb * FB * pt.nfb * FB;
    psoff += mixed_p_bytes(pt.nfb);
    npts += pt.npts;
    nb2g += pt.nb2g;
    nbg += pt.nbg;
  }
  NKX_REQUIRE(nslots < (int64_t)INT32_MAX, NKX_ERR_UNSUPPORTED,
              "too many live (block, batch) pairs for 32-bit slot numbers");
  const int nbf_pad = nfb * FB;
  const int64_t npad = nb * PT;
  std::vector<int> fperm(nbf_pad, 0), pidx(npad), b_off(nb + 1), b_fbl, slot_b, blk_mol(nfb),
      blk_local(nfb), tile_slot(std::max<int64_t>(ts, 1), -1);
  std::vector<FuncRec> fr(nbf_pad);
  std::vector<HarmTerm> terms;
  std::vector<double> px(npad), py(npad), pz(npad), pw(npad);
  pl.pair_I_h.clear();
  pl.pair_J_h.clear();
  b_fbl.reserve(nslots);
  slot_b.reserve(nslots);
  int nbf_total = 0;
  for (int m = 0; m < nmol; ++m) {
    const PlanPart &pt = parts[m];
    const int blk0 = pl.mol_blk0_h[m];
    const int64_t b0 = pl.mol_b0_h[m];
    const int s0 = (int)b_fbl.size();
    const int toff = (int)terms.size();
    for (int f = 0; f < pt.nfb; ++f) {
      blk_mol[blk0 + f] = m;
      blk_local[blk0 + f] = f;
    }
    for (int i = 0; i < pt.nbf; ++i) fperm[blk0 * FB + i] = pt.fperm[i];
    for (int i = 0; i < pt.nfb * FB; ++i) {
      FuncRec r = pt.fr[i];
      r.term_off += toff;
      fr[blk0 * FB + i] = r;
    }
    terms.insert(terms.end(), pt.terms.begin(), pt.terms.end());
    for (int64_t i = 0; i < pt.nb * PT; ++i) {
      px[b0 * PT + i] = pt.px[i];
      py[b0 * PT + i] = pt.py[i];
      pz[b0 * PT + i] = pt.pz[i];
      pw[b0 * PT + i] = pt.pw[i];
      pidx[b0 * PT + i] = pt.pidx[i];
    }
    for (int64_t b = 0; b < pt.nb; ++b) b_off[b0 + b] = s0 + pt.b_off[b];
    for (size_t k = 0; k < pt.b_fbl.size(); ++k) {
      b_fbl.push_back(blk0 + pt.b_fbl[k]);
      slot_b.push_back((int)(b0 + pt.slot_b[k]));
    }
    for (size_t k = 0; k < pt.tile_slot.size(); ++k)
      if (pt.tile_slot[k] >= 0) tile_slot[ts_off[m] + k] = s0 + pt.tile_slot[k];
    for (size_t k = 0; k < pt.pair_I.size(); ++k) {
      pl.pair_I_h.push_back(blk0 + pt.pair_I[k]);
      pl.pair_J_h.push_back(blk0 + pt.pair_J[k]);
    }
    nbf_total += pt.nbf;
  }
  b_off[nb] = (int)b_fbl.size();
  // super-pairs: blocks (2a, 2a + 1) x (2c, 2c + 1), a <= c, per molecule,
  // kept when one of the four tiles is a kept pair
  std::vector<int> spb, spp;
  {
    size_t pbase = 0;
    for (int m = 0; m < nmol; ++m) {
      const PlanPart &pt = parts[m];
      const int blk0 = pl.mol_blk0_h[m], nf = pt.nfb;
      std::vector<int> pid((size_t)nf * nf, -1);
      for (size_t k = 0; k < pt.pair_I.size(); ++k)
        pid[(size_t)pt.pair_I[k] * nf + pt.pair_J[k]] = (int)(pbase + k);
      auto at = [&](int I, int J) {
        return (I < 0 || J < 0 || I > J) ? -1 : pid[(size_t)I * nf + J];
      };
      const int ns = (nf + 1) / 2;
      for (int a = 0; a < ns; ++a)
        for (int c = a; c < ns; ++c) {
          const int I0 = 2 * a, I1 = 2 * a + 1 < nf ? 2 * a + 1 : -1;
          const int J0 = 2 * c, J1 = 2 * c + 1 < nf ? 2 * c + 1 : -1;
          const int px[4] = {at(I0, J0), at(I0, J1), a == c ? -1 : at(I1, J0), at(I1, J1)};
          if (px[0] < 0 && px[1] < 0 && px[2] < 0 && px[3] < 0) continue;
          spb.insert(spb.end(), {blk0 + I0, I1 >= 0 ? blk0 + I1 : -1, blk0 + J0,
                                 J1 >= 0 ? blk0 + J1 : -1});
          spp.insert(spp.end(), px, px + 4);
        }
      pbase += pt.pair_I.size();
    }
  }
  if (terms.empty()) terms.push_back(HarmTerm{0, 0, 0, 0, 0});
  if (b_fbl.empty()) {  // keep device buffers non-empty
    b_fbl.push_back(0);
    slot_b.push_back(0);
  }

  pl.nbf = nmol == 1 ? parts[0].nbf : nbf_total;
  pl.nbf_pad = nbf_pad;
  pl.nfb = nfb;
  pl.fb = FB;
  pl.npts = npts;
  pl.npts_pad = npad;
  pl.nb = nb;
  pl.b_off_h = b_off;
  pl.nslots = nslots;
  pl.pairs_kept = nslots;
  pl.npairs = (int)pl.pair_I_h.size();
  pl.work_nb2g = nb2g;
  pl.work_nbg = nbg;
  pl.P_total = poff;
  pl.fperm = fperm;
  pl.fperm_d.upload(fperm.data(), fperm.size(), s);
  pl.funcs.upload(fr.data(), fr.size(), s);
  pl.terms.upload(terms.data(), terms.size(), s);
  pl.px.upload(px.data(), npad, s);
  pl.py.upload(py.data(), npad, s);
  pl.pz.upload(pz.data(), npad, s);
  pl.pw.upload(pw.data(), npad, s);
  pl.pidx.upload(pidx.data(), npad, s);
  pl.b_off.upload(b_off.data(), b_off.size(), s);
  pl.b_fbl.upload(b_fbl.data(), b_fbl.size(), s);
  pl.slot_b.upload(slot_b.data(), slot_b.size(), s);
  pl.tile_slot.upload(tile_slot.data(), tile_slot.size(), s);
  if (pl.npairs) {
    pl.pair_I.upload(pl.pair_I_h.data(), pl.npairs, s);
    pl.pair_J.upload(pl.pair_J_h.data(), pl.npairs, s);
  }
  pl.nspairs = (int)(spb.size() / 4);
  pl.max_mol_nfb = 0;
  for (int m = 0; m < nmol; ++m) pl.max_mol_nfb = std::max(pl.max_mol_nfb, parts[m].nfb);
  if (pl.nspairs) {
    pl.sp_blk.upload(spb.data(), spb.size(), s);
    pl.sp_pidx.upload(spp.data(), spp.size(), s);
  }
  pl.blk_mol.upload(blk_mol.data(), nfb, s);
  pl.blk_local.upload(blk_local.data(), nfb, s);
  pl.mol_ld.upload(pl.mol_ld_h.data(), nmol, s);
  pl.mol_nbf.upload(pl.mol_nbf_h.data(), nmol, s);
  pl.mol_nfb.upload(pl.mol_nfb_h.data(), nmol, s);
  pl.mol_b0.upload(pl.mol_b0_h.data(), nmol, s);
  pl.mol_nb.upload(pl.mol_nb_h.data(), nmol, s);
  pl.mol_ts_off.upload(ts_off.data(), nmol, s);
  pl.mol_P_off.upload(pl.mol_P_off_h.data(), nmol, s);
  pl.mol_Ps_off.upload(pl.mol_Ps_off_h.data(), nmol, s);
  NKX_HIP_CHECK(hipStreamSynchronize(s));
  pl.pt = PT;
}

}  // namespace

void ensure_plan(Ctx &c) {
  Plan &pl = c.plan;
  NKX_REQUIRE(c.grid.npts > 0, NKX_ERR_STATE, "no grid: call nkx_make_flat_grid or nkx_set_grid");
  NKX_REQUIRE(c.basis.nbf > 0, NKX_ERR_STATE, "no basis: call nkx_set_basis");
  if (pl.grid_version == c.grid.version && pl.basis_version == c.basis_version &&
      pl.screen_tol == c.opt.screen_tol && pl.pt == PT)
    return;
  hipStream_t s = c.stream;
  const int64_t npts = c.grid.npts;
  std::vector<double> P(3 * npts), W(npts);
  NKX_HIP_CHECK(hipMemcpyAsync(P.data(), c.grid.pts.p, 3 * npts * sizeof(double),
                               hipMemcpyDeviceToHost, s));
  NKX_HIP_CHECK(hipMemcpyAsync(W.data(), c.grid.w.p, npts * sizeof(double),
                               hipMemcpyDeviceToHost, s));
  NKX_HIP_CHECK(hipStreamSynchronize(s));
  std::vector<PlanPart> parts;
  parts.push_back(make_part(P, W, c.basis, c.opt.screen_tol));
  upload_plan(pl, parts, s);
  pl.grid_version = c.grid.version;
  pl.basis_version = c.basis_version;
  pl.screen_tol = c.opt.screen_tol;
}

// Plan of a molecule set: molecule m has points mol_pts[m] (3 npts), weights
// mol_w[m] and basis bases[m].
void build_molset_plan(Ctx &c, const std::vector<std::vector<double>> &mol_pts,
                       const std::vector<std::vector<double>> &mol_w,
                       const std::vector<BasisHost> &bases) {
  std::vector<PlanPart> parts;
  for (size_t m = 0; m < bases.size(); ++m)
    parts.push_back(make_part(mol_pts[m], mol_w[m], bases[m], c.opt.screen_tol));
  upload_plan(c.mplan, parts, c.stream);
  c.mplan.screen_tol = c.opt.screen_tol;
  ++c.mplan_version;
}

namespace {

// Number of slots per chunk so that the chunk's tiles fit the budget.
// cap: the ctx's tile-store budget in bytes, 0 for min(48 GB, 40 % of free).
int64_t chunk_slot_budget(size_t slot_bytes, size_t cap) {
  size_t budget = cap;
  if (budget == 0) {
    size_t freeb = 0, total = 0;
    NKX_HIP_CHECK(hipMemGetInfo(&freeb, &total));
    budget = std::min<size_t>((size_t)48 << 30, freeb / 5 * 2);
  }
  budget = std::max<size_t>(budget, (size_t)64 << 20);
  return (int64_t)(budget / slot_bytes);
}

// Split count of the V contraction over batches: about 8192 workgroups of
// the FP64 kernel's pair tiles or of the int8 kernel's super-pair tiles
// (a few per CU at a time, the rest balancing the tail), or the autotuned
// value.
int choose_nsplit(const Plan &pl, int tuned = 0, bool super = false) {
  const int64_t units = super ? pl.nspairs : pl.npairs;
  if (units == 0) return 1;
  // about 8192 workgroups, and (automatic choice) at least 32 batches per
  // split: on small molecules one batch per workgroup is mostly set-up
  const int64_t target = 8192;
  int64_t ns = tuned > 0 ? tuned : (target + units - 1) / units;
  if (tuned <= 0) ns = std::min<int64_t>(ns, pl.nb / 32);
  ns = std::max<int64_t>(1, std::min<int64_t>(ns, pl.nb));
  return (int)ns;
}

template <typename F>
void for_each_chunk(const Plan &pl, int64_t max_slots, F &&fn) {
  int64_t b0 = 0;
  while (b0 < pl.nb) {
    int64_t b1 = b0 + 1;
    // grow while the slot range fits (at least one batch per chunk)
    int64_t lo = b0 + 1, hi = pl.nb;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) / 2;
      if (pl.b_off_h[mid] - pl.b_off_h[b0] <= max_slots) lo = mid;
      else hi = mid - 1;
    }
    b1 = lo;
    fn(b0, b1, (int64_t)pl.b_off_h[b0], (int64_t)pl.b_off_h[b1]);
    b0 = b1;
  }
}

}  // namespace

namespace {
// Device table of output pointers (kept alive in the workspace until the
// build synchronises).
double *const *upload_ptrs(Ctx &c, const std::vector<double *> &ptrs) {
  c.ws.vptr_h = ptrs;
  c.ws.vptr.upload(c.ws.vptr_h.data(), c.ws.vptr_h.size(), c.stream);
  return c.ws.vptr.p;
}
}  // namespace

// Sizes and workspace of one build on plan pl (allocations happen here, so a
// captured build performs none).
BuildSetup prepare_build(Workspace &ws, const Plan &pl, int nspin, bool gga, int precision,
                         int tuned_split, int tuned_group, size_t mem_cap) {
  BuildSetup b;
  b.nspin = nspin;
  // density workgroup shape: tuned, else 2 batches per workgroup (FP64) /
  // the split form with 32 rows per wave and one batch (int8 form)
  b.group = tuned_group > 0 ? tuned_group : (precision == NKX_PREC_MIXED ? 65 : 2);
  b.gga = gga;
  const int nmol = pl.nmol;
  b.narr = (gga ? 4 : 1) + nspin;
  b.nsplit = choose_nsplit(pl, tuned_split, precision == NKX_PREC_MIXED);
  b.bf16 = precision == NKX_PREC_MIXED_BF16;
  b.mixed = precision == NKX_PREC_MIXED || b.bf16;
  NKX_REQUIRE(!b.bf16 || nspin == 1, NKX_ERR_UNSUPPORTED,
              "the spin-polarised build is available in the FP64 and int8 forms");
  NKX_REQUIRE(!b.bf16 || nmol == 1, NKX_ERR_UNSUPPORTED,
              "the split-bf16 form builds one molecule at a time");
  const size_t slot_bytes = b.bf16    ? (size_t)bf16_slot_bytes(gga ? 1 : 0)
                            : b.mixed ? (size_t)mixed_slot_bytes(gga ? 1 : 0)
                                      : (size_t)b.narr * FB * PT * sizeof(double);
  // per slot beyond the tile store: density partials per spin and the second
  // spin's Z section (int8 form)
  const size_t slot_extra =
      b.mixed && !b.bf16
          ? (size_t)nspin * 4 * PT * sizeof(double) +
                (nspin == 2 ? (size_t)mixed_zsection_bytes() : 0)
          : 0;
  int64_t max_slots = chunk_slot_budget(slot_bytes + slot_extra, mem_cap);
  // at least the largest single batch
  for (int64_t bb = 0; bb < pl.nb; ++bb)
    max_slots = std::max<int64_t>(max_slots, pl.b_off_h[bb + 1] - pl.b_off_h[bb]);
  b.slots_alloc = std::min<int64_t>(max_slots, std::max<int64_t>(pl.nslots, 1));
  if (b.mixed) {
    ws.mstore.reserve((size_t)b.slots_alloc * slot_bytes);
    if (!b.bf16) ws.dpart.reserve((size_t)nspin * b.slots_alloc * 4 * PT);
    if (!b.bf16 && nspin == 2)
      ws.zstore_b.reserve((size_t)b.slots_alloc * mixed_zsection_bytes());
    int64_t psb = 0;
    for (int m = 0; m < nmol; ++m) psb += mixed_p_bytes(pl.mol_nfb_h[m]);
    b.psb = psb;
    ws.Ps.reserve(b.bf16 ? (size_t)6 * pl.nbf_pad * pl.nbf_pad : (size_t)nspin * psb);
  } else {
    ws.tiles64.reserve((size_t)b.slots_alloc * b.narr * FB * PT);
    ws.Pd.reserve((size_t)nspin * pl.P_total);
  }
  ws.coef.reserve((size_t)4 * nspin * pl.npts_pad);
  ws.blockEN.reserve((size_t)3 * pl.nb);
  b.vsz = (size_t)b.nsplit * std::max(pl.npairs, 1) * FB * FB;
  ws.vpart.reserve(nspin * b.vsz);
  ws.red.reserve((size_t)3 * pl.nmol);
  return b;
}

// Every launch of one build, stream-ordered, ending with the copy of the
// (E, N_alpha, N_beta) triples of all molecules into en_dst (host).  No
// allocation, no memset, no synchronisation: the sequence can be captured.
void enqueue_build(Ctx &c, Workspace &ws, const Plan &pl, const BuildSetup &b,
                   const double *const *P, const XcList &xl, double *const *V,
                   double *const *Vtab, double *en_dst) {
  hipStream_t s = c.stream;
  const int nmol = pl.nmol, nspin = b.nspin, narr = b.narr, nsplit = b.nsplit;
  const bool bf16 = b.bf16, mixed = b.mixed;
  const size_t vsz = b.vsz;
  Timing &tm = c.timing;
  const int t_build = tm.begin(kStageBuild, s);
  int t = tm.begin(kStageFinal, s);
  uint16_t *bstore = reinterpret_cast<uint16_t *>(ws.mstore.p);
  uint16_t *bPs = reinterpret_cast<uint16_t *>(ws.Ps.p);
  for (int sp = 0; sp < nspin; ++sp)
    for (int m = 0; m < nmol; ++m) {
      const double *Pm = P[(size_t)sp * nmol + m];
      const int nbf = pl.mol_nbf_h[m];
      const int *fp = pl.fperm_d.p + (size_t)pl.mol_blk0_h[m] * FB;
      if (bf16)
        vkb_gather_p(s, nbf, pl.nfb, fp, Pm, bPs);
      else if (mixed)
        vkm_gather_p(s, nbf, pl.mol_nfb_h[m], fp, Pm,
                     ws.Ps.p + (size_t)sp * b.psb + pl.mol_Ps_off_h[m]);
      else
        vk_gather_p(s, nbf, pl.mol_ld_h[m], fp, Pm,
                    ws.Pd.p + (size_t)sp * pl.P_total + pl.mol_P_off_h[m]);
      vk_fill0(s, V[(size_t)sp * nmol + m], (int64_t)nbf * nbf);
    }
  vk_fill0(s, ws.vpart.p, (int64_t)(nspin * vsz));
  tm.end(t, s);
  for_each_chunk(pl, b.slots_alloc, [&](int64_t b0, int64_t b1, int64_t s0, int64_t s1) {
    const int ig = b.gga ? 1 : 0;
    int tt = tm.begin(kStageColloc, s);
    if (bf16)
      vkb_colloc(s, s0, s1, pl, ig, bstore);
    else if (mixed) {
      vkm_colloc(s, s0, s1, pl, ig, ws.mstore.p);
      vkm_pm_group(s, b0, b1, s0, pl, ig, ws.mstore.p);
    } else
      vk_colloc(s, s0, s1, pl, narr, ig, ws.tiles64.p);
    tm.end(tt, s);
    tt = tm.begin(kStageDensity, s);
    if (bf16)
      vkb_density(s, 6, b0, b1, s0, pl, bPs, ig, bstore, xl, ws.coef.p, ws.blockEN.p);
    else if (mixed)
      for (int sp = 0; sp < nspin; ++sp)  // one contraction per spin
        vkm_density(s, b.group, b0, b1, s0, pl, ws.Ps.p + (size_t)sp * b.psb, ig,
                    ws.mstore.p, xl, ws.coef.p, ws.blockEN.p,
                    ws.dpart.p + (size_t)sp * b.slots_alloc * 4 * PT);
    else
      vk_density(s, nspin, b.group, b0, b1, s0, pl, ws.Pd.p, narr, ig, ws.tiles64.p, xl,
                 ws.coef.p, ws.blockEN.p, nullptr);
    tm.end(tt, s);
    if (mixed && !bf16) {
      tt = tm.begin(kStageZ, s);
      if (nspin == 2)
        vkm_density_finish_pol(s, b0, b1, s0, pl, ig, ws.mstore.p, ws.zstore_b.p, xl,
                               ws.coef.p, ws.coef.p + (size_t)4 * pl.npts_pad, ws.blockEN.p,
                               ws.dpart.p, ws.dpart.p + (size_t)b.slots_alloc * 4 * PT);
      else
        vkm_density_finish(s, b.group, b0, b1, s0, pl, ig, ws.mstore.p, xl, ws.coef.p,
                           ws.blockEN.p, ws.dpart.p);
      tm.end(tt, s);
    }
    tt = tm.begin(kStageVgemm, s);
    if (bf16)
      vkb_vgemm(s, 6, b0, b1, s0, pl, nsplit, ig, bstore, ws.vpart.p);
    else if (mixed) {
      vkm_vgemm(s, b0, b1, s0, pl, nsplit, ig, ws.mstore.p, ws.vpart.p);
      if (nspin == 2)
        vkm_vgemm(s, b0, b1, s0, pl, nsplit, ig, ws.mstore.p, ws.vpart.p + vsz, ws.zstore_b.p);
    }
    else
      vk_vgemm(s, b0, b1, s0, pl, nsplit, narr, nspin, ws.tiles64.p, ws.vpart.p);
    tm.end(tt, s);
  });
  t = tm.begin(kStageFinal, s);
  for (int sp = 0; sp < nspin; ++sp)
    vk_vreduce(s, pl, nsplit, ws.vpart.p + sp * vsz, Vtab + (size_t)sp * nmol, pl.mol_nbf.p);
  vk_en_reduce(s, pl, ws.blockEN.p, ws.red.p);
  tm.end(t, s);
  tm.end(t_build, s);
  NKX_HIP_CHECK(hipMemcpyAsync(en_dst, ws.red.p, (size_t)3 * nmol * sizeof(double),
                               hipMemcpyDeviceToHost, s));
}

// One V_xc build on plan pl.  Molecule m, spin sp: density matrix
// P[sp * nmol + m] and potential V[sp * nmol + m] (device, nbf_m x nbf_m,
// row-major); E[m] and N[m * nspin + sp] on return.  nspin = 2 is FP64 only.
void build_on_plan(Ctx &c, Plan &pl, int nspin, const double *const *P, const XcList &xl,
                   bool gga, int precision, double *E, double *N, double *const *V) {
  const int nmol = pl.nmol;
  const BuildSetup b =
      prepare_build(c.ws, pl, nspin, gga, precision, c.opt.vgemm_split, c.opt.density_group,
                    c.opt.mem_budget);
  double *const *Vtab = upload_ptrs(c, std::vector<double *>(V, V + (size_t)nspin * nmol));
  std::vector<double> en((size_t)3 * nmol);
  enqueue_build(c, c.ws, pl, b, P, xl, V, Vtab, en.data());
  NKX_HIP_CHECK(hipStreamSynchronize(c.stream));
  c.timing.collect();
  c.work.flops_density = 2.0 * nspin * pl.work_nb2g;
  c.work.flops_v = 2.0 * nspin * pl.work_nb2g;
  c.work.colloc_elems = pl.work_nbg;
  for (int m = 0; m < nmol; ++m) {
    E[m] = en[3 * m];
    for (int sp = 0; sp < nspin; ++sp) N[(size_t)m * nspin + sp] = en[3 * m + 1 + sp];
  }
}

// ---- autotuner ------------------------------------------------------------
// Times one build per candidate (after one untimed build): first the batches
// per density workgroup (1, 2, 4), then the split count of the V contraction
// (1/4 .. 4 x the automatic choice), and keeps the fastest in the ctx options.
namespace {
// Times each candidate on plan pl (build() runs one build with the ctx
// options): first the density shape, then the V split count; keeps the
// fastest pair in the ctx options and returns its build time.
double autotune_on(Ctx &c, const Plan &pl, int precision, const std::function<void()> &build,
                   int *best_split, int *best_group) {
  const int base = choose_nsplit(pl, 0, precision == NKX_PREC_MIXED);
  std::vector<int> splits;
  for (int k : {base / 4, base / 2, base, 2 * base, 4 * base}) {
    const int v = std::max(1, std::min<int>(k, (int)std::max<int64_t>(pl.nb, 1)));
    if (std::find(splits.begin(), splits.end(), v) == splits.end()) splits.push_back(v);
  }
  const bool bf16 = precision == NKX_PREC_MIXED_BF16;  // one batch per workgroup
  hipEvent_t a, b;
  NKX_HIP_CHECK(hipEventCreate(&a));
  NKX_HIP_CHECK(hipEventCreate(&b));
  const int saved_split = c.opt.vgemm_split, saved_group = c.opt.density_group;
  const bool timing = c.timing.on;
  c.timing.on = false;
  // the faster of two builds per candidate (one build alone is noisy at the
  // few-percent level that separates neighbouring split counts)
  auto timed = [&]() {
    double best_ms = 1e300;
    for (int rep = 0; rep < 2; ++rep) {
      NKX_HIP_CHECK(hipEventRecord(a, c.stream));
      build();
      NKX_HIP_CHECK(hipEventRecord(b, c.stream));
      NKX_HIP_CHECK(hipEventSynchronize(b));
      float ms = 0.f;
      NKX_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
      best_ms = std::min(best_ms, (double)ms);
    }
    return best_ms;
  };
  double best = 1e300;
  try {
    c.opt.vgemm_split = base;
    build();  // warm caches and workspace
    int best_g = bf16 ? 1 : (precision == NKX_PREC_MIXED ? 65 : 2);
    // density shapes: point batches (64, 128 or 256 points) per workgroup
    // in the FP64 form; in the int8 form 65 (one batch) and 71 (one batch,
    // absolute block pairs in XCD tiles) -- 66 and 72 (two batches, 8 x 8
    // tiles) are never the fastest on large grids and are not timed
    const std::vector<int> shapes =
        precision == NKX_PREC_MIXED ? std::vector<int>{65, 71}
                                    : std::vector<int>{1, 2, 4};
    if (!bf16)
      for (int g : shapes) {
        c.opt.density_group = g;
        const double ms = timed();
        if (ms < best) {
          best = ms;
          best_g = g;
        }
      }
    c.opt.density_group = best_g;
    int best_k = base;
    best = 1e300;
    for (int k : splits) {
      c.opt.vgemm_split = k;
      const double ms = timed();
      if (ms < best) {
        best = ms;
        best_k = k;
      }
    }
    c.opt.vgemm_split = best_k;
    if (best_split) *best_split = best_k;
    if (best_group) *best_group = best_g;
  } catch (...) {
    c.opt.vgemm_split = saved_split;
    c.opt.density_group = saved_group;
    c.timing.on = timing;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    throw;
  }
  c.timing.on = timing;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return best;
}
}  // namespace

double autotune(Ctx &c, const double *P, const XcList &xl, bool gga, int precision, double *V,
                int *best_split, int *best_group) {
  ensure_plan(c);
  double E, N;
  return autotune_on(
      c, c.plan, precision, [&]() { vxc_build(c, P, xl, gga, precision, &E, &N, V); },
      best_split, best_group);
}

double molset_autotune(Ctx &c, const double *const *P, const XcList &xl, bool gga,
                       int precision, double *const *V, int *best_split, int *best_group) {
  NKX_REQUIRE(c.mplan.nmol > 0, NKX_ERR_STATE, "no molecule set: call nkx_molset_create");
  std::vector<double> E(c.mplan.nmol), N(c.mplan.nmol);
  return autotune_on(
      c, c.mplan, precision,
      [&]() { build_on_plan(c, c.mplan, 1, P, xl, gga, precision, E.data(), N.data(), V); },
      best_split, best_group);
}

// ---- HIP graph of one build on the ctx grid and basis ----------------------
namespace {
// Capture one build on plan pl (nmol molecules: P[m], V[m]) into c.graph.
void capture_on_plan(Ctx &c, Plan &pl, int nspin, const double *const *P, const XcList &xl,
                     bool gga, int precision, double *const *V) {
  GraphState &gs = c.graph;
  gs.release();
  const int nmol = pl.nmol;
  const BuildSetup b =
      prepare_build(gs.ws, pl, nspin, gga, precision, c.opt.vgemm_split, c.opt.density_group,
                    c.opt.mem_budget);
  NKX_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&gs.en_pinned),
                              3 * (size_t)nmol * sizeof(double), hipHostMallocDefault));
  gs.nmol = nmol;
  gs.nspin = nspin;
  gs.vtab_h.assign(V, V + (size_t)nspin * nmol);
  gs.vtab.upload(gs.vtab_h.data(), (size_t)nspin * nmol, c.stream);
  NKX_HIP_CHECK(hipStreamSynchronize(c.stream));
  const bool timing = c.timing.on;
  c.timing.on = false;  // no event records inside the graph
  NKX_HIP_CHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  try {
    enqueue_build(c, gs.ws, pl, b, P, xl, V, gs.vtab.p, gs.en_pinned);
  } catch (...) {
    hipGraph_t broken = nullptr;
    (void)hipStreamEndCapture(c.stream, &broken);
    if (broken) (void)hipGraphDestroy(broken);
    c.timing.on = timing;
    throw;
  }
  c.timing.on = timing;
  NKX_HIP_CHECK(hipStreamEndCapture(c.stream, &gs.graph));
  NKX_HIP_CHECK(hipGraphInstantiate(&gs.exec, gs.graph, nullptr, nullptr, 0));
}
}  // namespace

void graph_capture(Ctx &c, int nspin, const double *const *P, const XcList &xl, bool gga,
                   int precision, double *const *V) {
  ensure_plan(c);
  capture_on_plan(c, c.plan, nspin, P, xl, gga, precision, V);
  GraphState &gs = c.graph;
  gs.molset = false;
  gs.grid_version = c.grid.version;
  gs.basis_version = c.basis_version;
  gs.screen_tol = c.opt.screen_tol;
}

void molset_graph_capture(Ctx &c, const double *const *P, const XcList &xl, bool gga,
                          int precision, double *const *V) {
  NKX_REQUIRE(c.mplan.nmol > 0, NKX_ERR_STATE, "no molecule set: call nkx_molset_create");
  capture_on_plan(c, c.mplan, 1, P, xl, gga, precision, V);
  GraphState &gs = c.graph;
  gs.molset = true;
  gs.mplan_version = c.mplan_version;
  gs.screen_tol = c.mplan.screen_tol;
}

// Replay the captured build; E[m], N[m] for each molecule of the capture
// (an open-shell capture: E[0], N[0] = N_alpha, N[1] = N_beta).
void graph_launch(Ctx &c, double *E, double *N) {
  GraphState &gs = c.graph;
  NKX_REQUIRE(gs.exec, NKX_ERR_STATE, "no captured build: call nkx_graph_capture");
  if (gs.molset)
    NKX_REQUIRE(gs.mplan_version == c.mplan_version, NKX_ERR_STATE,
                "the molecule set changed since the capture");
  else
    NKX_REQUIRE(gs.grid_version == c.grid.version && gs.basis_version == c.basis_version &&
                    gs.screen_tol == c.opt.screen_tol,
                NKX_ERR_STATE, "grid, basis or screening changed since the capture");
  NKX_HIP_CHECK(hipGraphLaunch(gs.exec, c.stream));
  NKX_HIP_CHECK(hipStreamSynchronize(c.stream));
  if (gs.nspin == 2) {
    E[0] = gs.en_pinned[0];
    if (N) {
      N[0] = gs.en_pinned[1];
      N[1] = gs.en_pinned[2];
    }
    return;
  }
  for (int m = 0; m < gs.nmol; ++m) {
    E[m] = gs.en_pinned[3 * m];
    if (N) N[m] = gs.en_pinned[3 * m + 1];
  }
}

void GraphState::release() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (en_pinned) (void)hipHostFree(en_pinned);
  exec = nullptr;
  graph = nullptr;
  en_pinned = nullptr;
}

// One V_xc build on the ctx grid and basis.  nspin = 1: P[0] is the total
// density matrix; nspin = 2: P[0], P[1] the spin density matrices.
void vxc_build_spin(Ctx &c, int nspin, const double *const *P, const XcList &xl, bool gga,
                    int precision, double *E, double *N, double *const *V) {
  ensure_plan(c);
  build_on_plan(c, c.plan, nspin, P, xl, gga, precision, E, N, V);
}

void vxc_build(Ctx &c, const double *P, const XcList &xl, bool gga, int precision, double *E,
               double *N, double *V) {
  vxc_build_spin(c, 1, &P, xl, gga, precision, E, N, &V);
}

// Nuclear gradient of E_xc at fixed P (kernels_grad.hip).  nspin = 1: P[0]
// is the total density matrix; nspin = 2: P[0], P[1] the spin density
// matrices.  Every basis function must sit on an atom of the grid; grad is
// host memory, 3 natoms.
void xc_grad(Ctx &c, int nspin, const double *const *P, const XcList &xl, bool gga,
             int partition, bool mixed, double *E, double *grad) {
  ensure_plan(c);
  Plan &pl = c.plan;
  Workspace &ws = c.ws;
  hipStream_t s = c.stream;
  const int natoms = c.grid.natoms;
  NKX_REQUIRE(natoms > 0 && natoms <= 2048, NKX_ERR_UNSUPPORTED,
              "the gradient supports 1 .. 2048 atoms");
  NKX_REQUIRE(partition >= 0 && partition <= 2, NKX_ERR_INVALID_ARGUMENT,
              "partition: 0 fixed weights, 1 capped Becke, 2 Becke");
  NKX_REQUIRE(nspin == 1 || nspin == 2, NKX_ERR_NSPIN, "nspin must be 1 or 2");
  // function (sorted order) -> atom
  std::vector<int> fatom(pl.nbf, -1);
  for (int sp = 0; sp < pl.nbf; ++sp) {
    const int i = pl.fperm[sp];
    const double ctr[3] = {c.basis.cx[i], c.basis.cy[i], c.basis.cz[i]};
    for (int a = 0; a < natoms && fatom[sp] < 0; ++a) {
      double d2 = 0.0;
      for (int k = 0; k < 3; ++k) d2 += std::pow(ctr[k] - c.grid.centers_host[3 * a + k], 2);
      if (d2 <= 1e-20) fatom[sp] = a;
    }
    NKX_REQUIRE(fatom[sp] >= 0, NKX_ERR_INVALID_ARGUMENT,
                "basis function " + std::to_string(i) + " is not centred on an atom of the grid");
  }
  NKX_REQUIRE(pl.nmol == 1, NKX_ERR_STATE, "the gradient runs on the ctx grid and basis");
  const int narr = 4 + nspin;  // Phi, grad Phi (3), Z per spin
  // FP64: the tile arrays; mixed: the int8 slot (sections and FP32 grad Phi),
  // the second spin's Y sections and the density partials; both plus the
  // per-slot gradient partials
  const size_t slot_bytes =
      mixed ? (size_t)mixed_slot_bytes(gga ? 1 : 0) + (size_t)nspin * 4 * PT * sizeof(double) +
                  (nspin == 2 ? (size_t)mixed_zsection_bytes() : 0)
            : (size_t)narr * FB * PT * sizeof(double);
  int64_t max_slots =
      chunk_slot_budget(slot_bytes + 3 * FB * sizeof(double), c.opt.mem_budget);
  for (int64_t b = 0; b < pl.nb; ++b)
    max_slots = std::max<int64_t>(max_slots, pl.b_off_h[b + 1] - pl.b_off_h[b]);
  const int64_t slots_alloc = std::min<int64_t>(max_slots, std::max<int64_t>(pl.nslots, 1));
  const int nseg = 32;
  // one-wave workgroups of the weight-derivative kernel: one wave of points
  // each up to 16384 workgroups, so that the dispatcher balances waves of
  // unequal cost (points near many atoms cost more); fewer for large
  // molecules, so that the per-workgroup partials stay within 2^25 doubles
  const int64_t nblk_cap =
      std::max<int64_t>(1024, ((int64_t)1 << 25) / std::max<int64_t>(3 * (int64_t)natoms, 1));
  const int nblk = (int)std::min<int64_t>(std::max<int64_t>(pl.npts_pad / 64, 1),
                                          std::min<int64_t>(16384, nblk_cap));
  const size_t psb = (size_t)mixed_p_bytes(pl.mol_nfb_h[0]);
  const size_t dsz = (size_t)slots_alloc * 4 * PT;
  if (mixed) {
    ws.mstore.reserve((size_t)slots_alloc * mixed_slot_bytes(gga ? 1 : 0));
    ws.dpart.reserve(nspin * dsz);
    ws.Ps.reserve(nspin * psb);
    if (nspin == 2) ws.zstore_b.reserve((size_t)slots_alloc * mixed_zsection_bytes());
  } else {
    ws.tiles64.reserve((size_t)slots_alloc * narr * FB * PT);
    ws.Pd.reserve((size_t)nspin * pl.P_total);
  }
  ws.coef.reserve((size_t)4 * nspin * pl.npts_pad);
  ws.blockEN.reserve((size_t)3 * pl.nb);
  ws.red.reserve(3);
  // gradient buffers: ptab, epoint, tpt, cpart, gacc, gfun, wpart, wsum
  const size_t n_ep = pl.npts_pad, n_tp = 3 * pl.npts_pad, n_cp = (size_t)slots_alloc * 3 * FB;
  const size_t n_ga = (size_t)nseg * pl.nbf_pad * 3, n_gf = (size_t)pl.nbf_pad * 3;
  const size_t n_wp = (size_t)nblk * 3 * natoms, n_ws = (size_t)3 * natoms;
  // atom-pair table first, rounded to 256 bytes, so the others keep the
  // alignment of the allocation
  const size_t n_pt = ((size_t)2 * natoms * natoms + 31) & ~(size_t)31;
  ws.gbuf.reserve(n_pt + n_ep + n_tp + n_cp + n_ga + n_gf + n_wp + n_ws);
  double *ptab = ws.gbuf.p, *epoint = ptab + n_pt, *tpt = epoint + n_ep, *cpart = tpt + n_tp,
         *gacc = cpart + n_cp, *gfun = gacc + n_ga, *wpart = gfun + n_gf, *wsum = wpart + n_wp;
  ws.gint.reserve(pl.npts_pad);

  for (int sp = 0; sp < nspin; ++sp) {
    if (mixed)
      vkm_gather_p(s, pl.nbf, pl.mol_nfb_h[0], pl.fperm_d.p, P[sp], ws.Ps.p + sp * psb);
    else
      vk_gather_p(s, pl.nbf, pl.nbf_pad, pl.fperm_d.p, P[sp],
                  ws.Pd.p + (size_t)sp * pl.P_total);
  }
  vk_fill0(s, gacc, (int64_t)n_ga);
  const int ig = gga ? 1 : 0;
  double *coef_b = ws.coef.p + (size_t)4 * pl.npts_pad;
  for_each_chunk(pl, slots_alloc, [&](int64_t b0, int64_t b1, int64_t s0, int64_t s1) {
    if (mixed) {
      // density by the split form (contraction per spin, then the functional
      // with w e per point), Y slices (the first spin's over the Z sections),
      // then the basis-centre term per spin
      vkm_colloc(s, s0, s1, pl, ig, ws.mstore.p);
      vkm_pm_group(s, b0, b1, s0, pl, ig, ws.mstore.p);
      for (int sp = 0; sp < nspin; ++sp)
        vkm_density(s, 65, b0, b1, s0, pl, ws.Ps.p + sp * psb, ig, ws.mstore.p, xl, ws.coef.p,
                    ws.blockEN.p, ws.dpart.p + sp * dsz);
      if (nspin == 1) {
        vkm_density_functional(s, b0, b1, s0, pl, ig, xl, ws.coef.p, ws.blockEN.p, ws.dpart.p,
                               epoint);
      } else {
        vkm_density_finish_pol(s, b0, b1, s0, pl, ig, nullptr, nullptr, xl, ws.coef.p, coef_b,
                               ws.blockEN.p, ws.dpart.p, ws.dpart.p + dsz, epoint);
        vkgm_y_slots(s, s0, s1, pl, ig, coef_b, ws.mstore.p, ws.zstore_b.p);
      }
      vkgm_y_slots(s, s0, s1, pl, ig, ws.coef.p, ws.mstore.p);
      vkgm_basis(s, b0, b1, s0, pl, ws.Ps.p, ig, ws.mstore.p, ws.coef.p, nullptr, false, cpart,
                 tpt);
      vkg_fold(s, b0, b1, s0, pl, nseg, cpart, gacc);
      if (nspin == 2) {
        vkgm_basis(s, b0, b1, s0, pl, ws.Ps.p + psb, ig, ws.mstore.p, coef_b, ws.zstore_b.p,
                   true, cpart, tpt);
        vkg_fold(s, b0, b1, s0, pl, nseg, cpart, gacc);
      }
    } else {
      // the density kernel writes coef per point as [spin][4]
      vk_colloc(s, s0, s1, pl, narr, 1, ws.tiles64.p);
      vk_density(s, nspin, 4, b0, b1, s0, pl, ws.Pd.p, narr, ig, ws.tiles64.p, xl, ws.coef.p,
                 ws.blockEN.p, epoint);
      for (int sp = 0; sp < nspin; ++sp) {
        vkg_basis(s, b0, b1, s0, pl, ws.Pd.p + (size_t)sp * pl.P_total, narr, ig, ws.tiles64.p,
                  ws.coef.p + 4 * sp, 4 * nspin, sp > 0, cpart, tpt);
        vkg_fold(s, b0, b1, s0, pl, nseg, cpart, gacc);
      }
    }
  });
  vkg_colsum(s, nseg, pl.nbf_pad * 3, gacc, gfun);
  vk_en_reduce(s, pl, ws.blockEN.p, ws.red.p);
  vkg_owners(s, pl, c.grid.owner.p, ws.gint.p);
  vkg_weights(s, pl, natoms, c.grid.centers.p, ws.gint.p, epoint, tpt, partition, nblk, ptab,
              wpart);
  vkg_colsum(s, nblk, 3 * natoms, wpart, wsum);
  std::vector<double> gf(n_gf), wsh(n_ws);
  double en[3];
  NKX_HIP_CHECK(hipMemcpyAsync(gf.data(), gfun, n_gf * sizeof(double), hipMemcpyDeviceToHost, s));
  NKX_HIP_CHECK(hipMemcpyAsync(wsh.data(), wsum, n_ws * sizeof(double), hipMemcpyDeviceToHost, s));
  NKX_HIP_CHECK(hipMemcpyAsync(en, ws.red.p, sizeof(en), hipMemcpyDeviceToHost, s));
  NKX_HIP_CHECK(hipStreamSynchronize(s));
  for (int k = 0; k < 3 * natoms; ++k) grad[k] = wsh[k];
  for (int sp = 0; sp < pl.nbf; ++sp)
    for (int k = 0; k < 3; ++k) grad[3 * fatom[sp] + k] += gf[3 * sp + k];
  *E = en[0];
}

// rho, grad rho and sigma of P at every grid point, in the grid's order: the
// FP64 batched density with an empty functional list, then a scatter through
// the point permutation.  Stream-ordered; outputs other than rho may be NULL.
void density_points(Ctx &c, const double *P, double *rho, double *gx, double *gy, double *gz,
                    double *sigma) {
  ensure_plan(c);
  Plan &pl = c.plan;
  NKX_REQUIRE(pl.nmol == 1, NKX_ERR_STATE, "the per-point density runs on the ctx grid and basis");
  Workspace &ws = c.ws;
  hipStream_t s = c.stream;
  const int narr = 5;
  int64_t max_slots = chunk_slot_budget((size_t)narr * FB * PT * sizeof(double), c.opt.mem_budget);
  for (int64_t b = 0; b < pl.nb; ++b)
    max_slots = std::max<int64_t>(max_slots, pl.b_off_h[b + 1] - pl.b_off_h[b]);
  const int64_t slots_alloc = std::min<int64_t>(max_slots, std::max<int64_t>(pl.nslots, 1));
  ws.tiles64.reserve((size_t)slots_alloc * narr * FB * PT);
  ws.Pd.reserve((size_t)pl.P_total);
  ws.coef.reserve((size_t)4 * pl.npts_pad);
  ws.blockEN.reserve((size_t)3 * pl.nb);
  ws.gbuf.reserve((size_t)4 * pl.npts_pad);
  XcList none{};
  none.n = 0;
  vk_gather_p(s, pl.nbf, pl.nbf_pad, pl.fperm_d.p, P, ws.Pd.p);
  for_each_chunk(pl, slots_alloc, [&](int64_t b0, int64_t b1, int64_t s0, int64_t s1) {
    vk_colloc(s, s0, s1, pl, narr, 1, ws.tiles64.p);
    vk_density(s, 1, 4, b0, b1, s0, pl, ws.Pd.p, narr, 1, ws.tiles64.p, none, ws.coef.p,
               ws.blockEN.p, nullptr, ws.gbuf.p);
  });
  vk_scatter_density(s, pl, ws.gbuf.p, rho, gx, gy, gz, sigma);
}

int Timing::begin(int stage, hipStream_t s) {
  if (!on) return -1;
  const size_t k = pending_stage.size();
  if (k >= pool.size()) {
    hipEvent_t a, b;
    NKX_HIP_CHECK(hipEventCreate(&a));
    NKX_HIP_CHECK(hipEventCreate(&b));
    pool.emplace_back(a, b);
  }
  pending_stage.push_back(stage);
  NKX_HIP_CHECK(hipEventRecord(pool[k].first, s));
  return (int)k;
}

void Timing::end(int slot, hipStream_t s) {
  if (slot < 0) return;
  NKX_HIP_CHECK(hipEventRecord(pool[slot].second, s));
}

void Timing::collect() {
  for (size_t k = 0; k < pending_stage.size(); ++k) {
    float ms_k = 0.f;
    NKX_HIP_CHECK(hipEventElapsedTime(&ms_k, pool[k].first, pool[k].second));
    ms[pending_stage[k]] += ms_k;
    launches[pending_stage[k]] += 1;
  }
  pending_stage.clear();
}

void Timing::release() {
  for (auto &e : pool) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  pool.clear();
  pending_stage.clear();
}

void overlap_build(Ctx &c, double *S) {
  ensure_plan(c);
  Plan &pl = c.plan;
  Workspace &ws = c.ws;
  hipStream_t s = c.stream;
  const int narr = 2;
  const int nsplit = choose_nsplit(pl);
  int64_t max_slots =
      chunk_slot_budget((size_t)narr * FB * PT * sizeof(double), c.opt.mem_budget);
  for (int64_t b = 0; b < pl.nb; ++b)
    max_slots = std::max<int64_t>(max_slots, pl.b_off_h[b + 1] - pl.b_off_h[b]);
  const int64_t slots_alloc = std::min<int64_t>(max_slots, std::max<int64_t>(pl.nslots, 1));
  ws.tiles64.reserve((size_t)slots_alloc * narr * FB * PT);
  const size_t vsz = (size_t)nsplit * std::max(pl.npairs, 1) * FB * FB;
  ws.vpart.reserve(vsz);
  vk_fill0(s, ws.vpart.p, (int64_t)vsz);
  vk_fill0(s, S, (int64_t)pl.nbf * pl.nbf);
  double *const *Stab = upload_ptrs(c, std::vector<double *>{S});
  for_each_chunk(pl, slots_alloc, [&](int64_t b0, int64_t b1, int64_t s0, int64_t s1) {
    vk_colloc(s, s0, s1, pl, narr, 0, ws.tiles64.p);
    vk_zoverlap(s, s0, s1, pl, narr, ws.tiles64.p);
    vk_vgemm(s, b0, b1, s0, pl, nsplit, narr, 1, ws.tiles64.p, ws.vpart.p);
  });
  vk_vreduce(s, pl, nsplit, ws.vpart.p, Stab, pl.mol_nbf.p);
  NKX_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace nkx
