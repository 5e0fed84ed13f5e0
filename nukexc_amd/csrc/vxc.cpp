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

void ensure_plan(Ctx &c) {
  Plan &pl = c.plan;
  NKX_REQUIRE(c.grid.npts > 0, NKX_ERR_STATE, "no grid: call nkx_make_flat_grid or nkx_set_grid");
  NKX_REQUIRE(c.basis.nbf > 0, NKX_ERR_STATE, "no basis: call nkx_set_basis");
  if (pl.grid_version == c.grid.version && pl.basis_version == c.basis_version &&
      pl.screen_tol == c.opt.screen_tol && pl.pt == PT)
    return;
  hipStream_t s = c.stream;
  const BasisHost &B = c.basis;
  const int64_t npts = c.grid.npts;

  // ---- points ------------------------------------------------------------
  std::vector<double> P(3 * npts), W(npts);
  NKX_HIP_CHECK(hipMemcpyAsync(P.data(), c.grid.pts.p, 3 * npts * sizeof(double),
                               hipMemcpyDeviceToHost, s));
  NKX_HIP_CHECK(hipMemcpyAsync(W.data(), c.grid.w.p, npts * sizeof(double),
                               hipMemcpyDeviceToHost, s));
  NKX_HIP_CHECK(hipStreamSynchronize(s));
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
  std::vector<double> px(npad), py(npad), pz(npad), pw(npad);
  std::vector<int> pidx(npad);
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
      px[sp] = P[3 * src];
      py[sp] = P[3 * src + 1];
      pz[sp] = P[3 * src + 2];
      pw[sp] = sp < npts ? W[src] : 0.0;
      pidx[sp] = sp < npts ? (int)src : -1;
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
  std::vector<double> rad(nbf);
  Box cb{{1e300, 1e300, 1e300}, {-1e300, -1e300, -1e300}};
  for (int i = 0; i < nbf; ++i) {
    double nrm, cut;
    sto_norm_cutoff(B.n[i], B.zeta[i], c.opt.screen_tol, &nrm, &cut);
    // Same radial bound with the stored normalisation (callers may scale it).
    if (B.norm[i] != nrm && B.norm[i] > 0) {
      double r = cut;
      for (int it = 0; it < 60; ++it) {
        const double v = std::log(B.norm[i]) + (B.n[i] - 1) * std::log(r) - B.zeta[i] * r -
                         std::log(c.opt.screen_tol);
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
  pl.fperm.resize(nbf);
  for (int i = 0; i < nbf; ++i) pl.fperm[i] = std::get<2>(fkey[i]);

  std::vector<FuncRec> fr(nbf_pad);
  std::vector<HarmTerm> terms;
  for (int sp = 0; sp < nbf_pad; ++sp) {
    FuncRec r{};
    if (sp < nbf) {
      const int i = pl.fperm[sp];
      const auto &t = solid_harmonic_terms(B.l[i], B.m[i]);
      r.cx = B.cx[i];
      r.cy = B.cy[i];
      r.cz = B.cz[i];
      r.zeta = B.zeta[i];
      r.norm = B.norm[i];
      r.k = B.n[i] - B.l[i] - 1;
      r.term_off = (int)terms.size();
      r.term_cnt = (int)t.size();
      terms.insert(terms.end(), t.begin(), t.end());
    } else {
      r.zeta = 1.0;
      r.norm = 0.0;
      r.term_off = 0;
      r.term_cnt = 0;
    }
    fr[sp] = r;
  }
  if (terms.empty()) terms.push_back(HarmTerm{0, 0, 0, 0, 0});

  // ---- screening: live (block, batch) pairs -------------------------------
  std::vector<int> b_off(nb + 1, 0), b_fbl, slot_b;
  b_fbl.reserve((size_t)nb * 4);
  for (int64_t b = 0; b < nb; ++b) {
    b_off[b] = (int)b_fbl.size();
    for (int f = 0; f < nfb; ++f) {
      bool live = false;
      for (int k = f * FB; k < std::min(nbf, (f + 1) * FB) && !live; ++k) {
        const int i = pl.fperm[k];
        live = box_dist2(boxes[b], B.cx[i], B.cy[i], B.cz[i]) < rad[i] * rad[i];
      }
      if (live) {
        b_fbl.push_back(f);
        slot_b.push_back((int)b);
      }
    }
    NKX_REQUIRE(b_fbl.size() < (size_t)INT32_MAX, NKX_ERR_UNSUPPORTED,
                "too many live (block, batch) pairs for 32-bit slot numbers");
  }
  b_off[nb] = (int)b_fbl.size();
  // algorithmic work: true functions of the live blocks, true points
  {
    double nb2g = 0, nbg = 0;
    for (int64_t b = 0; b < nb; ++b) {
      double nfun = 0;
      for (int k = b_off[b]; k < b_off[b + 1]; ++k)
        nfun += std::min(nbf, (b_fbl[k] + 1) * FB) - b_fbl[k] * FB;
      const double g = (double)std::min<int64_t>(PT, npts - b * PT);
      nb2g += nfun * nfun * g;
      nbg += nfun * g;
    }
    pl.work_nb2g = nb2g;
    pl.work_nbg = nbg;
  }
  if (b_fbl.empty()) {  // keep device buffers non-empty
    b_fbl.push_back(0);
    slot_b.push_back(0);
  }
  std::vector<int> tile_slot((size_t)nfb * nb, -1);
  std::vector<char> pair_live((size_t)nfb * nfb, 0);
  for (int64_t b = 0; b < nb; ++b) {
    for (int k = b_off[b]; k < b_off[b + 1]; ++k) {
      tile_slot[(size_t)b_fbl[k] * nb + b] = k;
      for (int k2 = k; k2 < b_off[b + 1]; ++k2) pair_live[(size_t)b_fbl[k] * nfb + b_fbl[k2]] = 1;
    }
  }
  pl.pair_I_h.clear();
  pl.pair_J_h.clear();
  for (int I = 0; I < nfb; ++I)
    for (int J = I; J < nfb; ++J)
      if (pair_live[(size_t)I * nfb + J]) {
        pl.pair_I_h.push_back(I);
        pl.pair_J_h.push_back(J);
      }

  // ---- upload -------------------------------------------------------------
  pl.nbf = nbf;
  pl.nbf_pad = nbf_pad;
  pl.nfb = nfb;
  pl.fb = FB;
  pl.npts = npts;
  pl.npts_pad = npad;
  pl.nb = nb;
  pl.b_off_h = b_off;
  pl.nslots = b_off[nb];
  pl.pairs_kept = b_off[nb];
  pl.npairs = (int)pl.pair_I_h.size();
  pl.fperm_d.upload(pl.fperm.data(), nbf, s);
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
  NKX_HIP_CHECK(hipStreamSynchronize(s));
  pl.grid_version = c.grid.version;
  pl.basis_version = c.basis_version;
  pl.screen_tol = c.opt.screen_tol;
  pl.pt = PT;
}

namespace {

// Number of slots per chunk so that the chunk's tiles fit the budget.
int64_t chunk_slot_budget(size_t slot_bytes) {
  size_t freeb = 0, total = 0;
  NKX_HIP_CHECK(hipMemGetInfo(&freeb, &total));
  size_t budget = std::min<size_t>((size_t)48 << 30, freeb / 5 * 2);
  budget = std::max<size_t>(budget, (size_t)64 << 20);
  return (int64_t)(budget / slot_bytes);
}

int choose_nsplit(const Plan &pl) {
  if (pl.npairs == 0) return 1;
  int64_t ns = (2048 + pl.npairs - 1) / pl.npairs;
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

// One V_xc build.  nspin = 1: P is the total density matrix, V the potential.
// nspin = 2 (FP64 only): P[0], P[1] are the spin density matrices and V[0],
// V[1] the spin potentials.  N[0..nspin-1] receives the electron counts.
void vxc_build_spin(Ctx &c, int nspin, const double *const *P, const XcList &xl, bool gga,
                    int precision, double *E, double *N, double *const *V) {
  ensure_plan(c);
  Plan &pl = c.plan;
  Workspace &ws = c.ws;
  hipStream_t s = c.stream;
  const int nval = gga ? 4 : 1;
  const int narr = nval + nspin;
  const int nsplit = choose_nsplit(pl);
  const bool bf16 = precision == NKX_PREC_MIXED_BF16;
  const bool mixed = precision == NKX_PREC_MIXED || bf16;
  NKX_REQUIRE(!mixed || nspin == 1, NKX_ERR_UNSUPPORTED,
              "the spin-polarised build is available in the FP64 form only");
  const size_t slot_bytes = bf16    ? (size_t)bf16_slot_bytes(gga ? 1 : 0)
                            : mixed ? (size_t)mixed_slot_bytes(gga ? 1 : 0)
                                    : (size_t)narr * FB * PT * sizeof(double);
  int64_t max_slots = chunk_slot_budget(slot_bytes);
  // at least the largest single batch
  for (int64_t b = 0; b < pl.nb; ++b)
    max_slots = std::max<int64_t>(max_slots, pl.b_off_h[b + 1] - pl.b_off_h[b]);
  const int64_t slots_alloc = std::min<int64_t>(max_slots, std::max<int64_t>(pl.nslots, 1));
  const size_t psz = (size_t)pl.nbf_pad * pl.nbf_pad;
  if (mixed) {
    ws.mstore.reserve((size_t)slots_alloc * slot_bytes);
    ws.Ps.reserve(bf16 ? (size_t)6 * pl.nbf_pad * pl.nbf_pad : (size_t)mixed_p_bytes(pl.nfb));
  } else {
    ws.tiles64.reserve((size_t)slots_alloc * narr * FB * PT);
    ws.Pd.reserve(nspin * psz);
  }
  ws.coef.reserve((size_t)4 * nspin * pl.npts_pad);
  ws.blockEN.reserve((size_t)3 * pl.nb);
  const size_t vsz = (size_t)nsplit * std::max(pl.npairs, 1) * FB * FB;
  ws.vpart.reserve(nspin * vsz);
  ws.red.reserve(3);

  Timing &tm = c.timing;
  const int t_build = tm.begin(kStageBuild, s);
  int t = tm.begin(kStageFinal, s);
  uint16_t *bstore = reinterpret_cast<uint16_t *>(ws.mstore.p);
  uint16_t *bPs = reinterpret_cast<uint16_t *>(ws.Ps.p);
  for (int sp = 0; sp < nspin; ++sp) {
    if (bf16)
      vkb_gather_p(s, pl.nbf, pl.nfb, pl.fperm_d.p, P[sp], bPs);
    else if (mixed)
      vkm_gather_p(s, pl.nbf, pl.nfb, pl.fperm_d.p, P[sp], ws.Ps.p);
    else
      vk_gather_p(s, pl.nbf, pl.nbf_pad, pl.fperm_d.p, P[sp], ws.Pd.p + sp * psz);
    NKX_HIP_CHECK(hipMemsetAsync(V[sp], 0, (size_t)pl.nbf * pl.nbf * sizeof(double), s));
  }
  NKX_HIP_CHECK(hipMemsetAsync(ws.vpart.p, 0, nspin * vsz * sizeof(double), s));
  tm.end(t, s);
  for_each_chunk(pl, slots_alloc, [&](int64_t b0, int64_t b1, int64_t s0, int64_t s1) {
    const int ig = gga ? 1 : 0;
    int tt = tm.begin(kStageColloc, s);
    if (bf16)
      vkb_colloc(s, s0, s1, pl, ig, bstore);
    else if (mixed)
      vkm_colloc(s, s0, s1, pl, ig, ws.mstore.p);
    else
      vk_colloc(s, s0, s1, pl, narr, ig, ws.tiles64.p);
    tm.end(tt, s);
    tt = tm.begin(kStageDensity, s);
    if (bf16)
      vkb_density(s, 6, b0, b1, s0, pl, bPs, ig, bstore, xl, ws.coef.p, ws.blockEN.p);
    else if (mixed)
      vkm_density(s, b0, b1, s0, pl, ws.Ps.p, ig, ws.mstore.p, xl, ws.coef.p, ws.blockEN.p);
    else
      vk_density(s, nspin, b0, b1, s0, pl, ws.Pd.p, narr, ig, ws.tiles64.p, xl, ws.coef.p,
                 ws.blockEN.p, nullptr);
    tm.end(tt, s);
    tt = tm.begin(kStageVgemm, s);
    if (bf16)
      vkb_vgemm(s, 6, b0, b1, s0, pl, nsplit, ig, bstore, ws.vpart.p);
    else if (mixed)
      vkm_vgemm(s, b0, b1, s0, pl, nsplit, ig, ws.mstore.p, ws.vpart.p);
    else
      vk_vgemm(s, b0, b1, s0, pl, nsplit, narr, nspin, ws.tiles64.p, ws.vpart.p);
    tm.end(tt, s);
  });
  t = tm.begin(kStageFinal, s);
  for (int sp = 0; sp < nspin; ++sp) vk_vreduce(s, pl, nsplit, ws.vpart.p + sp * vsz, V[sp]);
  vk_en_reduce(s, pl.nb, ws.blockEN.p, ws.red.p);
  tm.end(t, s);
  tm.end(t_build, s);
  double en[3];
  NKX_HIP_CHECK(hipMemcpyAsync(en, ws.red.p, sizeof(en), hipMemcpyDeviceToHost, s));
  NKX_HIP_CHECK(hipStreamSynchronize(s));
  tm.collect();
  c.work.flops_density = 2.0 * nspin * pl.work_nb2g;
  c.work.flops_v = 2.0 * nspin * pl.work_nb2g;
  c.work.colloc_elems = pl.work_nbg;
  *E = en[0];
  for (int sp = 0; sp < nspin; ++sp) N[sp] = en[1 + sp];
}

void vxc_build(Ctx &c, const double *P, const XcList &xl, bool gga, int precision, double *E,
               double *N, double *V) {
  vxc_build_spin(c, 1, &P, xl, gga, precision, E, N, &V);
}

// Nuclear gradient of E_xc at fixed P (kernels_grad.hip).  Every basis
// function must sit on an atom of the grid; grad is host memory, 3 natoms.
void xc_grad(Ctx &c, const double *P, const XcList &xl, bool gga, int partition, double *E,
             double *grad) {
  ensure_plan(c);
  Plan &pl = c.plan;
  Workspace &ws = c.ws;
  hipStream_t s = c.stream;
  const int natoms = c.grid.natoms;
  NKX_REQUIRE(natoms > 0 && natoms <= 2048, NKX_ERR_UNSUPPORTED,
              "the gradient supports 1 .. 2048 atoms");
  NKX_REQUIRE(partition >= 0 && partition <= 2, NKX_ERR_INVALID_ARGUMENT,
              "partition: 0 fixed weights, 1 capped Becke, 2 Becke");
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
  const int narr = 5;  // Phi, grad Phi (3), Z
  const size_t slot_bytes = (size_t)narr * FB * PT * sizeof(double);
  // tiles plus the per-slot gradient partials
  int64_t max_slots = chunk_slot_budget(slot_bytes + 3 * FB * sizeof(double));
  for (int64_t b = 0; b < pl.nb; ++b)
    max_slots = std::max<int64_t>(max_slots, pl.b_off_h[b + 1] - pl.b_off_h[b]);
  const int64_t slots_alloc = std::min<int64_t>(max_slots, std::max<int64_t>(pl.nslots, 1));
  const int nseg = 32;
  const int nblk = 1024;
  ws.tiles64.reserve((size_t)slots_alloc * narr * FB * PT);
  ws.Pd.reserve((size_t)pl.nbf_pad * pl.nbf_pad);
  ws.coef.reserve((size_t)4 * pl.npts_pad);
  ws.blockEN.reserve((size_t)3 * pl.nb);
  ws.red.reserve(3);
  // gradient buffers: epoint, tpt, cpart, gacc, gfun, wpart, wsum
  const size_t n_ep = pl.npts_pad, n_tp = 3 * pl.npts_pad, n_cp = (size_t)slots_alloc * 3 * FB;
  const size_t n_ga = (size_t)nseg * pl.nbf_pad * 3, n_gf = (size_t)pl.nbf_pad * 3;
  const size_t n_wp = (size_t)nblk * 3 * natoms, n_ws = (size_t)3 * natoms;
  ws.gbuf.reserve(n_ep + n_tp + n_cp + n_ga + n_gf + n_wp + n_ws);
  double *epoint = ws.gbuf.p, *tpt = epoint + n_ep, *cpart = tpt + n_tp, *gacc = cpart + n_cp,
         *gfun = gacc + n_ga, *wpart = gfun + n_gf, *wsum = wpart + n_wp;
  ws.gint.reserve(pl.npts_pad);

  vk_gather_p(s, pl.nbf, pl.nbf_pad, pl.fperm_d.p, P, ws.Pd.p);
  NKX_HIP_CHECK(hipMemsetAsync(gacc, 0, n_ga * sizeof(double), s));
  const int ig = gga ? 1 : 0;
  for_each_chunk(pl, slots_alloc, [&](int64_t b0, int64_t b1, int64_t s0, int64_t s1) {
    vk_colloc(s, s0, s1, pl, narr, 1, ws.tiles64.p);
    vk_density(s, 1, b0, b1, s0, pl, ws.Pd.p, narr, ig, ws.tiles64.p, xl, ws.coef.p,
               ws.blockEN.p, epoint);
    vkg_basis(s, b0, b1, s0, pl, ws.Pd.p, narr, ig, ws.tiles64.p, ws.coef.p, cpart, tpt);
    vkg_fold(s, b0, b1, s0, pl, nseg, cpart, gacc);
  });
  vkg_colsum(s, nseg, pl.nbf_pad * 3, gacc, gfun);
  vk_en_reduce(s, pl.nb, ws.blockEN.p, ws.red.p);
  vkg_owners(s, pl, c.grid.owner.p, ws.gint.p);
  vkg_weights(s, pl, natoms, c.grid.centers.p, ws.gint.p, epoint, tpt, partition, nblk, wpart);
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
  int64_t max_slots = chunk_slot_budget((size_t)narr * FB * PT * sizeof(double));
  for (int64_t b = 0; b < pl.nb; ++b)
    max_slots = std::max<int64_t>(max_slots, pl.b_off_h[b + 1] - pl.b_off_h[b]);
  const int64_t slots_alloc = std::min<int64_t>(max_slots, std::max<int64_t>(pl.nslots, 1));
  ws.tiles64.reserve((size_t)slots_alloc * narr * FB * PT);
  const size_t vsz = (size_t)nsplit * std::max(pl.npairs, 1) * FB * FB;
  ws.vpart.reserve(vsz);
  NKX_HIP_CHECK(hipMemsetAsync(ws.vpart.p, 0, vsz * sizeof(double), s));
  NKX_HIP_CHECK(hipMemsetAsync(S, 0, (size_t)pl.nbf * pl.nbf * sizeof(double), s));
  for_each_chunk(pl, slots_alloc, [&](int64_t b0, int64_t b1, int64_t s0, int64_t s1) {
    vk_colloc(s, s0, s1, pl, narr, 0, ws.tiles64.p);
    vk_zoverlap(s, s0, s1, pl, narr, ws.tiles64.p);
    vk_vgemm(s, b0, b1, s0, pl, nsplit, narr, 1, ws.tiles64.p, ws.vpart.p);
  });
  vk_vreduce(s, pl, nsplit, ws.vpart.p, S);
  NKX_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace nkx
