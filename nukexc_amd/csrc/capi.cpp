// extern "C" entry points of libnukexc_hip.so (include/nukexc_hip.h).
// Every call converts internal errors to an nkx_status and records the
// message on the ctx; no exception leaves this file.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "nkx_internal.hpp"
#include "vxc_kernels.hpp"

struct nkx_ctx {
  nkx::Ctx c;
};

namespace {

thread_local std::string g_thread_error;

template <typename F>
int guarded(nkx_ctx *ctx, F &&fn) {
  try {
    if (ctx) NKX_HIP_CHECK(hipSetDevice(ctx->c.device));
    fn();
    return NKX_OK;
  } catch (const nkx::Error &e) {
    (ctx ? ctx->c.last_error : g_thread_error) = e.msg;
    return e.code;
  } catch (const std::bad_alloc &) {
    (ctx ? ctx->c.last_error : g_thread_error) = "host allocation failed";
    return NKX_ERR_OUT_OF_MEMORY;
  } catch (const std::exception &e) {
    (ctx ? ctx->c.last_error : g_thread_error) = e.what();
    return NKX_ERR_INVALID_ARGUMENT;
  }
}

#define NKX_CTX_CHECK(ctx)                                                           \
  do {                                                                               \
    if (!(ctx)) {                                                                    \
      g_thread_error = "null nkx_ctx";                                               \
      return NKX_ERR_INVALID_ARGUMENT;                                               \
    }                                                                                \
  } while (0)

int functional_info(int id, int *family, int *kind) {
  switch (id) {
    case 1: *family = NKX_FAMILY_LDA; *kind = 0; return 0;
    case 12: case 13: *family = NKX_FAMILY_LDA; *kind = 1; return 0;
    case 101: *family = NKX_FAMILY_GGA; *kind = 0; return 0;
    case 130: *family = NKX_FAMILY_GGA; *kind = 1; return 0;
    case 106: *family = NKX_FAMILY_GGA; *kind = 0; return 0;
    case 7: *family = NKX_FAMILY_LDA; *kind = 1; return 0;
    case 131: *family = NKX_FAMILY_GGA; *kind = 1; return 0;
    case 109: *family = NKX_FAMILY_GGA; *kind = 0; return 0;
    case 134: *family = NKX_FAMILY_GGA; *kind = 1; return 0;
    default: return 1;
  }
}

double default_threshold(int id) { return id == 130 || id == 134 ? 1e-12 : 1e-15; }

}  // namespace

namespace nkx {
bool xc_list_make(int nfunc, const int *ids, double thr_override, XcList *out, bool *any_gga) {
  if (nfunc < 1 || nfunc > 4 || !ids) return false;
  out->n = nfunc;
  *any_gga = false;
  for (int k = 0; k < nfunc; ++k) {
    int fam, kind;
    if (functional_info(ids[k], &fam, &kind)) return false;
    out->ids[k] = ids[k];
    out->thr[k] = thr_override > 0 ? thr_override : default_threshold(ids[k]);
    if (fam == NKX_FAMILY_GGA) *any_gga = true;
  }
  for (int k = nfunc; k < 4; ++k) {
    out->ids[k] = 0;
    out->thr[k] = 1.0;
  }
  return true;
}
}  // namespace nkx

using nkx::Error;

extern "C" {

int nkx_version(int *major, int *minor, int *patch) {
  if (major) *major = 0;
  if (minor) *minor = 1;
  if (patch) *patch = 0;
  return NKX_OK;
}

int nkx_create(int device, void *hip_stream, nkx_ctx **out) {
  if (!out) {
    g_thread_error = "nkx_create: out is NULL";
    return NKX_ERR_INVALID_ARGUMENT;
  }
  *out = nullptr;
  nkx_ctx *ctx = nullptr;
  int rc = guarded(nullptr, [&] {
    int n = 0;
    NKX_HIP_CHECK(hipGetDeviceCount(&n));
    NKX_REQUIRE(device >= 0 && device < n, NKX_ERR_INVALID_ARGUMENT, "no such device");
    NKX_HIP_CHECK(hipSetDevice(device));
    nkx::becke_scratch_init(device);
    ctx = new nkx_ctx();
    ctx->c.device = device;
    if (hip_stream) {
      ctx->c.stream = static_cast<hipStream_t>(hip_stream);
    } else {
      // A blocking stream: ordered with the legacy default stream that
      // callers (and torch) use for their own copies.
      NKX_HIP_CHECK(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamDefault));
      ctx->c.own_stream = true;
    }
  });
  if (rc != NKX_OK) {
    delete ctx;
    return rc;
  }
  *out = ctx;
  return NKX_OK;
}

int nkx_destroy(nkx_ctx *ctx) {
  if (!ctx) return NKX_OK;
  (void)hipSetDevice(ctx->c.device);
  (void)hipStreamSynchronize(ctx->c.stream);
  if (ctx->c.own_stream) (void)hipStreamDestroy(ctx->c.stream);
  delete ctx;
  return NKX_OK;
}

int nkx_sync(nkx_ctx *ctx) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] { NKX_HIP_CHECK(hipStreamSynchronize(ctx->c.stream)); });
}

const char *nkx_last_error(const nkx_ctx *ctx) {
  return ctx ? ctx->c.last_error.c_str() : g_thread_error.c_str();
}

int nkx_malloc(nkx_ctx *ctx, size_t bytes, void **ptr) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    NKX_REQUIRE(ptr, NKX_ERR_INVALID_ARGUMENT, "nkx_malloc: ptr is NULL");
    *ptr = nullptr;
    if (bytes) NKX_HIP_CHECK(hipMalloc(ptr, bytes));
  });
}

int nkx_free(nkx_ctx *ctx, void *ptr) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    if (ptr) {
      NKX_HIP_CHECK(hipStreamSynchronize(ctx->c.stream));
      NKX_HIP_CHECK(hipFree(ptr));
    }
  });
}

int nkx_memcpy_h2d(nkx_ctx *ctx, void *dst, const void *src, size_t bytes) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    if (!bytes) return;
    NKX_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->c.stream));
    NKX_HIP_CHECK(hipStreamSynchronize(ctx->c.stream));
  });
}

int nkx_memcpy_d2h(nkx_ctx *ctx, void *dst, const void *src, size_t bytes) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    if (!bytes) return;
    NKX_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->c.stream));
    NKX_HIP_CHECK(hipStreamSynchronize(ctx->c.stream));
  });
}

int nkx_memset(nkx_ctx *ctx, void *dst, int value, size_t bytes) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    if (bytes) NKX_HIP_CHECK(hipMemsetAsync(dst, value, bytes, ctx->c.stream));
  });
}

int nkx_copy_d2d(nkx_ctx *ctx, void *dst, const void *src, size_t bytes) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    NKX_REQUIRE(bytes % 16 == 0 && ((uintptr_t)dst % 16) == 0 && ((uintptr_t)src % 16) == 0,
                NKX_ERR_INVALID_ARGUMENT, "copy_d2d: 16-byte aligned pointers and size");
    NKX_REQUIRE(bytes == 0 || (dst && src), NKX_ERR_INVALID_ARGUMENT, "copy_d2d: null pointer");
    nkx::launch_copy16(ctx->c.stream, dst, src, bytes);
  });
}

int nkx_lebedev_npts(int order, int *npts, int *actual) {
  return guarded(nullptr, [&] { nkx::lebedev_lookup(order, npts, actual); });
}

int nkx_lebedev_rule(int order, double *xyzw) {
  return guarded(nullptr, [&] {
    NKX_REQUIRE(xyzw, NKX_ERR_INVALID_ARGUMENT, "xyzw is NULL");
    int actual = 0;
    nkx::lebedev_lookup(order, nullptr, &actual);
    nkx::lebedev_expand(actual, xyzw);
  });
}

int nkx_radial_rule(int scheme, int n, double R, double alpha, double *r, double *w) {
  return guarded(nullptr, [&] {
    NKX_REQUIRE(r && w, NKX_ERR_INVALID_ARGUMENT, "output is NULL");
    nkx::radial_rule(scheme, n, R, alpha, r, w);
  });
}

int nkx_radial_scale(int scheme, int Z, double *R) {
  return guarded(nullptr, [&] {
    NKX_REQUIRE(R, NKX_ERR_INVALID_ARGUMENT, "R is NULL");
    NKX_REQUIRE(scheme == 0 || scheme == 1, NKX_ERR_INVALID_ARGUMENT, "unknown radial scheme");
    *R = nkx::radial_scale(scheme, Z);
  });
}

extern "C++" {
namespace {
// Angular rules of a grid: 0 the full order; pruned grids (IntegratorXX's
// Treutler and Robust schemes, grid.hpp:164-166, 256-265) also 1 low (order
// 7) and 2 medium (order 11 for Treutler, the tabulated full order - 6 for
// Robust).  The base is the order actually used,
// next_algebraic_order(nang_order), as
// convergence_studies/convergence_pruning.cxx:130-134 describes.  Each entry
// is then looked up as the smallest tabulated order >= it.
void pruned_rule_orders(int nang_order, int pruning, int orders[3]) {
  int base = 0;
  nkx::lebedev_lookup(nang_order, nullptr, &base);
  const int lo = std::min(7, base);
  const int me = pruning == 1 ? std::min(11, base) : std::max(base - 6, 1);
  int a = 0;
  nkx::lebedev_lookup(lo, nullptr, &a);
  orders[0] = base;
  orders[1] = a;
  nkx::lebedev_lookup(me, nullptr, &a);
  orders[2] = a;
}

// make_flat_grid (grid.hpp:172-348) into out: TA / Becke radial x Lebedev per
// atom, capped Becke partition, compaction of w > w_thr.
void build_grid(nkx::Ctx &c, int natoms, const double *xyz, const int *Z, int nrad,
                int nang_order, int radial_scheme, double ta_alpha, double w_thr,
                int radial_sizing, nkx::GridState &out, int pruning = 0) {
  NKX_REQUIRE(natoms > 0 && xyz && Z && nrad > 0, NKX_ERR_INVALID_ARGUMENT,
              "make_flat_grid: need atoms, coordinates, charges and nrad > 0");
  NKX_REQUIRE(radial_sizing == 0 || radial_sizing == 1, NKX_ERR_INVALID_ARGUMENT,
              "unknown radial sizing");
  NKX_REQUIRE(pruning >= 0 && pruning <= 2, NKX_ERR_INVALID_ARGUMENT,
              "pruning: 0 none, 1 Treutler, 2 Robust");
  int orders[3];
  pruned_rule_orders(nang_order, pruning, orders);
  const int nrules = pruning ? 3 : 1;
  std::vector<double> ang;
  std::vector<int> rule_off, rule_n;
  double wang_min[3], wang_max[3];
  for (int k = 0; k < nrules; ++k) {
    int nk = 0, actual = 0;
    nkx::lebedev_lookup(orders[k], &nk, &actual);
    rule_off.push_back((int)(ang.size() / 4));
    rule_n.push_back(nk);
    ang.resize(ang.size() + 4 * (size_t)nk);
    double *u = ang.data() + 4 * (size_t)rule_off[k];
    nkx::lebedev_expand(actual, u);
    wang_min[k] = wang_max[k] = 0.0;
    for (int j = 0; j < nk; ++j) {
      wang_min[k] = std::min(wang_min[k], u[4 * j + 3]);
      wang_max[k] = std::max(wang_max[k], u[4 * j + 3]);
    }
  }
  // radial shells on the host (small); the product points on the device.
  // Counting shells from the nucleus: shells below nr/3 + 1 (Treutler) or
  // nr/4 + 1 (Robust) take rule 1, shells below nr/2 + 1 rule 2.
  std::vector<int> sh_atom, sh_rule;
  std::vector<double> sh_r, sh_w;
  std::vector<int64_t> poff(1, 0);
  for (int a = 0; a < natoms; ++a) {
    int nr = nrad;
    if (radial_sizing == NKX_SIZING_PYSCF) {
      static const double ratio[7] = {1.0, 1.5, 1.6, 1.8, 1.9, 2.0, 2.1};
      nr = (int)(nrad * ratio[nkx::periodic_row(Z[a]) - 1] + 0.5);
    }
    std::vector<double> r(nr), wr(nr);
    nkx::radial_rule(radial_scheme, nr, nkx::radial_scale(radial_scheme, Z[a]), ta_alpha,
                     r.data(), wr.data());
    const int b1 = (pruning == 1 ? nr / 3 : nr / 4) + 1, b2 = nr / 2 + 1;
    for (int i = 0; i < nr; ++i) {
      const int rank = r[0] < r[nr - 1] ? i : nr - 1 - i;
      const int k = pruning == 0 ? 0 : (rank < b1 ? 1 : (rank < b2 ? 2 : 0));
      NKX_REQUIRE(!((wr[i] < 0 && wang_max[k] > 0) || (wr[i] > 0 && wang_min[k] < 0)),
                  NKX_ERR_NEGATIVE_WEIGHT, "Detected negative weights aborting");
      sh_atom.push_back(a);
      sh_rule.push_back(k);
      sh_r.push_back(r[i]);
      sh_w.push_back(wr[i]);
      poff.push_back(poff.back() + rule_n[k]);
    }
  }
  const int64_t n = poff.back();
  const int nshell = (int)sh_atom.size();
  hipStream_t s = c.stream;
  nkx::DevBuf<double> p0, w0, xr, xw, xa, xc;
  nkx::DevBuf<int> o0, xat, xru, xro, xrn;
  nkx::DevBuf<int64_t> cnt, xpo;
  xr.upload(sh_r.data(), sh_r.size(), s);
  xw.upload(sh_w.data(), sh_w.size(), s);
  xat.upload(sh_atom.data(), sh_atom.size(), s);
  xru.upload(sh_rule.data(), sh_rule.size(), s);
  xpo.upload(poff.data(), poff.size(), s);
  xro.upload(rule_off.data(), rule_off.size(), s);
  xrn.upload(rule_n.data(), rule_n.size(), s);
  xa.upload(ang.data(), ang.size(), s);
  xc.upload(xyz, 3 * (size_t)natoms, s);
  p0.reserve(3 * n);
  w0.reserve(n);
  o0.reserve(n);
  nkx::launch_grid_shells(s, nshell, xat.p, xr.p, xw.p, xru.p, xpo.p, xro.p, xrn.p, xc.p, xa.p,
                          p0.p, w0.p, o0.p);
  out.centers_host.assign(xyz, xyz + 3 * natoms);
  out.centers.upload(xyz, 3 * (size_t)natoms, s);
  out.natoms = natoms;
  c.becke_rtab.reserve((size_t)natoms * natoms);
  nkx::launch_becke(s, natoms, out.centers.p, n, p0.p, o0.p, w0.p, nkx::BECKE_CAPPED,
                    c.becke_rtab.p);
  cnt.reserve((n + 1023) / 1024 + 1);
  out.pts.reserve(3 * n);
  out.w.reserve(n);
  out.owner.reserve(n);
  out.npts = nkx::compact_points(s, n, w_thr, p0.p, w0.p, o0.p, out.pts.p, out.w.p,
                                 out.owner.p, nullptr, cnt.p);
  out.version++;
}

nkx::BasisHost make_basis(int nbf, const int *n, const int *l, const int *m, const double *zeta,
                          const double *norm, const double *centers, const double *cutoff) {
  NKX_REQUIRE(nbf > 0 && n && l && m && zeta && norm && centers, NKX_ERR_INVALID_ARGUMENT,
              "set_basis: bad arguments");
  nkx::BasisHost B;
  B.nbf = nbf;
  for (int i = 0; i < nbf; ++i) {
    NKX_REQUIRE(n[i] >= 1 && l[i] >= 0 && l[i] <= 7 && l[i] < n[i] && std::abs(m[i]) <= l[i] &&
                    zeta[i] > 0,
                NKX_ERR_INVALID_ARGUMENT,
                "set_basis: need n >= 1, 0 <= l < n, l <= 7, |m| <= l, zeta > 0");
    B.n.push_back(n[i]);
    B.l.push_back(l[i]);
    B.m.push_back(m[i]);
    B.zeta.push_back(zeta[i]);
    B.norm.push_back(norm[i]);
    B.cx.push_back(centers[3 * i]);
    B.cy.push_back(centers[3 * i + 1]);
    B.cz.push_back(centers[3 * i + 2]);
    B.cutoff.push_back(cutoff ? cutoff[i] : 0.0);
    (void)nkx::solid_harmonic_terms(l[i], m[i]);
  }
  return B;
}
}  // namespace
}  // extern "C++"

int nkx_make_flat_grid_pruned(nkx_ctx *ctx, int natoms, const double *xyz, const int *Z,
                              int nrad, int nang_order, int radial_scheme, double ta_alpha,
                              double w_thr, int radial_sizing, int pruning, int64_t *npts_out) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    build_grid(ctx->c, natoms, xyz, Z, nrad, nang_order, radial_scheme, ta_alpha, w_thr,
               radial_sizing, ctx->c.grid, pruning);
    if (npts_out) *npts_out = ctx->c.grid.npts;
  });
}

int nkx_pruning_orders(int nang_order, int pruning, int *orders3) {
  return guarded(nullptr, [&] {
    NKX_REQUIRE(orders3, NKX_ERR_INVALID_ARGUMENT, "orders3 is NULL");
    NKX_REQUIRE(pruning >= 0 && pruning <= 2, NKX_ERR_INVALID_ARGUMENT,
                "pruning: 0 none, 1 Treutler, 2 Robust");
    pruned_rule_orders(nang_order, pruning, orders3);
  });
}

int nkx_make_flat_grid(nkx_ctx *ctx, int natoms, const double *xyz, const int *Z, int nrad,
                       int nang_order, int radial_scheme, double ta_alpha, double w_thr,
                       int radial_sizing, int64_t *npts_out) {
  return nkx_make_flat_grid_pruned(ctx, natoms, xyz, Z, nrad, nang_order, radial_scheme,
                                   ta_alpha, w_thr, radial_sizing, 0, npts_out);
}

int nkx_molset_create(nkx_ctx *ctx, int nmol, const int *natoms, const double *xyz,
                      const int *Z, int nrad, int nang_order, int radial_scheme,
                      double ta_alpha, double w_thr, int radial_sizing, const int *nbf,
                      const int *n, const int *l, const int *m, const double *zeta,
                      const double *norm, const double *centers, int64_t *npts_out) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    nkx::Ctx &c = ctx->c;
    NKX_REQUIRE(nmol > 0 && natoms && xyz && Z && nbf, NKX_ERR_INVALID_ARGUMENT,
                "molset_create: bad arguments");
    std::vector<std::vector<double>> pts(nmol), wts(nmol);
    std::vector<nkx::BasisHost> bases;
    size_t aoff = 0, foff = 0;
    for (int k = 0; k < nmol; ++k) {
      nkx::GridState g;
      build_grid(c, natoms[k], xyz + 3 * aoff, Z + aoff, nrad, nang_order, radial_scheme,
                 ta_alpha, w_thr, radial_sizing, g);
      pts[k].resize(3 * (size_t)g.npts);
      wts[k].resize(g.npts);
      NKX_HIP_CHECK(hipMemcpyAsync(pts[k].data(), g.pts.p, pts[k].size() * sizeof(double),
                                   hipMemcpyDeviceToHost, c.stream));
      NKX_HIP_CHECK(hipMemcpyAsync(wts[k].data(), g.w.p, wts[k].size() * sizeof(double),
                                   hipMemcpyDeviceToHost, c.stream));
      NKX_HIP_CHECK(hipStreamSynchronize(c.stream));
      NKX_REQUIRE(g.npts > 0, NKX_ERR_INVALID_ARGUMENT, "molset_create: a molecule has no points");
      if (npts_out) npts_out[k] = g.npts;
      bases.push_back(make_basis(nbf[k], n + foff, l + foff, m + foff, zeta + foff, norm + foff,
                                 centers + 3 * foff, nullptr));
      aoff += natoms[k];
      foff += nbf[k];
    }
    nkx::build_molset_plan(c, pts, wts, bases);
  });
}

int nkx_molset_vxc(nkx_ctx *ctx, const double *const *P, int nfunc, const int *ids,
                   int precision, double *E, double *N, double *const *V) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    nkx::Ctx &c = ctx->c;
    NKX_REQUIRE(c.mplan.nmol > 0, NKX_ERR_STATE, "no molecule set: call nkx_molset_create");
    nkx::XcList xl;
    bool gga = false;
    NKX_REQUIRE(nkx::xc_list_make(nfunc, ids, c.opt.dens_threshold, &xl, &gga),
                NKX_ERR_UNKNOWN_FUNCTIONAL, "unknown functional id or list longer than 4");
    NKX_REQUIRE(precision == NKX_PREC_FP64 || precision == NKX_PREC_MIXED,
                NKX_ERR_UNSUPPORTED, "molecule sets build in the FP64 or the int8 mixed form");
    NKX_REQUIRE(P && E && V, NKX_ERR_INVALID_ARGUMENT, "molset_vxc: bad arguments");
    const int nmol = c.mplan.nmol;
    for (int k = 0; k < nmol; ++k)
      NKX_REQUIRE(P[k] && V[k], NKX_ERR_INVALID_ARGUMENT, "molset_vxc: null P or V");
    std::vector<double> nel(nmol);
    nkx::build_on_plan(c, c.mplan, 1, P, xl, gga, precision, E, nel.data(), V);
    if (N)
      for (int k = 0; k < nmol; ++k) N[k] = nel[k];
  });
}

int nkx_set_grid(nkx_ctx *ctx, int natoms, const double *centers, int64_t npts,
                 const double *pts, const double *w, const int *owner) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    nkx::Ctx &c = ctx->c;
    NKX_REQUIRE(natoms > 0 && centers && npts >= 0 && (npts == 0 || (pts && w)),
                NKX_ERR_INVALID_ARGUMENT, "set_grid: bad arguments");
    hipStream_t s = c.stream;
    c.grid.natoms = natoms;
    c.grid.centers.reserve(3 * (size_t)natoms);
    NKX_HIP_CHECK(hipMemcpyAsync(c.grid.centers.p, centers, 3 * natoms * sizeof(double),
                                 hipMemcpyDeviceToDevice, s));
    c.grid.centers_host.resize(3 * natoms);
    NKX_HIP_CHECK(hipMemcpyAsync(c.grid.centers_host.data(), centers,
                                 3 * natoms * sizeof(double), hipMemcpyDeviceToHost, s));
    c.grid.pts.reserve(3 * npts);
    c.grid.w.reserve(npts);
    c.grid.owner.reserve(npts);
    if (npts) {
      NKX_HIP_CHECK(hipMemcpyAsync(c.grid.pts.p, pts, 3 * npts * sizeof(double),
                                   hipMemcpyDeviceToDevice, s));
      NKX_HIP_CHECK(hipMemcpyAsync(c.grid.w.p, w, npts * sizeof(double),
                                   hipMemcpyDeviceToDevice, s));
      if (owner)
        NKX_HIP_CHECK(hipMemcpyAsync(c.grid.owner.p, owner, npts * sizeof(int),
                                     hipMemcpyDeviceToDevice, s));
      else
        NKX_HIP_CHECK(hipMemsetAsync(c.grid.owner.p, 0, npts * sizeof(int), s));
    }
    NKX_HIP_CHECK(hipStreamSynchronize(s));
    c.grid.npts = npts;
    c.grid.version++;
  });
}

int nkx_grid_npts(nkx_ctx *ctx, int64_t *npts) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    NKX_REQUIRE(npts, NKX_ERR_INVALID_ARGUMENT, "npts is NULL");
    *npts = ctx->c.grid.npts;
  });
}

int nkx_grid_export(nkx_ctx *ctx, double *pts, double *w, int *owner) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    nkx::Ctx &c = ctx->c;
    const int64_t n = c.grid.npts;
    hipStream_t s = c.stream;
    if (n == 0) return;
    if (pts)
      NKX_HIP_CHECK(hipMemcpyAsync(pts, c.grid.pts.p, 3 * n * sizeof(double),
                                   hipMemcpyDeviceToDevice, s));
    if (w)
      NKX_HIP_CHECK(hipMemcpyAsync(w, c.grid.w.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (owner)
      NKX_HIP_CHECK(hipMemcpyAsync(owner, c.grid.owner.p, n * sizeof(int),
                                   hipMemcpyDeviceToDevice, s));
  });
}

int nkx_becke(nkx_ctx *ctx, int natoms, const double *centers, int64_t npts, const double *pts,
              const int *owner, double *w, int capped) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    NKX_REQUIRE(natoms > 0 && centers && (npts == 0 || (pts && owner && w)),
                NKX_ERR_INVALID_ARGUMENT, "becke: bad arguments");
    const int mode = capped == NKX_BECKE_CAPPED_UNSCREENED
                         ? nkx::BECKE_CAPPED | nkx::BECKE_NO_SCREEN
                         : (capped ? nkx::BECKE_CAPPED : 0);
    ctx->c.becke_rtab.reserve((size_t)natoms * natoms);
    nkx::launch_becke(ctx->c.stream, natoms, centers, npts, pts, owner, w, mode,
                      ctx->c.becke_rtab.p);
  });
}

int nkx_becke_noalloc(nkx_ctx *ctx, int natoms, const double *centers, int64_t npts,
                      const double *pts, const int *owner, double *w, int capped) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    NKX_REQUIRE(natoms > 0 && centers && (npts == 0 || (pts && owner && w)),
                NKX_ERR_INVALID_ARGUMENT, "becke_noalloc: bad arguments");
    NKX_REQUIRE(natoms <= nkx::BECKE_SCRATCH_ATOMS, NKX_ERR_INVALID_ARGUMENT,
                "becke_noalloc: at most 512 atoms (its fixed table holds 512 x 512 "
                "inter-atomic distances); use nkx_becke for more");
    double *rt = nkx::becke_scratch(ctx->c.device);
    NKX_REQUIRE(rt, NKX_ERR_STATE, "becke_noalloc: scratch table not resolved on this device");
    const int mode = capped == NKX_BECKE_CAPPED_UNSCREENED
                         ? nkx::BECKE_CAPPED | nkx::BECKE_NO_SCREEN
                         : (capped ? nkx::BECKE_CAPPED : 0);
    nkx::launch_becke(ctx->c.stream, natoms, centers, npts, pts, owner, w, mode, rt);
  });
}

int nkx_sto_norm_cutoff(int n, double zeta, double tol, double *norm, double *cutoff) {
  return guarded(nullptr, [&] {
    NKX_REQUIRE(norm && cutoff, NKX_ERR_INVALID_ARGUMENT, "output is NULL");
    nkx::sto_norm_cutoff(n, zeta, tol, norm, cutoff);
  });
}

int nkx_set_basis(nkx_ctx *ctx, int nbf, const int *n, const int *l, const int *m,
                  const double *zeta, const double *norm, const double *centers,
                  const double *cutoff) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    ctx->c.basis = make_basis(nbf, n, l, m, zeta, norm, centers, cutoff);
    ctx->c.basis_version++;
  });
}

int nkx_basis_nbf(nkx_ctx *ctx, int *nbf) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    NKX_REQUIRE(nbf, NKX_ERR_INVALID_ARGUMENT, "nbf is NULL");
    *nbf = ctx->c.basis.nbf;
  });
}

int nkx_solid_harmonic(int l, int m, double x, double y, double z, double *out4) {
  return guarded(nullptr, [&] {
    NKX_REQUIRE(out4, NKX_ERR_INVALID_ARGUMENT, "out4 is NULL");
    const auto &t = nkx::solid_harmonic_terms(l, m);
    double v = 0, gx = 0, gy = 0, gz = 0;
    auto pw = [](double b, int e) {
      double r = 1.0;
      for (int i = 0; i < e; ++i) r *= b;
      return r;
    };
    for (const auto &h : t) {
      v += h.coef * pw(x, h.a) * pw(y, h.b) * pw(z, h.c);
      if (h.a) gx += h.coef * h.a * pw(x, h.a - 1) * pw(y, h.b) * pw(z, h.c);
      if (h.b) gy += h.coef * h.b * pw(x, h.a) * pw(y, h.b - 1) * pw(z, h.c);
      if (h.c) gz += h.coef * h.c * pw(x, h.a) * pw(y, h.b) * pw(z, h.c - 1);
    }
    out4[0] = v;
    out4[1] = gx;
    out4[2] = gy;
    out4[3] = gz;
  });
}

int nkx_collocation(nkx_ctx *ctx, int64_t npts, const double *pts, double *phi, double *gx,
                    double *gy, double *gz) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    nkx::Ctx &c = ctx->c;
    NKX_REQUIRE(c.basis.nbf > 0, NKX_ERR_STATE, "no basis: call nkx_set_basis");
    NKX_REQUIRE(npts >= 0 && (npts == 0 || (pts && (phi || gx))), NKX_ERR_INVALID_ARGUMENT,
                "collocation: bad arguments");
    NKX_REQUIRE((!gx && !gy && !gz) || (gx && gy && gz), NKX_ERR_INVALID_ARGUMENT,
                "collocation: pass all three gradient arrays or none");
    // Function records in basis order (no permutation).
    const nkx::BasisHost &B = c.basis;
    std::vector<nkx::FuncRec> fr(B.nbf);
    std::vector<nkx::HarmTerm> terms;
    for (int i = 0; i < B.nbf; ++i) {
      const auto &t = nkx::solid_harmonic_terms(B.l[i], B.m[i]);
      fr[i] = nkx::FuncRec{B.cx[i], B.cy[i], B.cz[i], B.zeta[i], B.norm[i],
                           B.n[i] - B.l[i] - 1, (int)terms.size(), (int)t.size(), 0};
      terms.insert(terms.end(), t.begin(), t.end());
    }
    nkx::DevBuf<nkx::FuncRec> fd;
    nkx::DevBuf<nkx::HarmTerm> td;
    fd.upload(fr.data(), fr.size(), c.stream);
    td.upload(terms.data(), terms.size(), c.stream);
    nkx::launch_collocation_ref(c.stream, B.nbf, fd.p, td.p, npts, pts, phi, gx, gy, gz);
    NKX_HIP_CHECK(hipStreamSynchronize(c.stream));
  });
}

int nkx_density(nkx_ctx *ctx, int nbf, int64_t npts, const double *phi, const double *gx,
                const double *gy, const double *gz, int nocc, const double *C, const double *f,
                double *rho, double *grx, double *gry, double *grz, double *sigma) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    NKX_REQUIRE(nbf > 0 && npts >= 0 && nocc >= 0 && phi && C && f && rho,
                NKX_ERR_INVALID_ARGUMENT, "density: bad arguments");
    NKX_REQUIRE((!gx && !gy && !gz) || (gx && gy && gz), NKX_ERR_INVALID_ARGUMENT,
                "density: pass all three gradient arrays or none");
    nkx::launch_density_ref(ctx->c.stream, nbf, npts, phi, gx, gy, gz, nocc, C, f, rho,
                            gx ? grx : nullptr, gry, grz, gx ? sigma : nullptr);
  });
}

int nkx_functional_info(int id, int *family, int *kind) {
  int fam = 0, k = 0;
  if (functional_info(id, &fam, &k)) {
    g_thread_error = "unknown functional id " + std::to_string(id);
    return NKX_ERR_UNKNOWN_FUNCTIONAL;
  }
  if (family) *family = fam;
  if (kind) *kind = k;
  return NKX_OK;
}

int nkx_set_dens_threshold(nkx_ctx *ctx, double thr) {
  NKX_CTX_CHECK(ctx);
  ctx->c.opt.dens_threshold = thr;
  return NKX_OK;
}

int nkx_functional_eval(nkx_ctx *ctx, int nfunc, const int *ids, int64_t npts,
                        const double *rho, const double *sigma, double *exc, double *vrho,
                        double *vsigma) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    nkx::XcList xl;
    bool gga = false;
    NKX_REQUIRE(nkx::xc_list_make(nfunc, ids, ctx->c.opt.dens_threshold, &xl, &gga),
                NKX_ERR_UNKNOWN_FUNCTIONAL, "unknown functional id or list longer than 4");
    NKX_REQUIRE(!gga || (sigma && vsigma), NKX_ERR_FUNCTIONAL_FAMILY,
                "GGA functionals need sigma and vsigma");
    NKX_REQUIRE(npts >= 0 && (npts == 0 || (rho && exc && vrho)), NKX_ERR_INVALID_ARGUMENT,
                "functional_eval: bad arguments");
    nkx::launch_functional_eval(ctx->c.stream, xl, npts, rho, sigma, exc, vrho, vsigma);
  });
}

extern "C++" {
namespace {
// compute_lda / compute_gga family checks (xc_integrals.hpp:85-91, :437-447)
void family_require(int family_check, int nfunc, const int *ids, bool gga) {
  if (family_check == NKX_FAMILY_LDA)
    NKX_REQUIRE(!gga, NKX_ERR_FUNCTIONAL_FAMILY,
                "Provided functional is not part of the LDA family");
  if (family_check == NKX_FAMILY_GGA) {
    bool all_gga = true;
    for (int k = 0; k < nfunc; ++k) {
      int fam, kind;
      functional_info(ids[k], &fam, &kind);
      all_gga = all_gga && fam == NKX_FAMILY_GGA;
    }
    NKX_REQUIRE(all_gga, NKX_ERR_FUNCTIONAL_FAMILY,
                "Provided funtional is not a part of the GGA or Hybrid GGA family");
  }
}
}  // namespace
}  // extern "C++"

int nkx_functional_eval_polarized(nkx_ctx *ctx, int nfunc, const int *ids, int64_t npts,
                                  const double *rho, const double *sigma, double *exc,
                                  double *vrho, double *vsigma) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    nkx::XcList xl;
    bool gga = false;
    NKX_REQUIRE(nkx::xc_list_make(nfunc, ids, ctx->c.opt.dens_threshold, &xl, &gga),
                NKX_ERR_UNKNOWN_FUNCTIONAL, "unknown functional id or list longer than 4");
    NKX_REQUIRE(!gga || (sigma && vsigma), NKX_ERR_FUNCTIONAL_FAMILY,
                "GGA functionals need sigma and vsigma");
    NKX_REQUIRE(npts >= 0 && (npts == 0 || (rho && exc && vrho)), NKX_ERR_INVALID_ARGUMENT,
                "functional_eval_polarized: bad arguments");
    nkx::launch_functional_eval_pol(ctx->c.stream, xl, npts, rho, gga ? sigma : nullptr, exc,
                                    vrho, gga ? vsigma : nullptr);
  });
}

int nkx_xc_from_collocation_polarized(nkx_ctx *ctx, int nbf, int64_t npts, const double *phi,
                                      const double *gx, const double *gy, const double *gz,
                                      const double *w, int nocca, const double *Ca,
                                      const double *fa, int noccb, const double *Cb,
                                      const double *fb, int nfunc, const int *ids,
                                      int family_check, double *E, double *Va, double *Vb) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    nkx::XcList xl;
    bool gga = false;
    NKX_REQUIRE(nkx::xc_list_make(nfunc, ids, ctx->c.opt.dens_threshold, &xl, &gga),
                NKX_ERR_UNKNOWN_FUNCTIONAL, "unknown functional id or list longer than 4");
    family_require(family_check, nfunc, ids, gga);
    NKX_REQUIRE(!gga || (gx && gy && gz), NKX_ERR_INVALID_ARGUMENT,
                "GGA functionals need the collocation gradients");
    NKX_REQUIRE(nbf > 0 && npts >= 0 && nocca >= 0 && noccb >= 0 && phi && w && E && Va && Vb &&
                    (nocca == 0 || (Ca && fa)) && (noccb == 0 || (Cb && fb)),
                NKX_ERR_INVALID_ARGUMENT, "xc_from_collocation_polarized: bad arguments");
    nkx::xc_from_collocation_pol(ctx->c, nbf, npts, phi, gx, gy, gz, w, nocca, Ca, fa, noccb, Cb,
                                 fb, xl, gga, E, Va, Vb);
  });
}

int nkx_xc_from_collocation(nkx_ctx *ctx, int nbf, int64_t npts, const double *phi,
                            const double *gx, const double *gy, const double *gz,
                            const double *w, int nocc, const double *C, const double *f,
                            int nfunc, const int *ids, int family_check, double *E, double *V) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    nkx::XcList xl;
    bool gga = false;
    NKX_REQUIRE(nkx::xc_list_make(nfunc, ids, ctx->c.opt.dens_threshold, &xl, &gga),
                NKX_ERR_UNKNOWN_FUNCTIONAL, "unknown functional id or list longer than 4");
    family_require(family_check, nfunc, ids, gga);
    NKX_REQUIRE(!gga || (gx && gy && gz), NKX_ERR_INVALID_ARGUMENT,
                "GGA functionals need the collocation gradients");
    NKX_REQUIRE(nbf > 0 && npts >= 0 && phi && w && C && f && E && V,
                NKX_ERR_INVALID_ARGUMENT, "xc_from_collocation: bad arguments");
    nkx::xc_from_collocation(ctx->c, nbf, npts, phi, gx, gy, gz, w, nocc, C, f, xl, gga, E, V);
  });
}

int nkx_vxc(nkx_ctx *ctx, const double *P, int nfunc, const int *ids, int precision, double *E,
            double *N, double *V) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    nkx::XcList xl;
    bool gga = false;
    NKX_REQUIRE(nkx::xc_list_make(nfunc, ids, ctx->c.opt.dens_threshold, &xl, &gga),
                NKX_ERR_UNKNOWN_FUNCTIONAL, "unknown functional id or list longer than 4");
    NKX_REQUIRE(precision == NKX_PREC_FP64 || precision == NKX_PREC_MIXED ||
                    precision == NKX_PREC_MIXED_BF16,
                NKX_ERR_INVALID_ARGUMENT, "unknown precision");
    NKX_REQUIRE(P && E && N && V, NKX_ERR_INVALID_ARGUMENT, "vxc: bad arguments");
    nkx::vxc_build(ctx->c, P, xl, gga, precision, E, N, V);
  });
}

int nkx_vxc_polarized(nkx_ctx *ctx, const double *Pa, const double *Pb, int nfunc,
                      const int *ids, int precision, double *E, double *N, double *Va,
                      double *Vb) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    nkx::XcList xl;
    bool gga = false;
    NKX_REQUIRE(nkx::xc_list_make(nfunc, ids, ctx->c.opt.dens_threshold, &xl, &gga),
                NKX_ERR_UNKNOWN_FUNCTIONAL, "unknown functional id or list longer than 4");
    NKX_REQUIRE(precision == NKX_PREC_FP64 || precision == NKX_PREC_MIXED ||
                    precision == NKX_PREC_MIXED_BF16,
                NKX_ERR_INVALID_ARGUMENT, "unknown precision");
    NKX_REQUIRE(precision != NKX_PREC_MIXED_BF16, NKX_ERR_UNSUPPORTED,
                "the spin-polarised build is available in the FP64 and int8 forms");
    NKX_REQUIRE(Pa && Pb && E && Va && Vb && Va != Vb, NKX_ERR_INVALID_ARGUMENT,
                "vxc_polarized: bad arguments");
    const double *P[2] = {Pa, Pb};
    double *V[2] = {Va, Vb};
    double n2[2];
    nkx::vxc_build_spin(ctx->c, 2, P, xl, gga, precision, E, n2, V);
    if (N) {
      N[0] = n2[0];
      N[1] = n2[1];
    }
  });
}

static int xc_grad_entry(nkx_ctx *ctx, int nspin, const double *const *P, int nfunc,
                         const int *ids, int partition, int precision, double *E,
                         double *grad) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    NKX_REQUIRE(precision == NKX_PREC_FP64 || precision == NKX_PREC_MIXED,
                NKX_ERR_INVALID_ARGUMENT, "xc_grad: precision must be FP64 or MIXED");
    const bool mixed = precision == NKX_PREC_MIXED;
    nkx::XcList xl;
    bool gga = false;
    NKX_REQUIRE(nkx::xc_list_make(nfunc, ids, ctx->c.opt.dens_threshold, &xl, &gga),
                NKX_ERR_UNKNOWN_FUNCTIONAL, "unknown functional id or list longer than 4");
    NKX_REQUIRE(grad, NKX_ERR_INVALID_ARGUMENT, "xc_grad: bad arguments");
    for (int sp = 0; sp < nspin; ++sp)
      NKX_REQUIRE(P[sp], NKX_ERR_INVALID_ARGUMENT, "xc_grad: bad arguments");
    double e = 0.0;
    nkx::xc_grad(ctx->c, nspin, P, xl, gga, partition, mixed, &e, grad);
    if (E) *E = e;
  });
}

int nkx_xc_grad(nkx_ctx *ctx, const double *P, int nfunc, const int *ids, int partition,
                double *E, double *grad) {
  return xc_grad_entry(ctx, 1, &P, nfunc, ids, partition, NKX_PREC_FP64, E, grad);
}

int nkx_xc_grad_mixed(nkx_ctx *ctx, const double *P, int nfunc, const int *ids, int partition,
                      double *E, double *grad) {
  return xc_grad_entry(ctx, 1, &P, nfunc, ids, partition, NKX_PREC_MIXED, E, grad);
}

int nkx_xc_grad_polarized(nkx_ctx *ctx, const double *Pa, const double *Pb, int nfunc,
                          const int *ids, int partition, int precision, double *E,
                          double *grad) {
  const double *P[2] = {Pa, Pb};
  return xc_grad_entry(ctx, 2, P, nfunc, ids, partition, precision, E, grad);
}

int nkx_bounding_boxes(nkx_ctx *ctx, int64_t npts, double *points, double *weights, int *owner,
                       int max_points_per_box, double *boxes) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    NKX_REQUIRE(npts >= 0 && max_points_per_box > 0 && (npts == 0 || (points && boxes)),
                NKX_ERR_INVALID_ARGUMENT, "bounding_boxes: bad arguments");
    nkx::sparse_bounding_boxes(ctx->c.stream, npts, points, weights, owner, max_points_per_box,
                               boxes);
  });
}

int nkx_neighbor_list(nkx_ctx *ctx, int64_t nboxes, const double *boxes, int nbf,
                      const double *centers, const double *radii, int *offsets, int *neighbors,
                      int64_t *total) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    NKX_REQUIRE(nboxes >= 0 && nbf >= 0 && offsets && (nboxes == 0 || boxes) &&
                    (nbf == 0 || (centers && radii)),
                NKX_ERR_INVALID_ARGUMENT, "neighbor_list: bad arguments");
    const int64_t t = nkx::sparse_neighbor_list(ctx->c.stream, nboxes, boxes, nbf, centers,
                                                radii, offsets, neighbors);
    if (total) *total = t;
  });
}

int nkx_density_matrix(nkx_ctx *ctx, int nbf, int nocc, const double *C, const double *f,
                       double *P) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    NKX_REQUIRE(nbf >= 0 && nocc >= 0 && P && (nocc == 0 || (C && f)),
                NKX_ERR_INVALID_ARGUMENT, "density_matrix: bad arguments");
    nkx::vk_density_matrix(ctx->c.stream, nbf, nocc, C, f, P);
  });
}

int nkx_density_points(nkx_ctx *ctx, const double *P, double *rho, double *gx, double *gy,
                       double *gz, double *sigma) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    NKX_REQUIRE(P && rho, NKX_ERR_INVALID_ARGUMENT, "density_points: P and rho are required");
    nkx::density_points(ctx->c, P, rho, gx, gy, gz, sigma);
  });
}

int nkx_graph_capture(nkx_ctx *ctx, const double *P, int nfunc, const int *ids, int precision,
                      double *V) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    nkx::XcList xl;
    bool gga = false;
    NKX_REQUIRE(nkx::xc_list_make(nfunc, ids, ctx->c.opt.dens_threshold, &xl, &gga),
                NKX_ERR_UNKNOWN_FUNCTIONAL, "unknown functional id or list longer than 4");
    NKX_REQUIRE(precision == NKX_PREC_FP64 || precision == NKX_PREC_MIXED ||
                    precision == NKX_PREC_MIXED_BF16,
                NKX_ERR_INVALID_ARGUMENT, "unknown precision");
    NKX_REQUIRE(P && V, NKX_ERR_INVALID_ARGUMENT, "graph_capture: bad arguments");
    nkx::graph_capture(ctx->c, 1, &P, xl, gga, precision, &V);
  });
}

int nkx_graph_capture_polarized(nkx_ctx *ctx, const double *Pa, const double *Pb, int nfunc,
                                const int *ids, int precision, double *Va, double *Vb) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    nkx::XcList xl;
    bool gga = false;
    NKX_REQUIRE(nkx::xc_list_make(nfunc, ids, ctx->c.opt.dens_threshold, &xl, &gga),
                NKX_ERR_UNKNOWN_FUNCTIONAL, "unknown functional id or list longer than 4");
    NKX_REQUIRE(precision == NKX_PREC_FP64 || precision == NKX_PREC_MIXED,
                NKX_ERR_UNSUPPORTED, "the open-shell build takes FP64 or NKX_PREC_MIXED");
    NKX_REQUIRE(Pa && Pb && Va && Vb && Va != Vb, NKX_ERR_INVALID_ARGUMENT,
                "graph_capture_polarized: bad arguments");
    const double *P[2] = {Pa, Pb};
    double *V[2] = {Va, Vb};
    nkx::graph_capture(ctx->c, 2, P, xl, gga, precision, V);
  });
}

int nkx_molset_graph_capture(nkx_ctx *ctx, const double *const *P, int nfunc, const int *ids,
                             int precision, double *const *V) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    nkx::Ctx &c = ctx->c;
    NKX_REQUIRE(c.mplan.nmol > 0, NKX_ERR_STATE, "no molecule set: call nkx_molset_create");
    nkx::XcList xl;
    bool gga = false;
    NKX_REQUIRE(nkx::xc_list_make(nfunc, ids, c.opt.dens_threshold, &xl, &gga),
                NKX_ERR_UNKNOWN_FUNCTIONAL, "unknown functional id or list longer than 4");
    NKX_REQUIRE(precision == NKX_PREC_FP64 || precision == NKX_PREC_MIXED,
                NKX_ERR_UNSUPPORTED, "molecule sets build in the FP64 or the int8 mixed form");
    NKX_REQUIRE(P && V, NKX_ERR_INVALID_ARGUMENT, "molset_graph_capture: bad arguments");
    for (int k = 0; k < c.mplan.nmol; ++k)
      NKX_REQUIRE(P[k] && V[k], NKX_ERR_INVALID_ARGUMENT, "molset_graph_capture: null P or V");
    nkx::molset_graph_capture(c, P, xl, gga, precision, V);
  });
}

int nkx_graph_launch(nkx_ctx *ctx, double *E, double *N) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    NKX_REQUIRE(E, NKX_ERR_INVALID_ARGUMENT, "graph_launch: E is NULL");
    nkx::graph_launch(ctx->c, E, N);
  });
}

int nkx_autotune(nkx_ctx *ctx, const double *P, int nfunc, const int *ids, int precision,
                 double *V, int *vgemm_split, int *density_group, double *best_ms) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    nkx::XcList xl;
    bool gga = false;
    NKX_REQUIRE(nkx::xc_list_make(nfunc, ids, ctx->c.opt.dens_threshold, &xl, &gga),
                NKX_ERR_UNKNOWN_FUNCTIONAL, "unknown functional id or list longer than 4");
    NKX_REQUIRE(precision == NKX_PREC_FP64 || precision == NKX_PREC_MIXED ||
                    precision == NKX_PREC_MIXED_BF16,
                NKX_ERR_INVALID_ARGUMENT, "unknown precision");
    NKX_REQUIRE(P && V, NKX_ERR_INVALID_ARGUMENT, "autotune: bad arguments");
    const double ms = nkx::autotune(ctx->c, P, xl, gga, precision, V, vgemm_split, density_group);
    if (best_ms) *best_ms = ms;
  });
}

int nkx_molset_autotune(nkx_ctx *ctx, const double *const *P, int nfunc, const int *ids,
                        int precision, double *const *V, int *vgemm_split, int *density_group,
                        double *best_ms) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    nkx::Ctx &c = ctx->c;
    NKX_REQUIRE(c.mplan.nmol > 0, NKX_ERR_STATE, "no molecule set: call nkx_molset_create");
    nkx::XcList xl;
    bool gga = false;
    NKX_REQUIRE(nkx::xc_list_make(nfunc, ids, c.opt.dens_threshold, &xl, &gga),
                NKX_ERR_UNKNOWN_FUNCTIONAL, "unknown functional id or list longer than 4");
    NKX_REQUIRE(precision == NKX_PREC_FP64 || precision == NKX_PREC_MIXED,
                NKX_ERR_INVALID_ARGUMENT, "molecule sets take FP64 or NKX_PREC_MIXED");
    NKX_REQUIRE(P && V, NKX_ERR_INVALID_ARGUMENT, "molset_autotune: bad arguments");
    for (int k = 0; k < c.mplan.nmol; ++k)
      NKX_REQUIRE(P[k] && V[k], NKX_ERR_INVALID_ARGUMENT, "molset_autotune: null P or V");
    const double ms =
        nkx::molset_autotune(c, P, xl, gga, precision, V, vgemm_split, density_group);
    if (best_ms) *best_ms = ms;
  });
}

int nkx_set_launch(nkx_ctx *ctx, int vgemm_split, int density_shape) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    NKX_REQUIRE(vgemm_split >= 0, NKX_ERR_INVALID_ARGUMENT, "vgemm_split must be >= 0");
    const int sh = density_shape;
    NKX_REQUIRE(sh == 0 || sh == 1 || sh == 2 || sh == 4 || nkx::mixed_density_shape_ok(sh),
                NKX_ERR_INVALID_ARGUMENT, "density shape must be 0, 1, 2, 4, 65, 66, 71 or 72");
    ctx->c.opt.vgemm_split = vgemm_split;
    ctx->c.opt.density_group = density_shape;
  });
}

int nkx_set_memory_budget(nkx_ctx *ctx, int64_t bytes) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    NKX_REQUIRE(bytes >= 0, NKX_ERR_INVALID_ARGUMENT, "memory budget must be >= 0");
    ctx->c.opt.mem_budget = (size_t)bytes;
  });
}

int nkx_get_launch(nkx_ctx *ctx, int *vgemm_split, int *density_shape) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    if (vgemm_split) *vgemm_split = ctx->c.opt.vgemm_split;
    if (density_shape) *density_shape = ctx->c.opt.density_group;
  });
}

int nkx_launch_points(int precision, int density_shape, int *points) {
  return guarded(nullptr, [&] {
    NKX_REQUIRE(points, NKX_ERR_INVALID_ARGUMENT, "points is NULL");
    NKX_REQUIRE(precision >= NKX_PREC_FP64 && precision <= NKX_PREC_MIXED_BF16,
                NKX_ERR_INVALID_ARGUMENT, "unknown precision");
    const int sh = density_shape;
    int batches = 1;
    if (precision == NKX_PREC_FP64)  // 1, 2, 4 batches; the int8 codes as the build reads them
      batches = sh == 66 ? 2 : (sh >= 65 ? 1 : (sh == 1 || sh == 2 ? sh : (sh == 0 ? 2 : 4)));
    else if (precision == NKX_PREC_MIXED)
      batches = sh == 66 ? 2 : 1;
    *points = 64 * batches;
  });
}

int nkx_set_screening(nkx_ctx *ctx, double tol, int batch_points) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    NKX_REQUIRE(tol > 0 && tol < 1, NKX_ERR_INVALID_ARGUMENT, "screening tolerance in (0, 1)");
    NKX_REQUIRE(batch_points == 0 || batch_points == 64, NKX_ERR_UNSUPPORTED,
                "batch size 64 is the only one built");
    ctx->c.opt.screen_tol = tol;
  });
}

int nkx_plan_stats(nkx_ctx *ctx, int64_t *npts, int64_t *npad, int64_t *nb, int64_t *kept,
                   int64_t *total) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    nkx::ensure_plan(ctx->c);
    const nkx::Plan &p = ctx->c.plan;
    if (npts) *npts = p.npts;
    if (npad) *npad = p.npts_pad;
    if (nb) *nb = p.nb;
    if (kept) *kept = p.pairs_kept;
    if (total) *total = p.nb * p.nfb;
  });
}

int nkx_set_timing(nkx_ctx *ctx, int enable) {
  NKX_CTX_CHECK(ctx);
  ctx->c.timing.on = enable != 0;
  return NKX_OK;
}

int nkx_reset_timing(nkx_ctx *ctx) {
  NKX_CTX_CHECK(ctx);
  for (int k = 0; k < nkx::kNumStages; ++k) {
    ctx->c.timing.ms[k] = 0;
    ctx->c.timing.launches[k] = 0;
  }
  return NKX_OK;
}

int nkx_get_timing(nkx_ctx *ctx, int stage, double *ms, int64_t *launches) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    NKX_REQUIRE(stage >= 0 && stage < nkx::kNumStages, NKX_ERR_INVALID_ARGUMENT,
                "unknown timing stage");
    if (ms) *ms = ctx->c.timing.ms[stage];
    if (launches) *launches = ctx->c.timing.launches[stage];
  });
}

int nkx_plan_work(nkx_ctx *ctx, double *fd, double *fv, double *ce) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    nkx::ensure_plan(ctx->c);
    const nkx::Plan &p = ctx->c.plan;
    if (fd) *fd = 2.0 * p.work_nb2g;
    if (fv) *fv = 2.0 * p.work_nb2g;
    if (ce) *ce = p.work_nbg;
  });
}

int nkx_overlap(nkx_ctx *ctx, double *S) {
  NKX_CTX_CHECK(ctx);
  return guarded(ctx, [&] {
    NKX_REQUIRE(S, NKX_ERR_INVALID_ARGUMENT, "S is NULL");
    nkx::overlap_build(ctx->c, S);
  });
}

}  // extern "C"
