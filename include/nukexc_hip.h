/*
 * nukexc_hip.h -- C ABI of the MI355X-native NuKEXC exchange-correlation path.
 *
 * The reference is a header-only C++ library over Kokkos views; it has no ABI.
 * This header is the stable boundary a host driver binds instead (C++ through
 * the headers under include/nukexc/, or any FFI: ctypes, cgo, JNI).  Each entry point names
 * the reference interface it replaces (paths relative to the NuKEXC tree).
 *
 * Conventions
 *  - Every function returns an nkx_status; no C++ exception crosses the ABI.
 *    nkx_last_error(ctx) gives the message of the last failure on that ctx
 *    (ctx == NULL: the calling thread's last failure, e.g. from nkx_create).
 *  - "_dev" arguments are device pointers on the ctx's GPU; "_host" arguments
 *    are host pointers.  Device work is ordered on the ctx's stream; calls
 *    that return host scalars (energies, point counts) synchronise that stream.
 *  - A ctx is bound to one device and one stream and owns all workspaces, so
 *    repeated V_xc builds (one per SCF iteration) reuse memory and launch
 *    plans.  One ctx per host thread.
 *  - Matrices are row-major unless stated.  Collocation arrays passed to the
 *    reference-shaped entry points use the reference's LayoutLeft (N_bf, N_pts)
 *    view: element (mu, g) at mu + g * N_bf.
 *  - Functional ids follow libxc numbering: LDA_X = 1, LDA_C_PW = 12,
 *    LDA_C_PW_MOD = 13, LDA_C_VWN = 7, GGA_X_PBE = 101, GGA_C_PBE = 130,
 *    GGA_X_B88 = 106, GGA_C_LYP = 131, GGA_X_PW91 = 109, GGA_C_PW91 = 134
 *    (tests/standalone_helpers.hpp:184-225).
 */
#ifndef NUKEXC_HIP_H
#define NUKEXC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nkx_ctx nkx_ctx;

typedef enum nkx_status {
  NKX_OK = 0,
  NKX_ERR_INVALID_ARGUMENT = 1,
  NKX_ERR_HIP = 2,
  NKX_ERR_OUT_OF_MEMORY = 3,
  NKX_ERR_FUNCTIONAL_FAMILY = 4, /* xc_integrals.hpp:81-84, :434-438 */
  NKX_ERR_NSPIN = 5,             /* xc_integrals.hpp:86-90 */
  NKX_ERR_NEGATIVE_WEIGHT = 6,   /* grid.hpp:284-286 */
  NKX_ERR_UNKNOWN_FUNCTIONAL = 7,
  NKX_ERR_STATE = 8,             /* call-order errors (no grid / basis yet) */
  NKX_ERR_UNSUPPORTED = 9
} nkx_status;

typedef enum nkx_precision {
  NKX_PREC_FP64 = 0,       /* FP64 collocation and FP64 MFMA contractions */
  NKX_PREC_MIXED = 1,      /* operands as four int8 slices with per-row power-of-two
                              scales, exact int32 MFMA accumulation; FP64
                              collocation arithmetic, functional and reductions */
  NKX_PREC_MIXED_BF16 = 2  /* operands as three bf16 terms (split-BF16), six
                              products per contraction on FP32-accumulating MFMA */
} nkx_precision;

typedef enum nkx_radial_scheme {
  NKX_RADIAL_TREUTLER_AHLRICHS = 0, /* grid.hpp:104-108, TA_XI :98-102 */
  NKX_RADIAL_BECKE = 1              /* half Bragg-Slater radius, :71-96 */
} nkx_radial_scheme;

typedef enum nkx_radial_sizing {
  NKX_SIZING_UNIFORM = 0, /* grid.hpp:129-133 */
  NKX_SIZING_PYSCF = 1    /* grid.hpp:135-148 */
} nkx_radial_sizing;

typedef enum nkx_family { NKX_FAMILY_LDA = 1, NKX_FAMILY_GGA = 2 } nkx_family;

/* ---------------------------------------------------------------- context */
/* hip_stream == NULL: the ctx creates its own blocking stream, which is
   ordered with the legacy default stream. */
int nkx_create(int device, void *hip_stream, nkx_ctx **out);
int nkx_destroy(nkx_ctx *ctx);
int nkx_sync(nkx_ctx *ctx);
const char *nkx_last_error(const nkx_ctx *ctx);
int nkx_version(int *major, int *minor, int *patch);

/* Device memory helpers so C/C++ callers need no HIP headers
   (the reference's Kokkos::View allocation / deep_copy). */
int nkx_malloc(nkx_ctx *ctx, size_t bytes, void **ptr_dev);
int nkx_free(nkx_ctx *ctx, void *ptr_dev);
int nkx_memcpy_h2d(nkx_ctx *ctx, void *dst_dev, const void *src_host, size_t bytes);
int nkx_memcpy_d2h(nkx_ctx *ctx, void *dst_host, const void *src_dev, size_t bytes);
int nkx_memset(nkx_ctx *ctx, void *dst_dev, int value, size_t bytes);
/* Stream-ordered device-to-device copy by a plain 16-byte vector load/store
   kernel (no copy engine); pointers and size 16-byte aligned.  Also the
   known-size kernel against which HBM traffic counters are calibrated. */
int nkx_copy_d2d(nkx_ctx *ctx, void *dst_dev, const void *src_dev, size_t bytes);

/* ------------------------------------------------------- angular / radial */
/* Lebedev rule for the smallest tabulated algebraic order >= order
   (IntegratorXX next_algebraic_order, grid.hpp:186-187). */
int nkx_lebedev_npts(int order, int *npts, int *actual_order);
/* Fills xyzw_host (npts rows of x, y, z, w; sum w = 4 pi). */
int nkx_lebedev_rule(int order, double *xyzw_host);
/* Radial rule with r^2 and the Jacobian folded into w (IntegratorXX radial
   quadratures called at grid.hpp:247-251). */
int nkx_radial_rule(int scheme, int n, double R, double alpha, double *r_host,
                    double *w_host);
/* Per-element radial scale (grid.hpp:227-235). */
int nkx_radial_scale(int scheme, int Z, double *R);

/* ------------------------------------------------------------------- grid */
/* make_flat_grid<TA|Becke, LebedevLaikov>(mol, nrad, nang_order, w_thr,
   ta_alpha, Unpruned, sizing)  (grid.hpp:172-348).  The grid is kept in the
   ctx: atom-major, radial, angular order; Becke-partitioned with the capped
   partition (partition_becke_team); points with w <= w_thr removed by an
   order-preserving compaction.  Returns the number of points. */
/* nkx_make_flat_grid with angular pruning (grid.hpp:164-166, 256-265):
   pruning 0 none, 1 Treutler (radial shells counted from the nucleus: below
   nrad/3 + 1 Lebedev order 7, below nrad/2 + 1 order 11, the rest the full
   order), 2 Robust (below nrad/4 + 1 order 7, below nrad/2 + 1 the full
   order - 6).  The point counts match the reference's pruning study. */
int nkx_make_flat_grid_pruned(nkx_ctx *ctx, int natoms, const double *xyz, const int *Z,
                              int nrad, int nang_order, int radial_scheme, double ta_alpha,
                              double w_thr, int radial_sizing, int pruning, int64_t *npts_out);
/* Lebedev orders a grid of this nang_order uses (host; no GPU needed):
   orders3[0] the full rule, next_algebraic_order(nang_order)
   (grid.hpp:186-187); for pruning 1 (Treutler) / 2 (Robust) orders3[1] the
   inner-shell rule (order 7) and orders3[2] the middle-shell rule (11 for
   Treutler, for Robust the full rule's order - 6, as
   convergence_studies/convergence_pruning.cxx:130-134 states), each the
   smallest tabulated order at or above it. */
int nkx_pruning_orders(int nang_order, int pruning, int *orders3);
int nkx_make_flat_grid(nkx_ctx *ctx, int natoms, const double *xyz_host,
                       const int *Z_host, int nrad, int nang_order, int radial_scheme,
                       double ta_alpha, double weight_threshold, int radial_sizing,
                       int64_t *npts);
/* Replace the ctx grid with caller data (device pointers, copied). */
int nkx_set_grid(nkx_ctx *ctx, int natoms, const double *centers_dev, int64_t npts,
                 const double *points_dev, const double *weights_dev,
                 const int *owner_dev);
int nkx_grid_npts(nkx_ctx *ctx, int64_t *npts);
/* Copy the ctx grid out: points (npts x 3), weights, owners (any may be NULL). */
int nkx_grid_export(nkx_ctx *ctx, double *points_dev, double *weights_dev,
                    int *owner_dev);

/* partition_becke (mode NKX_BECKE_PLAIN, partitioning.hpp:63-120) and
   partition_becke_team (NKX_BECKE_CAPPED, :126-213): weights_dev[g] *=
   P_owner(g) / sum_i P_i on caller arrays.  points_dev is npts x 3.  The
   capped form skips atoms and factors whose contribution is exactly 0 or 1;
   NKX_BECKE_CAPPED_UNSCREENED evaluates all of them (same result bit for
   bit; for checking that).  Any other nonzero mode counts as capped. */
enum { NKX_BECKE_PLAIN = 0, NKX_BECKE_CAPPED = 1, NKX_BECKE_CAPPED_UNSCREENED = 3 };
int nkx_becke(nkx_ctx *ctx, int natoms, const double *centers_dev, int64_t npts,
              const double *points_dev, const int *owner_dev, double *weights_dev,
              int capped);
/* nkx_becke without any device allocation (no counterpart in the reference):
   for callers that cannot allocate -- code under HIP graph capture, serving
   loops, harnesses that forbid hipMalloc.  Same arguments and results (bit
   for bit) as nkx_becke; its inter-atomic distance table lives in a fixed
   per-GPU device array, so natoms <= 512 (larger is NKX_ERR_INVALID_ARGUMENT).
   Every call on one GPU shares that array: calls that overlap on one GPU
   (other streams or contexts) race.  Stream-ordered. */
int nkx_becke_noalloc(nkx_ctx *ctx, int natoms, const double *centers_dev, int64_t npts,
                      const double *points_dev, const int *owner_dev, double *weights_dev,
                      int capped);

/* ------------------------------------------------------------------ basis */
/* make_manual_basis normalisation and compute_cutoff radius
   (stobasis.hpp:90-173, :200-201). */
int nkx_sto_norm_cutoff(int n, double zeta, double tol, double *norm, double *cutoff);
/* STOBasisSet (stobasis.hpp:54-64) for the ctx; host arrays, copied.
   l <= 7. centers_host is nbf x 3. */
int nkx_set_basis(nkx_ctx *ctx, int nbf, const int *n, const int *l, const int *m,
                  const double *zeta, const double *norm, const double *centers_host,
                  const double *cutoff);
int nkx_basis_nbf(nkx_ctx *ctx, int *nbf);
/* Real solid harmonic S_lm = r^l Y_lm and its gradient at (x, y, z), l <= 7
   (real_solid_harmonic_cart_and_grad_precomputed, spherical_harmonics.hpp:
   750-1608); host.  out4 = {S, dS/dx, dS/dy, dS/dz}. */
int nkx_solid_harmonic(int l, int m, double x, double y, double z, double *out4);

/* ------------------------------------------------------------ collocation */
/* fill_collocation / fill_grad_collocation (stobasis.hpp:494-584) on caller
   points (npts x 3): LayoutLeft (N_bf, N_pts) outputs; grad pointers may be
   NULL.  FP64. */
int nkx_collocation(nkx_ctx *ctx, int64_t npts, const double *points_dev,
                    double *phi_dev, double *gx_dev, double *gy_dev, double *gz_dev);

/* --------------------------------------------------------------- density */
/* compute_density / compute_density_and_sigma (density.hpp:150-255) on
   caller collocation arrays (LayoutLeft (N_bf, N_pts)) and orbitals
   C (LayoutLeft (N_bf, N_occ)), occupations f.  gx/gy/gz NULL: rho only. */
int nkx_density(nkx_ctx *ctx, int nbf, int64_t npts, const double *phi_dev,
                const double *gx_dev, const double *gy_dev, const double *gz_dev,
                int nocc, const double *C_dev, const double *f_dev, double *rho_dev,
                double *grx_dev, double *gry_dev, double *grz_dev, double *sigma_dev);

/* ------------------------------------------------------------- functional */
/* libxc family of an id (func.info->family); kind 0 exchange, 1 correlation. */
int nkx_functional_info(int id, int *family, int *kind);
/* Sets the density threshold (libxc xc_func_set_dens_threshold, used by
   standalone.cxx:336).  thr <= 0 restores each functional's default. */
int nkx_set_dens_threshold(nkx_ctx *ctx, double thr);
/* Sum over the listed functionals of exc (per particle), vrho, vsigma on
   device arrays (the host libxc calls at xc_integrals.hpp:113, :471-472, done
   on the GPU).  sigma/vsigma may be NULL for an LDA-only list. */
int nkx_functional_eval(nkx_ctx *ctx, int nfunc, const int *func_ids, int64_t npts,
                        const double *rho_dev, const double *sigma_dev,
                        double *exc_dev, double *vrho_dev, double *vsigma_dev);

/* Spin-polarised evaluation (libxc XC_POLARIZED, the host calls at
   xc_integrals.hpp:203 and :617, done on the GPU).  Interleaved layouts:
   rho[2g+s], sigma[3g+k] (k: aa, ab, bb), vrho[2g+s], vsigma[3g+k]; exc per
   particle.  Inputs are clamped as libxc does (rho_s >= thr,
   sigma_ss >= thr^(8/3), |sigma_ab| <= (sigma_aa + sigma_bb)/2). */
int nkx_functional_eval_polarized(nkx_ctx *ctx, int nfunc, const int *func_ids,
                                  int64_t npts, const double *rho_dev,
                                  const double *sigma_dev, double *exc_dev,
                                  double *vrho_dev, double *vsigma_dev);

/* ----------------------------------------------------- reference-shaped XC */
/* compute_lda (gx == NULL, xc_integrals.hpp:73-166) and compute_gga
   (:424-541): E_xc and V_xc (row-major N_bf x N_bf, device) from caller
   collocation arrays, weights, orbitals and occupations.  family_check:
   0 = no check, 1 = require LDA functionals, 2 = require GGA functionals. */
int nkx_xc_from_collocation(nkx_ctx *ctx, int nbf, int64_t npts, const double *phi_dev,
                            const double *gx_dev, const double *gy_dev,
                            const double *gz_dev, const double *w_dev, int nocc,
                            const double *C_dev, const double *f_dev, int nfunc,
                            const int *func_ids, int family_check, double *E_xc,
                            double *V_dev);

/* compute_lsda (gx == NULL, xc_integrals.hpp:169-262) and compute_gga_lsda
   (:543-718): E_xc and the two spin potentials V_alpha, V_beta from caller
   collocation arrays, weights and per-spin orbitals / occupations. */
int nkx_xc_from_collocation_polarized(nkx_ctx *ctx, int nbf, int64_t npts,
                                      const double *phi_dev, const double *gx_dev,
                                      const double *gy_dev, const double *gz_dev,
                                      const double *w_dev, int nocc_a, const double *Ca_dev,
                                      const double *fa_dev, int nocc_b, const double *Cb_dev,
                                      const double *fb_dev, int nfunc, const int *func_ids,
                                      int family_check, double *E_xc, double *Va_dev,
                                      double *Vb_dev);

/* ---------------------------------------------------------------- hot path */
/* One V_xc build on the ctx grid and basis (the SCF-iteration XC call,
   tests/standalone.cxx:724-741 -> compute_gga): density from the density
   matrix P (row-major N_bf x N_bf, symmetric, device), on-device functional,
   E_xc = sum_g w rho eps, N_el = sum_g w rho, and
   V_xc = Phi^T diag(w v) Phi (+ GGA terms), row-major N_bf x N_bf (device).
   Points are processed in spatially compact batches against screened
   function lists; precision selects the FP64 or the mixed form. */
int nkx_vxc(nkx_ctx *ctx, const double *P_dev, int nfunc, const int *func_ids,
            int precision, double *E_xc, double *N_el, double *V_dev);
/* Spin-polarised build (the open-shell SCF iteration, tests/standalone.cxx:
   321-360 -> standalone_helpers.hpp:229-256): spin density matrices Pa, Pb,
   E_xc = sum_g w (rho_a + rho_b) eps, N_el[0..1] = per-spin electron counts
   (may be NULL), V_alpha and V_beta (row-major N_bf x N_bf, device).
   precision: NKX_PREC_FP64 or NKX_PREC_MIXED (int8 slices: one density
   contraction and one V contraction per spin, the polarised functional in
   FP64); NKX_PREC_MIXED_BF16 is rejected with NKX_ERR_UNSUPPORTED. */
int nkx_vxc_polarized(nkx_ctx *ctx, const double *Pa_dev, const double *Pb_dev, int nfunc,
                      const int *func_ids, int precision, double *E_xc, double *N_el,
                      double *Va_dev, double *Vb_dev);
/* ------------------------------------------------------- screened grid */
/* create_bounding_boxes (octree.hpp:71-138): permutes the points (device,
   3 doubles per point), weights and owners (device; either may be NULL) in
   place into Morton order (63-bit codes, 21 bits per axis of the position in
   the points' bounding box; equal codes keep their order) and writes the
   bounding box of each run of max_points_per_box consecutive points to boxes
   (device, 6 doubles per box: min x, y, z, max x, y, z; ceil(npts /
   max_points_per_box) boxes, the last one ragged).  Data derived from the old
   point order is invalid afterwards, as in the reference.  Synchronous. */
int nkx_bounding_boxes(nkx_ctx *ctx, int64_t npts, double *points_dev, double *weights_dev,
                       int *owner_dev, int max_points_per_box, double *boxes_dev);
/* build_neighbor_list (octree.hpp:317-389): for each box, the ascending
   indices of the basis functions whose sphere (centre, cutoff radius; host
   arrays, 3 and 1 doubles per function) meets the box (distance from the
   centre to the box <= radius).  offsets (device, nboxes + 1 ints) is always
   written; neighbors (device, offsets[nboxes] ints) when not NULL; *total
   (may be NULL) = offsets[nboxes].  Call with neighbors NULL to size the list.
   Synchronous. */
int nkx_neighbor_list(nkx_ctx *ctx, int64_t nboxes, const double *boxes_dev, int nbf,
                      const double *centers_host, const double *radii_host, int *offsets_dev,
                      int *neighbors_dev, int64_t *total);
/* P = C diag(f) C^T (C LayoutLeft N_bf x N_occ, f N_occ, P N_bf x N_bf; all
   device): the density matrix of the (C, f) arguments of the reference's
   density and XC calls (density.hpp:150-186). */
int nkx_density_matrix(nkx_ctx *ctx, int nbf, int nocc, const double *C_dev, const double *f_dev,
                       double *P_dev);
/* rho, grad rho (x, y, z) and sigma = |grad rho|^2 of the density matrix P
   at every point of the ctx grid, in the grid's order (the outputs of
   compute_density_and_sigma(_sparse), density.hpp:188-328), by the batched
   FP64 path with its own screening; every output but rho may be NULL.
   Stream-ordered. */
int nkx_density_points(nkx_ctx *ctx, const double *P_dev, double *rho_dev, double *gx_dev,
                       double *gy_dev, double *gz_dev, double *sigma_dev);

/* Nuclear gradient dE_xc/dR_A at fixed P (FP64; no counterpart in the
   reference, which has no gradients): basis-centre, grid-point-motion and
   partition-weight terms for a grid whose points move with their owner atom.
   partition: 0 = weights held fixed, 1 = capped Becke (the partition of
   nkx_make_flat_grid, partitioning.hpp:126-213), 2 = Becke
   (partitioning.hpp:63-120).  Every basis function must be centred on an atom
   of the ctx grid.  E_xc (may be NULL) and grad (host, 3 * natoms, atom-major
   x, y, z) are written on return. */
int nkx_xc_grad(nkx_ctx *ctx, const double *P_dev, int nfunc, const int *func_ids,
                int partition, double *E_xc, double *grad_host);
/* The same gradient in the mixed form (no counterpart in the reference): the
   density and the basis-centre contractions X = P Phi, W = P Y run on int8
   slices with exact integer MFMA accumulation (as NKX_PREC_MIXED builds V);
   the functional, the STO derivatives and every sum stay FP64.  Arguments
   and results as nkx_xc_grad; E_xc is the mixed-form energy. */
int nkx_xc_grad_mixed(nkx_ctx *ctx, const double *P_dev, int nfunc, const int *func_ids,
                      int partition, double *E_xc, double *grad_host);
/* Open-shell nuclear gradient (no counterpart in the reference): E_xc of the
   spin density matrices Pa, Pb (device, nbf x nbf) with the polarised
   functional, differentiated as nkx_xc_grad.  precision: NKX_PREC_FP64, or
   NKX_PREC_MIXED for the int8 contractions of nkx_xc_grad_mixed (one density
   and one basis-centre contraction per spin); anything else is
   NKX_ERR_INVALID_ARGUMENT. */
int nkx_xc_grad_polarized(nkx_ctx *ctx, const double *Pa_dev, const double *Pb_dev, int nfunc,
                          const int *func_ids, int partition, int precision, double *E_xc,
                          double *grad_host);
/* Molecule set (no counterpart in the reference, whose driver handles one
   molecule): nmol molecules of different sizes, each with its own grid built
   as nkx_make_flat_grid does (shared grid parameters) and its own basis.
   natoms[k]; xyz (3 per atom) and Z concatenated over molecules; nbf[k]; the
   basis arrays (as nkx_set_basis) concatenated.  npts_out (nmol, may be NULL)
   receives each grid's size.  Replaces any previous set on the ctx. */
int nkx_molset_create(nkx_ctx *ctx, int nmol, const int *natoms, const double *xyz,
                      const int *Z, int nrad, int nang_order, int radial_scheme,
                      double ta_alpha, double w_thr, int radial_sizing, const int *nbf,
                      const int *n, const int *l, const int *m, const double *zeta,
                      const double *norm, const double *centers, int64_t *npts_out);
/* One V_xc build over every molecule of the set: all batches of all molecules
   go through each kernel launch together (FP64 or NKX_PREC_MIXED).
   P_dev[k] and V_dev[k] (host arrays of device pointers, nbf[k]^2 each),
   E_xc[k], N_el[k] (may be NULL). */
int nkx_molset_vxc(nkx_ctx *ctx, const double *const *P_dev, int nfunc, const int *func_ids,
                   int precision, double *E_xc, double *N_el, double *const *V_dev);
/* HIP graph of one nkx_vxc build (no counterpart in the reference): captures
   every launch of a build on the ctx grid and basis for these P and V device
   pointers (nothing runs during the capture; the graph has its own
   workspace).  nkx_graph_launch replays it and returns E_xc and N_el (N may be
   NULL); P's contents may change between replays, as across SCF iterations.
   A later grid, basis or screening change makes nkx_graph_launch fail with
   NKX_ERR_STATE until the next capture. */
int nkx_graph_capture(nkx_ctx *ctx, const double *P_dev, int nfunc, const int *func_ids,
                      int precision, double *V_dev);
/* Open-shell form of nkx_graph_capture (no counterpart in the reference):
   one nkx_vxc_polarized build (FP64 or NKX_PREC_MIXED) for these spin
   density matrices and potentials; nkx_graph_launch then returns E_xc and,
   when N is not NULL, N[0] = N_alpha and N[1] = N_beta. */
int nkx_graph_capture_polarized(nkx_ctx *ctx, const double *Pa_dev, const double *Pb_dev,
                                int nfunc, const int *func_ids, int precision, double *Va_dev,
                                double *Vb_dev);
/* The same for one nkx_molset_vxc build of the ctx molecule set (no
   counterpart in the reference): P_dev[m], V_dev[m] per molecule; replays
   with nkx_graph_launch, which then fills E_xc[m] and N_el[m] for every
   molecule.  A later nkx_molset_create makes the replay fail with
   NKX_ERR_STATE. */
int nkx_molset_graph_capture(nkx_ctx *ctx, const double *const *P_dev, int nfunc,
                             const int *func_ids, int precision, double *const *V_dev);
/* Replays the last capture; E_xc and N_el hold one value per captured
   molecule (one for nkx_graph_capture). */
int nkx_graph_launch(nkx_ctx *ctx, double *E_xc, double *N_el);
/* Autotuner (no counterpart in the reference): on the ctx grid and basis for
   this P, times one nkx_vxc build per candidate -- first the density
   workgroup shape (FP64: 1, 2 or 4 point batches of 64 points per
   workgroup; int8 form: 65 and 66 for one and two batches per workgroup,
   the workgroups of one batch group on one XCD; 71 and 72 for one batch on
   absolute block pairs dealt to XCDs in tiles of 4 x 16 / 8 x 8 (pairs x
   batches); the autotuner times 1, 2, 4 or 65, 71), then the split
   count of the V contraction over batches (the automatic choice x 1/4 .. 4)
   -- keeps the fastest for later builds and captures, and returns them
   (either pointer may be NULL) with the best build time.  V receives the
   last build's potential. */
int nkx_autotune(nkx_ctx *ctx, const double *P_dev, int nfunc, const int *func_ids,
                 int precision, double *V_dev, int *vgemm_split, int *density_group,
                 double *best_ms);
/* The same search for the molecule set of the ctx (no counterpart in the
   reference): candidates timed on nkx_molset_vxc builds of every molecule
   (P_dev[k], V_dev[k] as there; FP64 or NKX_PREC_MIXED).  The pair found is
   kept in the ctx options, so later set builds and set graph captures use
   it. */
int nkx_molset_autotune(nkx_ctx *ctx, const double *const *P_dev, int nfunc,
                        const int *func_ids, int precision, double *const *V_dev,
                        int *vgemm_split, int *density_group, double *best_ms);
/* Launch parameters (no counterpart in the reference): set or read the
   values the autotuner picks, e.g. to reuse a tuned pair without timing
   again.  vgemm_split 0 = automatic; density_shape 0 = default, else one of
   1, 2, 4 (FP64) or 65, 66, 71, 72 (int8 form); the FP64 form reads 65 and
   71 as 1 and 66 as 2, the int8 form reads 1, 2 and 4 as 65. */
int nkx_set_launch(nkx_ctx *ctx, int vgemm_split, int density_shape);
/* Device-memory budget of a build's tile store in bytes (no counterpart in
   the reference): a build whose tiles do not fit runs in chunks of batches.
   0 (default) = min(48 GB, 40 % of the free memory at build time); values
   below 64 MB count as 64 MB.  For many contexts on one GPU. */
int nkx_set_memory_budget(nkx_ctx *ctx, int64_t bytes);
int nkx_get_launch(nkx_ctx *ctx, int *vgemm_split, int *density_shape);
/* Grid points per density workgroup (64, 128 or 256: the point-batch size
   the autotuner chose) that a build of the given precision runs with this
   density shape (0 = the default shape); no counterpart in the reference. */
int nkx_launch_points(int precision, int density_shape, int *points);
/* Screening tolerance on |Phi| for the batch function lists (default 1e-12)
   and the point-batch size: 64 is the only size (the kernels' tiles are 64
   points x 64 functions); 0 keeps it, any other value is
   NKX_ERR_UNSUPPORTED. */
int nkx_set_screening(nkx_ctx *ctx, double tol, int batch_points);
/* Statistics of the current plan: number of points, padded points, batches,
   screened (function-block, batch) pairs kept, total pairs. */
int nkx_plan_stats(nkx_ctx *ctx, int64_t *npts, int64_t *npts_padded, int64_t *nbatches,
                   int64_t *pairs_kept, int64_t *pairs_total);
/* Kernel timing (the reference driver's TimingRegistry,
   tests/standalone_helpers.hpp:71-156, at kernel granularity): when enabled,
   nkx_vxc records HIP events on the ctx stream around each launch.  Stages:
   0 collocation, 1 density contraction (+ functional and Z in the fused
   forms), 2 V contraction, 3 gather / reductions, 4 whole build, 5 functional
   and Z of the int8 split form.  ms is the summed event time since the last
   reset, launches the number of timed launches. */
int nkx_set_timing(nkx_ctx *ctx, int enable);
int nkx_reset_timing(nkx_ctx *ctx);
int nkx_get_timing(nkx_ctx *ctx, int stage, double *ms, int64_t *launches);
/* Algorithmic work of one nkx_vxc build on the current plan, counting true
   points g_b and true screened functions n_b per batch b:
   density flops 2 sum_b n_b^2 g_b, V flops 2 sum_b n_b^2 g_b,
   collocation elements sum_b n_b g_b. */
int nkx_plan_work(nkx_ctx *ctx, double *flops_density, double *flops_v,
                  double *colloc_elems);
/* Numerical overlap S = Phi diag(w) Phi^T on the ctx grid (FP64, row-major,
   device); core_hamiltonian.hpp:62-119 overlap_integral. */
int nkx_overlap(nkx_ctx *ctx, double *S_dev);

#ifdef __cplusplus
}
#endif

#endif /* NUKEXC_HIP_H */
