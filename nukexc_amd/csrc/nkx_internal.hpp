// Internal declarations shared by the host sources and HIP kernels of
// libnukexc_hip.so.  Nothing here is part of the ABI (include/nukexc_hip.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "nukexc_hip.h"

namespace nkx {

// Errors raised inside the library; converted to nkx_status at the ABI.
struct Error {
  int code;
  std::string msg;
};

#define NKX_HIP_CHECK(call)                                                   \
  do {                                                                        \
    hipError_t nkx_e_ = (call);                                               \
    if (nkx_e_ != hipSuccess)                                                 \
      throw ::nkx::Error{nkx_e_ == hipErrorOutOfMemory ? NKX_ERR_OUT_OF_MEMORY \
                                                       : NKX_ERR_HIP,         \
                         std::string(#call) + ": " + hipGetErrorString(nkx_e_)}; \
  } while (0)

#define NKX_REQUIRE(cond, code, msg)                                          \
  do {                                                                        \
    if (!(cond)) throw ::nkx::Error{(code), (msg)};                           \
  } while (0)

// Owning device buffer; grows on demand and never shrinks.
template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void reserve(size_t count) {
    if (count <= n) return;
    release();
    if (count == 0) return;
    NKX_HIP_CHECK(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  void upload(const T *host, size_t count, hipStream_t s) {
    reserve(count);
    if (count) NKX_HIP_CHECK(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
};

// STO basis on the device, one record per function.
struct BasisHost {
  int nbf = 0;
  std::vector<int> n, l, m;
  std::vector<double> zeta, norm, cx, cy, cz, cutoff;
};

// Solid-harmonic polynomial terms: S_lm = sum_t coef_t x^a y^b z^c.
struct HarmTerm {
  double coef;
  int a, b, c, pad;
};

// Per-function record consumed by the collocation kernels.
struct FuncRec {
  double cx, cy, cz, zeta, norm;
  int k;          // n - l - 1, power of r in the radial part
  int term_off;   // first HarmTerm of this function
  int term_cnt;   // number of HarmTerms
  int pad;
};

struct GridState {
  int natoms = 0;
  std::vector<double> centers_host;  // natoms x 3
  DevBuf<double> centers;            // natoms x 3
  int64_t npts = 0;
  DevBuf<double> pts;                // npts x 3
  DevBuf<double> w;
  DevBuf<int> owner;
  uint64_t version = 0;              // bumped on every change
};

// Per-molecule tables of a plan, passed to kernels by value.  Blocks and
// batches of all molecules are numbered globally (molecule by molecule);
// a block pair and a batch's live blocks always belong to one molecule.
struct MolTab {
  const int *blk_mol, *blk_local;  // per global block: molecule, block index in it
  const int *ld;                   // per molecule: padded function count (P row stride)
  const int *nfb;                  // per molecule: function blocks
  const int64_t *b0, *nb;          // per molecule: first global batch, batch count
  const int64_t *ts_off;           // per molecule: offset of its tile_slot table [nfb][nb]
  const int64_t *P_off;            // per molecule: offset of its gathered FP64 P
  const int64_t *Ps_off;           // per molecule: byte offset of its int8 P sections
};

// Launch plan of the batched V_xc build; rebuilt when grid, basis or the
// screening options change.  A plan holds one molecule (the ctx grid and
// basis) or a molecule set.
struct Plan {
  uint64_t grid_version = ~0ull, basis_version = ~0ull;
  double screen_tol = -1;
  int pt = 0;                 // points per batch
  int fb = 64;                // functions per function block
  int nbf = 0, nbf_pad = 0, nfb = 0;       // totals over molecules
  int64_t npts = 0, npts_pad = 0, nb = 0;  // nb = number of point batches
  std::vector<int> fperm;     // global sorted position -> original function index in its molecule
  DevBuf<int> fperm_d;
  DevBuf<FuncRec> funcs;      // nbf_pad records in sorted order (padding: norm 0)
  DevBuf<HarmTerm> terms;
  DevBuf<double> px, py, pz, pw;  // sorted, padded point SoA
  DevBuf<int> pidx;               // sorted position -> original point in its molecule (-1: pad)
  // batch -> active function blocks (CSR)
  std::vector<int> b_off_h;
  DevBuf<int> b_off, b_fbl, slot_b;  // slot s = CSR position; slot_b[s] = its batch
  // function-block pairs (I <= J, same molecule) with a common live batch
  int npairs = 0;
  std::vector<int> pair_I_h, pair_J_h;
  DevBuf<int> pair_I, pair_J;
  // 2 x 2 super-pairs of blocks (one molecule) for the int8 V contraction:
  // 4 ints each, blocks I0 I1 J0 J1 (-1: none) and the pair indices of the
  // tiles (I0,J0) (I0,J1) (I1,J0) (I1,J1) (-1: not a kept pair, or redundant)
  int nspairs = 0;
  int max_mol_nfb = 0;            // largest function-block count of one molecule
  DevBuf<int> sp_blk, sp_pidx;
  DevBuf<int> tile_slot;          // per molecule [block][batch] -> slot in the tile store or -1
  int64_t nslots = 0;
  int64_t pairs_kept = 0;
  double work_nb2g = 0, work_nbg = 0;  // sum_b n_b^2 g_b, sum_b n_b g_b
  // molecules
  int nmol = 0;
  std::vector<int> mol_nbf_h, mol_ld_h, mol_nfb_h, mol_blk0_h;
  std::vector<int64_t> mol_b0_h, mol_nb_h, mol_npts_h, mol_P_off_h, mol_Ps_off_h;
  int64_t P_total = 0;            // sum of ld^2 (one spin of gathered FP64 P)
  DevBuf<int> blk_mol, blk_local, mol_ld, mol_nfb, mol_nbf;
  DevBuf<int64_t> mol_b0, mol_nb, mol_ts_off, mol_P_off, mol_Ps_off;
  MolTab tab() const {
    return MolTab{blk_mol.p, blk_local.p, mol_ld.p, mol_nfb.p, mol_b0.p,
                  mol_nb.p,  mol_ts_off.p, mol_P_off.p, mol_Ps_off.p};
  }
};

struct Options {
  double screen_tol = 1e-12;
  int batch_points = 64;
  int vgemm_split = 0;          // V-contraction split count; 0 = automatic
  int density_group = 0;        // density workgroup shape code; 0 = form default
  double dens_threshold = 0.0;  // <= 0: per-functional libxc default
  size_t mem_budget = 0;        // tile-store bytes per build; 0: min(48 GB, 40 % of free)
};

struct Workspace {
  DevBuf<double> Pd;               // permuted padded density matrix (FP64)
  DevBuf<signed char> Ps;          // int8 slice sections of the permuted P
  DevBuf<double> tiles64;          // Phi, gx, gy, gz, Z tiles (FP64)
  DevBuf<signed char> mstore;      // mixed-form tile store (int8 slices + FP32 gradients)
  DevBuf<double> coef;             // per padded point and spin: a, bx, by, bz
  DevBuf<double> blockEN;          // per batch: E, N_alpha, N_beta partial sums
  DevBuf<double> vpart;            // split-K partial V tiles
  DevBuf<double> dpart;            // per spin and slot: column sums of the density contraction (int8 split form)
  DevBuf<signed char> zstore_b;    // int8 form, second spin: Z sections per slot
  DevBuf<double> red;              // small reduction outputs
  DevBuf<double> scratch;
  DevBuf<double> gbuf;             // nuclear-gradient buffers
  DevBuf<int> gint;
  std::vector<double *> vptr_h;    // output pointer table of a build (host copy)
  DevBuf<double *> vptr;
};

// Per-stage kernel timing with HIP events on the ctx stream
// (the reference's driver-level TimingRegistry, standalone_helpers.hpp:71-156,
// at kernel granularity).
enum Stage { kStageColloc = 0, kStageDensity = 1, kStageVgemm = 2, kStageFinal = 3,
             kStageBuild = 4, kStageZ = 5, kNumStages = 6 };
struct Timing {
  bool on = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
  std::vector<int> pending_stage;  // stage of each recorded pair this build
  double ms[kNumStages] = {0, 0, 0, 0, 0};
  int64_t launches[kNumStages] = {0, 0, 0, 0, 0};
  // Records the start of a timed launch; returns its slot (-1 when off).
  int begin(int stage, hipStream_t s);
  void end(int slot, hipStream_t s);
  void collect();  // after a stream synchronisation
  void release();
  ~Timing() { release(); }
};

// Algorithmic work of one build on the current plan (true functions and true
// points only): density contraction flops 2 sum_b n_b^2 g_b, V contraction
// flops 2 sum_b n_b^2 g_b, collocation elements sum_b n_b g_b.
struct Work {
  double flops_density = 0, flops_v = 0, colloc_elems = 0;
};

// A captured build (nkx_graph_capture): its own workspace, so that builds
// outside the graph never reallocate memory the graph's kernels use.
struct GraphState {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double *en_pinned = nullptr;  // (E, N_alpha, N_beta) per molecule, written by the graph
  int nmol = 1;
  int nspin = 1;                // 2: an open-shell capture (one molecule)
  bool molset = false;          // captured on the molecule-set plan
  uint64_t mplan_version = 0;
  std::vector<double *> vtab_h;
  DevBuf<double *> vtab;
  Workspace ws;
  uint64_t grid_version = ~0ull, basis_version = ~0ull;
  double screen_tol = -1;
  void release();
  ~GraphState() { release(); }
};

// Sizes of one build (prepare_build).
struct BuildSetup {
  int nspin = 1, narr = 0, nsplit = 1, group = 4;
  bool gga = false, bf16 = false, mixed = false;
  int64_t slots_alloc = 0;
  int64_t psb = 0;  // int8 form: bytes of one spin's P sections (all molecules)
  size_t vsz = 0;
};

struct Ctx {
  int device = 0;
  Timing timing;
  Work work;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string last_error;
  Options opt;
  GridState grid;
  BasisHost basis;
  uint64_t basis_version = 0;
  Plan plan;
  Plan mplan;                    // plan of the molecule set (nkx_molset_create)
  uint64_t mplan_version = 0;    // bumped by every nkx_molset_create
  Workspace ws;
  GraphState graph;
  DevBuf<double> becke_rtab;     // inter-atomic distances of the last partition
};

// ---- host helpers (basis.cpp, lebedev.cpp, grid.cpp) --------------------
void sto_norm_cutoff(int n, double zeta, double tol, double *norm, double *cut);
// Monomial expansion of the real solid harmonic (l, m), l <= 7.
const std::vector<HarmTerm> &solid_harmonic_terms(int l, int m);
int lebedev_lookup(int order, int *npts, int *actual_order);
void lebedev_expand(int actual_order, double *xyzw);
void radial_rule(int scheme, int n, double R, double alpha, double *r, double *w);
double radial_scale(int scheme, int Z);
int periodic_row(int Z);

// ---- kernels (launchers) -------------------------------------------------
// radial shell x angular rule points, one workgroup per shell (kernels_basic.hip)
void launch_grid_shells(hipStream_t s, int nshell, const int *sh_atom, const double *sh_r,
                        const double *sh_w, const int *sh_rule, const int64_t *sh_poff,
                        const int *rule_off, const int *rule_n, const double *xyz,
                        const double *ang, double *P, double *W, int *O);
// Becke partition modes: BECKE_CAPPED (R~ = min(R, 5), partition_becke_team)
// or plain; BECKE_NO_SCREEN evaluates every factor of the capped form (the
// screened and unscreened forms agree bit for bit).  rtab: natoms^2 doubles.
constexpr int BECKE_CAPPED = 1, BECKE_NO_SCREEN = 2;
void launch_becke(hipStream_t s, int natoms, const double *centers, int64_t npts,
                  const double *pts, const int *owner, double *w, int mode, double *rtab);
// The allocation-free partition's table: at most BECKE_SCRATCH_ATOMS atoms.
// becke_scratch_init resolves the device array's address on the current
// device (once per device); becke_scratch returns it (NULL before that).
constexpr int BECKE_SCRATCH_ATOMS = 512;
void becke_scratch_init(int device);
double *becke_scratch(int device);
int64_t compact_points(hipStream_t s, int64_t n, double thr, const double *pts_in,
                       const double *w_in, const int *own_in, double *pts_out,
                       double *w_out, int *own_out, int *scratch_flags,
                       int64_t *scratch_scan);

void launch_collocation_ref(hipStream_t s, int nbf, const FuncRec *f,
                            const HarmTerm *t, int64_t npts, const double *pts,
                            double *phi, double *gx, double *gy, double *gz);

// Functionals: kind bitmask of ids (device constant table).
struct XcList {
  int n;
  int ids[4];
  double thr[4];
};
bool xc_list_make(int nfunc, const int *ids, double dens_thr_override, XcList *out,
                  bool *any_gga);
void launch_functional_eval(hipStream_t s, XcList xl, int64_t npts, const double *rho,
                            const double *sigma, double *exc, double *vrho,
                            double *vsigma);

void launch_copy16(hipStream_t s, void *dst, const void *src, size_t bytes);
void launch_functional_eval_pol(hipStream_t s, XcList xl, int64_t npts, const double *rho,
                                const double *sigma, double *exc, double *vrho,
                                double *vsigma);

void launch_density_ref(hipStream_t s, int nbf, int64_t npts, const double *phi,
                        const double *gx, const double *gy, const double *gz, int nocc,
                        const double *C, const double *f, double *rho, double *grx,
                        double *gry, double *grz, double *sigma);
void xc_from_collocation(Ctx &c, int nbf, int64_t npts, const double *phi,
                         const double *gx, const double *gy, const double *gz,
                         const double *w, int nocc, const double *C, const double *f,
                         const XcList &xl, bool gga, double *E, double *V);

void xc_from_collocation_pol(Ctx &c, int nbf, int64_t npts, const double *phi,
                             const double *gx, const double *gy, const double *gz,
                             const double *w, int nocca, const double *Ca, const double *fa,
                             int noccb, const double *Cb, const double *fb, const XcList &xl,
                             bool gga, double *E, double *Va, double *Vb);

// Batched build (vxc.cpp)
void ensure_plan(Ctx &c);
void vxc_build(Ctx &c, const double *P, const XcList &xl, bool gga, int precision,
               double *E, double *N, double *V);
void vxc_build_spin(Ctx &c, int nspin, const double *const *P, const XcList &xl, bool gga,
                    int precision, double *E, double *N, double *const *V);
void build_on_plan(Ctx &c, Plan &pl, int nspin, const double *const *P, const XcList &xl,
                   bool gga, int precision, double *E, double *N, double *const *V);
// nspin = 1: P[0] -> V[0]; nspin = 2: (P_alpha, P_beta) -> (V_alpha, V_beta)
void graph_capture(Ctx &c, int nspin, const double *const *P, const XcList &xl, bool gga,
                   int precision, double *const *V);
void graph_launch(Ctx &c, double *E, double *N);
void molset_graph_capture(Ctx &c, const double *const *P, const XcList &xl, bool gga,
                          int precision, double *const *V);
double autotune(Ctx &c, const double *P, const XcList &xl, bool gga, int precision, double *V,
                int *best_split, int *best_group);
double molset_autotune(Ctx &c, const double *const *P, const XcList &xl, bool gga,
                       int precision, double *const *V, int *best_split, int *best_group);
void build_molset_plan(Ctx &c, const std::vector<std::vector<double>> &mol_pts,
                       const std::vector<std::vector<double>> &mol_w,
                       const std::vector<BasisHost> &bases);
void overlap_build(Ctx &c, double *S);
// nuclear gradient of E_xc at fixed P (nspin spin density matrices, or the
// total one); mixed: int8-slice contractions
void xc_grad(Ctx &c, int nspin, const double *const *P, const XcList &xl, bool gga,
             int partition, bool mixed, double *E, double *grad);
// rho, grad rho, sigma of P per grid point (grid order); all but rho may be NULL
void density_points(Ctx &c, const double *P, double *rho, double *gx, double *gy, double *gz,
                    double *sigma);

}  // namespace nkx
