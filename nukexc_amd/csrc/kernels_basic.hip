// Grid partition, compaction, reference-shaped collocation / density / XC.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "nkx_internal.hpp"
#include "sto_eval.hpp"
#include "xc_functionals.hpp"
#include "xc_polarized.hpp"

namespace nkx {

// ---------------------------------------------------------------------------
// Becke partition (partitioning.hpp:44-55, :63-120, :126-213), one thread per
// point:
//   w_g *= P_own / sum_i P_i,  P_i = prod_(j != i) s(mu_ij),
//   mu_ij = clamp((r_i - r_j) / R~_ij, -1, 1),  s = (1 - p(p(p(mu)))) / 2,
// with R~ = min(R_ij, 5) in the capped form.  Every product runs over j in
// index order and the sum over i in index order, as the reference's loops.
//
// Capped form, exact screening: with R~ <= 5 every atom i whose distance
// exceeds the nearest atom's by >= 5 bohr has mu_(i,nearest) >= 1, so s = 0
// and P_i = 0 exactly; every j farther than r_i + 5 has mu_ij <= -1 and
// contributes the factor s(-1) = 1 exactly (p(+-1) = +-1 in FP64).  Skipping
// those factors and terms leaves every product and sum bit-identical (the
// unscreened mode, BECKE_NO_SCREEN, evaluates all of them and is compared
// with the screened one bit for bit by the tests).
//
// Work per point: distances are compared squared first, so a square root is
// taken only for atoms near enough to matter (the nearest distance is the
// root of the smallest square: the rounded root is monotonic).  Candidates i
// (at most BK_NC per pass, in index order) are held in registers; one sweep
// over j in index order multiplies every candidate's product by its factor,
// so each r_j is formed once per pass rather than once per (i, j).  The
// capped (or plain) inter-atomic distances come from a table (becke_rtab).
// ---------------------------------------------------------------------------
// The factor's polynomial and the distances are formed with rounded products
// and sums in the order written (no contraction into FMAs), the reference's
// compute_p / compute_s (partitioning.hpp:51-55) with x^3 = (x x) x: near
// mu = 1, 1 - p(p(p(mu))) keeps no digits below 2^-53, so which points end
// with exactly zero weight depends on this rounding, and fixing it makes the
// partition (and the compacted point set) reproducible bit for bit by a host
// loop with the same order.
__device__ __forceinline__ double becke_p(double x) {
#pragma clang fp contract(off)
  return (1.5 * x) - (0.5 * (x * x * x));
}

__device__ __forceinline__ double becke_s(double mu) {
#pragma clang fp contract(off)
  return 0.5 * (1.0 - becke_p(becke_p(becke_p(mu))));
}

// squared distance ((dx dx + dy dy) + dz dz), rounded at every step
__device__ __forceinline__ double bdist2(double px, double py, double pz, const double *c) {
#pragma clang fp contract(off)
  const double dx = px - c[0], dy = py - c[1], dz = pz - c[2];
  return dx * dx + dy * dy + dz * dz;
}

// R~_ij (capped) or R_ij (plain) for all pairs, row-major natoms x natoms
__global__ void __launch_bounds__(256) becke_rtab(int natoms, const double *__restrict__ C,
                                                  int capped, double *__restrict__ rt) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)natoms * natoms) return;
  const int i = (int)(e / natoms), j = (int)(e % natoms);
  double R = sqrt(bdist2(C[3 * i], C[3 * i + 1], C[3 * i + 2], C + 3 * j));
  if (capped) R = fmin(R, 5.0);
  rt[e] = R;
}

constexpr int BK_NC = 16;

__global__ void __launch_bounds__(256) becke_kernel(int natoms, const double *__restrict__ C,
                                                    const double *__restrict__ rt, int64_t npts,
                                                    const double *__restrict__ pts,
                                                    const int *__restrict__ owner,
                                                    double *__restrict__ w, int mode) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= npts) return;
  const bool capped = (mode & BECKE_CAPPED) != 0;
  const bool screen = capped && !(mode & BECKE_NO_SCREEN);
  const double px = pts[3 * g], py = pts[3 * g + 1], pz = pts[3 * g + 2];
  const int own = owner[g];
  double rmin = 0.0, lim_i2 = 0.0;
  if (screen) {
    double d2min = 1e300;
    for (int j = 0; j < natoms; ++j) d2min = fmin(d2min, bdist2(px, py, pz, C + 3 * j));
    rmin = sqrt(d2min);
    // beyond this square, r_i - r_min > 5 for sure (margin 1e-3 bohr)
    lim_i2 = (rmin + 5.001) * (rmin + 5.001);
  }
  double norm = 0.0, pown = 0.0;
  int next = 0;
  while (next < natoms) {
    double ri[BK_NC], Pi[BK_NC];
    int ii[BK_NC];
#pragma unroll
    for (int k = 0; k < BK_NC; ++k) {
      ri[k] = 0.0;
      Pi[k] = 1.0;
      ii[k] = -1;
    }
    int nc = 0;
    double rmax = 0.0;
    for (; next < natoms && nc < BK_NC; ++next) {
      const double d2 = bdist2(px, py, pz, C + 3 * next);
      if (screen && d2 > lim_i2) continue;
      const double r = sqrt(d2);
      if (screen && r - rmin >= 5.0) continue;  // P_i == 0 exactly
#pragma unroll
      for (int k = 0; k < BK_NC; ++k)
        if (k == nc) {
          ri[k] = r;
          ii[k] = next;
        }
      rmax = fmax(rmax, r);
      ++nc;
    }
    if (nc == 0) break;
    // beyond this square, r_i - r_j < -5 for every candidate i
    const double lim_j2 = (rmax + 5.001) * (rmax + 5.001);
    for (int j = 0; j < natoms; ++j) {
      const double d2 = bdist2(px, py, pz, C + 3 * j);
      if (screen && d2 > lim_j2) continue;  // every factor == 1 exactly
      const double rj = sqrt(d2);
#pragma unroll
      for (int k = 0; k < BK_NC; ++k) {
        const int i = ii[k];
        if (i < 0 || i == j) continue;
        // factor == 1 exactly, or the product already exactly zero
        if (screen && (ri[k] - rj <= -5.0 || Pi[k] == 0.0)) continue;
        double mu = (ri[k] - rj) / rt[(int64_t)i * natoms + j];
        mu = mu > 1.0 ? 1.0 : (mu < -1.0 ? -1.0 : mu);
        Pi[k] *= becke_s(mu);
      }
    }
#pragma unroll
    for (int k = 0; k < BK_NC; ++k) {
      if (ii[k] < 0) continue;
      if (ii[k] == own) pown = Pi[k];
      norm += Pi[k];
    }
  }
  w[g] *= pown / norm;
}

// Atom-centred product grid on the device, one workgroup per radial shell:
// shell q of atom sh_atom[q] at radius sh_r[q] (radial weight sh_w[q]) takes
// angular rule sh_rule[q] (points ang[4 (rule_off[k] + j) ..]) and writes
// its points from sh_poff[q]: centre + r u_j with weight w_r w_j.  Rounded
// products and sums, no contraction (the pragma: HIP's __dmul_rn is a plain
// product that may be fused into the following sum): bit-identical to a
// host loop.
__global__ void __launch_bounds__(256) grid_shells_kernel(
    const int *__restrict__ sh_atom, const double *__restrict__ sh_r,
    const double *__restrict__ sh_w, const int *__restrict__ sh_rule,
    const int64_t *__restrict__ sh_poff, const int *__restrict__ rule_off,
    const int *__restrict__ rule_n, const double *__restrict__ xyz,
    const double *__restrict__ ang, double *__restrict__ P, double *__restrict__ W,
    int *__restrict__ O) {
#pragma clang fp contract(off)
  const int q = blockIdx.x, a = sh_atom[q], k = sh_rule[q];
  const double r = sh_r[q], wr = sh_w[q];
  const double *u = ang + 4 * (int64_t)rule_off[k];
  for (int j = threadIdx.x; j < rule_n[k]; j += 256) {
    const int64_t g = sh_poff[q] + j;
    for (int d = 0; d < 3; ++d) P[3 * g + d] = xyz[3 * a + d] + r * u[4 * j + d];
    W[g] = wr * u[4 * j + 3];
    O[g] = a;
  }
}

void launch_grid_shells(hipStream_t s, int nshell, const int *sh_atom, const double *sh_r,
                        const double *sh_w, const int *sh_rule, const int64_t *sh_poff,
                        const int *rule_off, const int *rule_n, const double *xyz,
                        const double *ang, double *P, double *W, int *O) {
  if (nshell <= 0) return;
  hipLaunchKernelGGL(grid_shells_kernel, dim3((unsigned)nshell), dim3(256), 0, s, sh_atom, sh_r,
                     sh_w, sh_rule, sh_poff, rule_off, rule_n, xyz, ang, P, W, O);
  NKX_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Workspace of the allocation-free Becke partition (nkx_becke_noalloc): the
// pair-distance table in a file-scope device array, BECKE_SCRATCH_ATOMS^2
// doubles (2 MiB), one copy per GPU held until the process ends.  Host code
// never passes the array's name to a kernel: its device address is resolved
// once per device (becke_scratch_init, from nkx_create with that device
// current) and kept per device.
static __device__ __align__(256) double scratch[BECKE_SCRATCH_ATOMS * BECKE_SCRATCH_ATOMS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

void becke_scratch_init(int device) {
  NKX_REQUIRE(device >= 0 && device < kMaxDevices, NKX_ERR_INVALID_ARGUMENT,
              "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void *p = nullptr;
  NKX_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = static_cast<double *>(p);
}

double *becke_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

void launch_becke(hipStream_t s, int natoms, const double *centers, int64_t npts,
                  const double *pts, const int *owner, double *w, int mode, double *rtab) {
  if (npts == 0) return;
  const int64_t npair = (int64_t)natoms * natoms;
  hipLaunchKernelGGL(becke_rtab, dim3((unsigned)((npair + 255) / 256)), dim3(256), 0, s, natoms,
                     centers, mode & BECKE_CAPPED, rtab);
  NKX_HIP_CHECK(hipGetLastError());
  const int64_t blocks = (npts + 255) / 256;
  hipLaunchKernelGGL(becke_kernel, dim3((unsigned)blocks), dim3(256), 0, s, natoms, centers, rtab,
                     npts, pts, owner, w, mode);
  NKX_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Order-preserving compaction of (point, weight, owner) with w > thr.  The
// reference compacts with an atomic counter, so its surviving order is
// nondeterministic (grid.hpp:329-338); this keeps generation order.
// ---------------------------------------------------------------------------
constexpr int kCompactBlock = 1024;

__global__ void __launch_bounds__(kCompactBlock) compact_count(int64_t n, double thr,
                                                               const double *__restrict__ w,
                                                               int64_t *__restrict__ counts) {
  __shared__ int part[kCompactBlock / 64];
  const int64_t g = (int64_t)blockIdx.x * kCompactBlock + threadIdx.x;
  const bool keep = g < n && w[g] > thr;
  const unsigned long long mask = __ballot(keep);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = __popcll(mask);
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t s = 0;
    for (int i = 0; i < kCompactBlock / 64; ++i) s += part[i];
    counts[blockIdx.x] = s;
  }
}

__global__ void __launch_bounds__(kCompactBlock) compact_scatter(
    int64_t n, double thr, const double *__restrict__ pin, const double *__restrict__ win,
    const int *__restrict__ oin, const int64_t *__restrict__ offs, double *__restrict__ pout,
    double *__restrict__ wout, int *__restrict__ oout) {
  __shared__ int part[kCompactBlock / 64];
  const int64_t g = (int64_t)blockIdx.x * kCompactBlock + threadIdx.x;
  const bool keep = g < n && win[g] > thr;
  const unsigned long long mask = __ballot(keep);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) part[wv] = __popcll(mask);
  __syncthreads();
  int before = 0;
  for (int i = 0; i < wv; ++i) before += part[i];
  const unsigned long long lower = lane ? (mask & ((1ull << lane) - 1ull)) : 0ull;
  if (keep) {
    const int64_t d = offs[blockIdx.x] + before + __popcll(lower);
    pout[3 * d] = pin[3 * g];
    pout[3 * d + 1] = pin[3 * g + 1];
    pout[3 * d + 2] = pin[3 * g + 2];
    wout[d] = win[g];
    oout[d] = oin[g];
  }
}

int64_t compact_points(hipStream_t s, int64_t n, double thr, const double *pin,
                       const double *win, const int *oin, double *pout, double *wout,
                       int *oout, int * /*unused*/, int64_t *counts) {
  if (n == 0) return 0;
  const int64_t nb = (n + kCompactBlock - 1) / kCompactBlock;
  hipLaunchKernelGGL(compact_count, dim3((unsigned)nb), dim3(kCompactBlock), 0, s, n, thr, win,
                     counts);
  NKX_HIP_CHECK(hipGetLastError());
  std::vector<int64_t> h(nb);
  NKX_HIP_CHECK(hipMemcpyAsync(h.data(), counts, nb * sizeof(int64_t), hipMemcpyDeviceToHost, s));
  NKX_HIP_CHECK(hipStreamSynchronize(s));
  int64_t tot = 0;
  for (int64_t i = 0; i < nb; ++i) {
    const int64_t c = h[i];
    h[i] = tot;
    tot += c;
  }
  NKX_HIP_CHECK(hipMemcpyAsync(counts, h.data(), nb * sizeof(int64_t), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(compact_scatter, dim3((unsigned)nb), dim3(kCompactBlock), 0, s, n, thr, pin,
                     win, oin, counts, pout, wout, oout);
  NKX_HIP_CHECK(hipGetLastError());
  NKX_HIP_CHECK(hipStreamSynchronize(s));
  return tot;
}

// ---------------------------------------------------------------------------
// Collocation in the reference layout: LayoutLeft (N_bf, N_pts), element
// (mu, g) at mu + g * N_bf (stobasis.hpp:494-584).  One thread per element,
// mu fastest so the stores are contiguous.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) colloc_ref_kernel(int nbf, const FuncRec *__restrict__ F,
                                                         const HarmTerm *__restrict__ T,
                                                         int64_t npts,
                                                         const double *__restrict__ pts,
                                                         double *phi, double *gx, double *gy,
                                                         double *gz) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nbf * npts) return;
  const int mu = (int)(e % nbf);
  const int64_t g = e / nbf;
  const FuncRec f = F[mu];
  const double x = pts[3 * g], y = pts[3 * g + 1], z = pts[3 * g + 2];
  if (gx) {
    double v, a, b, c;
    sto_value_grad<double>(f, T, x, y, z, v, a, b, c);
    if (phi) phi[e] = v;
    gx[e] = a;
    gy[e] = b;
    gz[e] = c;
  } else {
    phi[e] = sto_value<double>(f, T, x, y, z);
  }
}

void launch_collocation_ref(hipStream_t s, int nbf, const FuncRec *f, const HarmTerm *t,
                            int64_t npts, const double *pts, double *phi, double *gx,
                            double *gy, double *gz) {
  const int64_t n = (int64_t)nbf * npts;
  if (n == 0) return;
  hipLaunchKernelGGL(colloc_ref_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nbf,
                     f, t, npts, pts, phi, gx, gy, gz);
  NKX_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Plain device copy with 16-byte vector loads and stores (grid-stride).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) copy16_kernel(int64_t n, const int4 *__restrict__ src,
                                                     int4 *__restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

void launch_copy16(hipStream_t s, void *dst, const void *src, size_t bytes) {
  const int64_t n = (int64_t)(bytes / 16);
  if (n == 0) return;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 256 * 32);
  hipLaunchKernelGGL(copy16_kernel, dim3((unsigned)blocks), dim3(256), 0, s, n,
                     reinterpret_cast<const int4 *>(src), reinterpret_cast<int4 *>(dst));
  NKX_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Functional evaluation on device arrays.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) functional_kernel(XcList xl, int64_t npts,
                                                         const double *__restrict__ rho,
                                                         const double *__restrict__ sigma,
                                                         double *exc, double *vrho,
                                                         double *vsigma) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= npts) return;
  const double r = rho[g];
  const XcPoint o = xc_eval(xl.n, xl.ids, xl.thr, r, sigma ? sigma[g] : 0.0);
  exc[g] = r > 0.0 ? o.e / r : 0.0;
  vrho[g] = o.vrho;
  if (vsigma) vsigma[g] = o.vsigma;
}

void launch_functional_eval(hipStream_t s, XcList xl, int64_t npts, const double *rho,
                            const double *sigma, double *exc, double *vrho, double *vsigma) {
  if (npts == 0) return;
  hipLaunchKernelGGL(functional_kernel, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, s, xl,
                     npts, rho, sigma, exc, vrho, vsigma);
  NKX_HIP_CHECK(hipGetLastError());
}

// Spin-polarised evaluation, libxc XC_POLARIZED layout (rho[2g+s],
// sigma[3g+k], vrho[2g+s], vsigma[3g+k]); exc per particle.
__global__ void __launch_bounds__(256) functional_pol_kernel(XcList xl, int64_t npts,
                                                             const double *__restrict__ rho,
                                                             const double *__restrict__ sigma,
                                                             double *exc, double *vrho,
                                                             double *vsigma) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= npts) return;
  const double r[2] = {rho[2 * g], rho[2 * g + 1]};
  const double sg[3] = {sigma ? sigma[3 * g] : 0.0, sigma ? sigma[3 * g + 1] : 0.0,
                        sigma ? sigma[3 * g + 2] : 0.0};
  double e, vr[2], vs[3];
  xc_eval_polarized(xl.n, xl.ids, xl.thr, r, sg, e, vr, vs);
  exc[g] = e;
  vrho[2 * g] = vr[0];
  vrho[2 * g + 1] = vr[1];
  if (vsigma) {
    vsigma[3 * g] = vs[0];
    vsigma[3 * g + 1] = vs[1];
    vsigma[3 * g + 2] = vs[2];
  }
}

void launch_functional_eval_pol(hipStream_t s, XcList xl, int64_t npts, const double *rho,
                                const double *sigma, double *exc, double *vrho,
                                double *vsigma) {
  if (npts == 0) return;
  hipLaunchKernelGGL(functional_pol_kernel, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0,
                     s, xl, npts, rho, sigma, exc, vrho, vsigma);
  NKX_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Reference-shaped density (density.hpp:150-255): psi_i = sum_mu C(mu,i)
// Phi(mu,g), rho = sum_i f_i psi_i^2, grad rho = 2 sum_i f_i psi_i grad psi_i.
// One thread per point; C and f staged through LDS in chunks.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) density_ref_kernel(
    int nbf, int64_t npts, const double *__restrict__ phi, const double *__restrict__ gx,
    const double *__restrict__ gy, const double *__restrict__ gz, int nocc,
    const double *__restrict__ C, const double *__restrict__ f, double *rho, double *grx,
    double *gry, double *grz, double *sigma) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = g < npts;
  double r = 0, a = 0, b = 0, c = 0;
  for (int i = 0; i < nocc; ++i) {
    double p0 = 0, p1 = 0, p2 = 0, p3 = 0;
    if (live) {
      const double *col = C + (size_t)i * nbf;  // LayoutLeft (N_bf, N_occ)
      const size_t base = (size_t)g * nbf;
      for (int mu = 0; mu < nbf; ++mu) {
        const double cm = col[mu];
        p0 += cm * phi[base + mu];
        if (gx) {
          p1 += cm * gx[base + mu];
          p2 += cm * gy[base + mu];
          p3 += cm * gz[base + mu];
        }
      }
    }
    const double fi = f[i];
    r += fi * p0 * p0;
    a += fi * p1 * p0;
    b += fi * p2 * p0;
    c += fi * p3 * p0;
  }
  if (!live) return;
  rho[g] = r;
  if (grx) {
    grx[g] = 2.0 * a;
    gry[g] = 2.0 * b;
    grz[g] = 2.0 * c;
  }
  if (sigma) sigma[g] = 4.0 * (a * a + b * b + c * c);
}

void launch_density_ref(hipStream_t s, int nbf, int64_t npts, const double *phi,
                        const double *gx, const double *gy, const double *gz, int nocc,
                        const double *C, const double *f, double *rho, double *grx,
                        double *gry, double *grz, double *sigma) {
  if (npts == 0) return;
  hipLaunchKernelGGL(density_ref_kernel, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, s,
                     nbf, npts, phi, gx, gy, gz, nocc, C, f, rho, grx, gry, grz, sigma);
  NKX_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Reference-shaped XC assembly (compute_lda xc_integrals.hpp:73-166,
// compute_gga :424-541): Z_mu,g = w_g (1/2 Phi v_rho + 2 (grad rho . grad Phi)
// v_sigma) written over Phi's storage pattern, then V = Phi Z^T + (Phi Z^T)^T
// by a split-K LDS-tiled FP64 kernel, and E = sum_g w rho eps.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) zmat_ref_kernel(int nbf, int64_t npts,
                                                       const double *__restrict__ phi,
                                                       const double *__restrict__ gx,
                                                       const double *__restrict__ gy,
                                                       const double *__restrict__ gz,
                                                       const double *__restrict__ w,
                                                       const double *__restrict__ rho,
                                                       const double *__restrict__ grx,
                                                       const double *__restrict__ gry,
                                                       const double *__restrict__ grz,
                                                       const double *__restrict__ vrho,
                                                       const double *__restrict__ vsig,
                                                       const double *__restrict__ exc,
                                                       double *Z, double *EN) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double econ = 0.0;
  if (e < (int64_t)nbf * npts) {
    const int mu = (int)(e % nbf);
    const int64_t g = e / nbf;
    double z = 0.5 * phi[e] * vrho[g];
    if (gx) z += 2.0 * vsig[g] * (grx[g] * gx[e] + gry[g] * gy[e] + grz[g] * gz[e]);
    Z[e] = w[g] * z;
    if (mu == 0) econ = w[g] * rho[g] * exc[g];
  }
  // block reduction of the energy terms
  for (int o = 32; o > 0; o >>= 1) econ += __shfl_down(econ, o);
  __shared__ double part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = econ;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(EN, part[0] + part[1] + part[2] + part[3]);
}

// V(mu, nu) += sum_g A(mu, g) B(nu, g) with A, B LayoutLeft (nbf, npts):
// 64 x 64 output tile per block, 16 x 16 threads with 4 x 4 micro-tiles,
// K split over blockIdx.z, partial tiles added with FP64 atomics.
__global__ void __launch_bounds__(256) gemm_nt_ref_kernel(int nbf, int64_t npts,
                                                          const double *__restrict__ A,
                                                          const double *__restrict__ B,
                                                          double *V, int64_t kchunk) {
  __shared__ double As[16][65];
  __shared__ double Bs[16][65];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
  const int64_t k0 = (int64_t)blockIdx.z * kchunk;
  const int64_t k1 = k0 + kchunk < npts ? k0 + kchunk : npts;
  double acc[4][4] = {};
  for (int64_t kk = k0; kk < k1; kk += 16) {
    for (int e = threadIdx.x; e < 16 * 64; e += 256) {
      const int kr = e / 64, mc = e % 64;
      const int64_t g = kk + kr;
      const bool ok = g < k1;
      As[kr][mc] = (ok && m0 + mc < nbf) ? A[g * nbf + m0 + mc] : 0.0;
      Bs[kr][mc] = (ok && n0 + mc < nbf) ? B[g * nbf + n0 + mc] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      double a[4], b[4];
      for (int i = 0; i < 4; ++i) a[i] = As[k][ty * 4 + i];
      for (int j = 0; j < 4; ++j) b[j] = Bs[k][tx * 4 + j];
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
    }
    __syncthreads();
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      const int mu = m0 + ty * 4 + i, nu = n0 + tx * 4 + j;
      if (mu < nbf && nu < nbf) atomicAdd(&V[(size_t)mu * nbf + nu], acc[i][j]);
    }
}

__global__ void symmetrize_kernel(int nbf, const double *__restrict__ Vb, double *V) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nbf * nbf) return;
  const int i = (int)(e / nbf), j = (int)(e % nbf);
  V[e] = Vb[(size_t)i * nbf + j] + Vb[(size_t)j * nbf + i];
}

// Spin-polarised Z for spin s (compute_lsda xc_integrals.hpp:218-240,
// compute_gga_lsda :653-712): Z = w (1/2 Phi vrho_s + vec_s . grad Phi) with
// vec_a = 2 vsigma_aa grad rho_a + vsigma_ab grad rho_b (and a <-> b).
// gr holds grad rho_a (x, y, z) then grad rho_b (x, y, z), npts each.
__global__ void __launch_bounds__(256) zmat_pol_kernel(int nbf, int64_t npts, int spin,
                                                       const double *__restrict__ phi,
                                                       const double *__restrict__ gx,
                                                       const double *__restrict__ gy,
                                                       const double *__restrict__ gz,
                                                       const double *__restrict__ w,
                                                       const double *__restrict__ gr,
                                                       const double *__restrict__ vrho,
                                                       const double *__restrict__ vsig,
                                                       double *Z) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nbf * npts) return;
  const int64_t g = e / nbf;
  double z = 0.5 * phi[e] * vrho[2 * g + spin];
  if (gx) {
    const double *gs = gr + (int64_t)(3 * spin) * npts, *go = gr + (int64_t)(3 - 3 * spin) * npts;
    const double vss = 2.0 * vsig[3 * g + 2 * spin], vab = vsig[3 * g + 1];
    const double vx = vss * gs[g] + vab * go[g];
    const double vy = vss * gs[npts + g] + vab * go[npts + g];
    const double vz = vss * gs[2 * npts + g] + vab * go[2 * npts + g];
    z += vx * gx[e] + vy * gy[e] + vz * gz[e];
  }
  Z[e] = w[g] * z;
}

// rho (2g + s) and sigma (3g + k) interleaved from per-spin densities and
// gradients; energy sum_g w (rho_a + rho_b) exc accumulated by a later pass.
__global__ void __launch_bounds__(256) pack_pol_kernel(int64_t npts, const double *__restrict__ ra,
                                                       const double *__restrict__ rb,
                                                       const double *__restrict__ gr,
                                                       double *rho2, double *sig3) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= npts) return;
  rho2[2 * g] = ra[g];
  rho2[2 * g + 1] = rb[g];
  if (sig3) {
    const double ax = gr[g], ay = gr[npts + g], az = gr[2 * npts + g];
    const double bx = gr[3 * npts + g], by = gr[4 * npts + g], bz = gr[5 * npts + g];
    sig3[3 * g] = ax * ax + ay * ay + az * az;
    sig3[3 * g + 1] = ax * bx + ay * by + az * bz;
    sig3[3 * g + 2] = bx * bx + by * by + bz * bz;
  }
}

__global__ void __launch_bounds__(256) energy_pol_kernel(int64_t npts, const double *__restrict__ w,
                                                         const double *__restrict__ rho2,
                                                         const double *__restrict__ exc,
                                                         double *EN) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (g < npts) e = w[g] * (rho2[2 * g] + rho2[2 * g + 1]) * exc[g];
  for (int o = 32; o > 0; o >>= 1) e += __shfl_down(e, o);
  __shared__ double part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(EN, part[0] + part[1] + part[2] + part[3]);
}

// V = Phi Z^T + (Phi Z^T)^T for one spin channel.
static void assemble_v(hipStream_t s, int nbf, int64_t npts, const double *phi, const double *Z,
                       double *Vb, double *V) {
  NKX_HIP_CHECK(hipMemsetAsync(Vb, 0, (size_t)nbf * nbf * sizeof(double), s));
  if (npts > 0) {
    const int tiles = (nbf + 63) / 64;
    int64_t nsplit = std::max<int64_t>(1, std::min<int64_t>(512 / (tiles * tiles) + 1,
                                                            (npts + 1023) / 1024));
    const int64_t kchunk = ((npts + nsplit - 1) / nsplit + 15) / 16 * 16;
    nsplit = (npts + kchunk - 1) / kchunk;
    hipLaunchKernelGGL(gemm_nt_ref_kernel, dim3(tiles, tiles, (unsigned)nsplit), dim3(256), 0, s,
                       nbf, npts, phi, Z, Vb, kchunk);
    NKX_HIP_CHECK(hipGetLastError());
  }
  const int64_t nn = (int64_t)nbf * nbf;
  hipLaunchKernelGGL(symmetrize_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, nbf,
                     Vb, V);
  NKX_HIP_CHECK(hipGetLastError());
}

void xc_from_collocation_pol(Ctx &c, int nbf, int64_t npts, const double *phi, const double *gx,
                             const double *gy, const double *gz, const double *w, int nocca,
                             const double *Ca, const double *fa, int noccb, const double *Cb,
                             const double *fb, const XcList &xl, bool gga, double *E,
                             double *Va, double *Vb) {
  hipStream_t s = c.stream;
  const bool grad = gga && gx;
  // per-point scratch: ra, rb, grad (6), rho2 (2), sig3 (3), exc, vrho (2),
  // vsig (3) = 19 npts; Z; Vbar; E
  const size_t per = 19 * (size_t)npts;
  c.ws.scratch.reserve(per + (size_t)nbf * npts + (size_t)nbf * nbf + 1);
  double *ra = c.ws.scratch.p, *rb = ra + npts, *gr = rb + npts, *rho2 = gr + 6 * npts,
         *sig3 = rho2 + 2 * npts, *exc = sig3 + 3 * npts, *vr = exc + npts, *vs = vr + 2 * npts;
  double *Z = c.ws.scratch.p + per;
  double *Vbar = Z + (size_t)nbf * npts;
  double *EN = Vbar + (size_t)nbf * nbf;
  launch_density_ref(s, nbf, npts, phi, grad ? gx : nullptr, gy, gz, nocca, Ca, fa, ra,
                     grad ? gr : nullptr, gr + npts, gr + 2 * npts, nullptr);
  launch_density_ref(s, nbf, npts, phi, grad ? gx : nullptr, gy, gz, noccb, Cb, fb, rb,
                     grad ? gr + 3 * npts : nullptr, gr + 4 * npts, gr + 5 * npts, nullptr);
  const unsigned nblk = (unsigned)((npts + 255) / 256);
  NKX_HIP_CHECK(hipMemsetAsync(EN, 0, sizeof(double), s));
  if (npts > 0) {
    hipLaunchKernelGGL(pack_pol_kernel, dim3(nblk), dim3(256), 0, s, npts, ra, rb, gr, rho2,
                       grad ? sig3 : nullptr);
    NKX_HIP_CHECK(hipGetLastError());
    launch_functional_eval_pol(s, xl, npts, rho2, grad ? sig3 : nullptr, exc, vr,
                               grad ? vs : nullptr);
    hipLaunchKernelGGL(energy_pol_kernel, dim3(nblk), dim3(256), 0, s, npts, w, rho2, exc, EN);
    NKX_HIP_CHECK(hipGetLastError());
  }
  const int64_t ne = (int64_t)nbf * npts;
  for (int spin = 0; spin < 2; ++spin) {
    if (ne > 0) {
      hipLaunchKernelGGL(zmat_pol_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s,
                         nbf, npts, spin, phi, grad ? gx : nullptr, gy, gz, w, gr, vr, vs, Z);
      NKX_HIP_CHECK(hipGetLastError());
    }
    assemble_v(s, nbf, npts, phi, Z, Vbar, spin ? Vb : Va);
  }
  NKX_HIP_CHECK(hipMemcpyAsync(E, EN, sizeof(double), hipMemcpyDeviceToHost, s));
  NKX_HIP_CHECK(hipStreamSynchronize(s));
}

void xc_from_collocation(Ctx &c, int nbf, int64_t npts, const double *phi, const double *gx,
                         const double *gy, const double *gz, const double *w, int nocc,
                         const double *C, const double *f, const XcList &xl, bool gga,
                         double *E, double *V) {
  hipStream_t s = c.stream;
  const bool grad = gga && gx;
  // per-point scratch: rho, grx, gry, grz, sigma, exc, vrho, vsigma; Z; Vbar; E
  const size_t per = 8 * (size_t)npts;
  c.ws.scratch.reserve(per + (size_t)nbf * npts + (size_t)nbf * nbf + 1);
  double *rho = c.ws.scratch.p, *grx = rho + npts, *gry = grx + npts, *grz = gry + npts,
         *sig = grz + npts, *exc = sig + npts, *vr = exc + npts, *vs = vr + npts;
  double *Z = c.ws.scratch.p + per;
  double *Vb = Z + (size_t)nbf * npts;
  double *EN = Vb + (size_t)nbf * nbf;
  launch_density_ref(s, nbf, npts, phi, grad ? gx : nullptr, gy, gz, nocc, C, f, rho,
                     grad ? grx : nullptr, gry, grz, grad ? sig : nullptr);
  launch_functional_eval(s, xl, npts, rho, grad ? sig : nullptr, exc, vr, grad ? vs : nullptr);
  NKX_HIP_CHECK(hipMemsetAsync(Vb, 0, ((size_t)nbf * nbf + 1) * sizeof(double), s));
  const int64_t ne = (int64_t)nbf * npts;
  if (ne > 0) {
    hipLaunchKernelGGL(zmat_ref_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, nbf,
                       npts, phi, grad ? gx : nullptr, gy, gz, w, rho, grx, gry, grz, vr, vs,
                       exc, Z, EN);
    NKX_HIP_CHECK(hipGetLastError());
    const int tiles = (nbf + 63) / 64;
    int64_t nsplit = std::max<int64_t>(1, std::min<int64_t>(512 / (tiles * tiles) + 1,
                                                            (npts + 1023) / 1024));
    const int64_t kchunk = ((npts + nsplit - 1) / nsplit + 15) / 16 * 16;
    nsplit = (npts + kchunk - 1) / kchunk;
    hipLaunchKernelGGL(gemm_nt_ref_kernel, dim3(tiles, tiles, (unsigned)nsplit), dim3(256), 0, s,
                       nbf, npts, phi, Z, Vb, kchunk);
    NKX_HIP_CHECK(hipGetLastError());
  }
  const int64_t nn = (int64_t)nbf * nbf;
  if (nn > 0) {
    hipLaunchKernelGGL(symmetrize_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s,
                       nbf, Vb, V);
    NKX_HIP_CHECK(hipGetLastError());
  }
  NKX_HIP_CHECK(hipMemcpyAsync(E, EN, sizeof(double), hipMemcpyDeviceToHost, s));
  NKX_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace nkx
