// Screened-grid set-up on the device (octree.hpp:70-389 of the reference,
// without ArborX): Morton order of the grid points, the bounding box of each
// run of consecutive points, and the CSR list of the basis functions whose
// cutoff sphere meets each box.  Run once per geometry; the V_xc build does its
// own screening on 64-point batches and does not need these lists.
#include <hipcub/hipcub.hpp>

#include "nkx_internal.hpp"
#include "vxc_kernels.hpp"

namespace nkx {

namespace {

constexpr int NT = 256;

__device__ __forceinline__ uint64_t spread21(uint64_t v) {
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffull;
  v = (v | v << 16) & 0x1f0000ff0000ffull;
  v = (v | v << 8) & 0x100f00f00f00f00full;
  v = (v | v << 4) & 0x10c30c30c30c30c3ull;
  v = (v | v << 2) & 0x1249249249249249ull;
  return v;
}

// Per-workgroup minimum and maximum of the points: part[6 g + (lo xyz, hi xyz)].
__global__ void __launch_bounds__(NT) bbox_partial(int64_t n, const double *__restrict__ pts,
                                                   double *__restrict__ part) {
  __shared__ double s[6][NT];
  double v[6] = {HUGE_VAL, HUGE_VAL, HUGE_VAL, -HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT)
    for (int d = 0; d < 3; ++d) {
      const double x = pts[3 * i + d];
      v[d] = fmin(v[d], x);
      v[3 + d] = fmax(v[3 + d], x);
    }
  for (int k = 0; k < 6; ++k) s[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int h = NT / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int k = 0; k < 6; ++k) {
        const double o = s[k][threadIdx.x + h];
        s[k][threadIdx.x] = k < 3 ? fmin(s[k][threadIdx.x], o) : fmax(s[k][threadIdx.x], o);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

// Reduces nparts partial boxes (one workgroup).
__global__ void __launch_bounds__(NT) bbox_final(int nparts, const double *__restrict__ part,
                                                 double *__restrict__ bb) {
  __shared__ double s[6][NT];
  double v[6] = {HUGE_VAL, HUGE_VAL, HUGE_VAL, -HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  for (int i = threadIdx.x; i < nparts; i += NT)
    for (int k = 0; k < 6; ++k)
      v[k] = k < 3 ? fmin(v[k], part[6 * i + k]) : fmax(v[k], part[6 * i + k]);
  for (int k = 0; k < 6; ++k) s[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int h = NT / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int k = 0; k < 6; ++k) {
        const double o = s[k][threadIdx.x + h];
        s[k][threadIdx.x] = k < 3 ? fmin(s[k][threadIdx.x], o) : fmax(s[k][threadIdx.x], o);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) bb[threadIdx.x] = s[threadIdx.x][0];
}

// 63-bit Morton code of each point: 21 bits per axis of its position in the
// bounding box, x in the highest bit of each triple.
__global__ void __launch_bounds__(NT) morton_keys(int64_t n, const double *__restrict__ pts,
                                                  const double *__restrict__ bb,
                                                  uint64_t *__restrict__ key, int *__restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (i >= n) return;
  uint64_t code = 0;
  for (int d = 0; d < 3; ++d) {
    const double ext = fmax(bb[3 + d] - bb[d], 1e-12);
    const double t = fmin(fmax((pts[3 * i + d] - bb[d]) / ext, 0.0), 1.0);
    code |= spread21((uint64_t)(t * 2097151.0)) << (2 - d);
  }
  key[i] = code;
  idx[i] = (int)i;
}

__global__ void __launch_bounds__(NT) permute_points(
    int64_t n, const int *__restrict__ perm, const double *__restrict__ pts,
    const double *__restrict__ w, const int *__restrict__ own, double *__restrict__ pts2,
    double *__restrict__ w2, int *__restrict__ own2) {
  const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (i >= n) return;
  const int64_t src = perm[i];
  for (int d = 0; d < 3; ++d) pts2[3 * i + d] = pts[3 * src + d];
  if (w) w2[i] = w[src];
  if (own) own2[i] = own[src];
}

// Box b: minimum and maximum over points b per .. min(n, (b + 1) per) - 1.
__global__ void __launch_bounds__(NT) tile_boxes(int64_t n, int per, int64_t nbox,
                                                 const double *__restrict__ pts,
                                                 double *__restrict__ boxes) {
  const int64_t b = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (b >= nbox) return;
  const int64_t g0 = b * per, g1 = g0 + per < n ? g0 + per : n;
  double lo[3], hi[3];
  for (int d = 0; d < 3; ++d) lo[d] = hi[d] = pts[3 * g0 + d];
  for (int64_t g = g0 + 1; g < g1; ++g)
    for (int d = 0; d < 3; ++d) {
      lo[d] = fmin(lo[d], pts[3 * g + d]);
      hi[d] = fmax(hi[d], pts[3 * g + d]);
    }
  for (int d = 0; d < 3; ++d) {
    boxes[6 * b + d] = lo[d];
    boxes[6 * b + 3 + d] = hi[d];
  }
}

// Sphere (c, r) meets box: distance from c to the box <= r (rounded
// products and sums, no contraction, so the test is reproducible on the host).
__device__ __forceinline__ bool meets(const double *box, const double *c, double r) {
#pragma clang fp contract(off)
  double d2 = 0.0;
  for (int d = 0; d < 3; ++d) {
    double e = 0.0;
    if (c[d] < box[d]) e = box[d] - c[d];
    else if (c[d] > box[3 + d]) e = c[d] - box[3 + d];
    d2 = d2 + e * e;
  }
  return sqrt(d2) <= r;
}

// One workgroup per box: count[b] = number of functions meeting box b.
__global__ void __launch_bounds__(NT) nl_count(const double *__restrict__ boxes, int nbf,
                                               const double *__restrict__ ctr,
                                               const double *__restrict__ rad,
                                               int *__restrict__ count) {
  __shared__ int s[NT];
  const int64_t b = blockIdx.x;
  double box[6];
  for (int k = 0; k < 6; ++k) box[k] = boxes[6 * b + k];
  int c = 0;
  for (int j = threadIdx.x; j < nbf; j += NT) c += meets(box, ctr + 3 * j, rad[j]) ? 1 : 0;
  s[threadIdx.x] = c;
  __syncthreads();
  for (int h = NT / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) count[b] = s[0];
}

// One workgroup per box: the meeting functions in ascending order at
// nbr[off[b] ..]; 256 candidates per round, compacted by wave ballots.
__global__ void __launch_bounds__(NT) nl_fill(const double *__restrict__ boxes, int nbf,
                                              const double *__restrict__ ctr,
                                              const double *__restrict__ rad,
                                              const int *__restrict__ off,
                                              int *__restrict__ nbr) {
  __shared__ int wtot[NT / 64];
  const int64_t b = blockIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double box[6];
  for (int k = 0; k < 6; ++k) box[k] = boxes[6 * b + k];
  int base = off[b];
  for (int j0 = 0; j0 < nbf; j0 += NT) {
    const int j = j0 + (int)threadIdx.x;
    const bool hit = j < nbf && meets(box, ctr + 3 * j, rad[j]);
    const uint64_t m = __ballot(hit);
    if (lane == 0) wtot[wv] = __popcll(m);
    __syncthreads();
    int before = 0, round = 0;
    for (int k = 0; k < NT / 64; ++k) {
      if (k < wv) before += wtot[k];
      round += wtot[k];
    }
    if (hit) nbr[base + before + __popcll(m & ((1ull << lane) - 1ull))] = j;
    base += round;
    __syncthreads();
  }
}

// P = C diag(f) C^T, C column-major (nbf x nocc); 16 x 16 threads per tile.
__global__ void __launch_bounds__(NT) density_matrix(int nbf, int nocc,
                                                     const double *__restrict__ C,
                                                     const double *__restrict__ f,
                                                     double *__restrict__ P) {
  const int i = blockIdx.x * 16 + (threadIdx.x & 15), j = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (i >= nbf || j >= nbf) return;
  double acc = 0.0;
  for (int k = 0; k < nocc; ++k)
    acc = fma(C[(int64_t)k * nbf + i] * f[k], C[(int64_t)k * nbf + j], acc);
  P[(int64_t)i * nbf + j] = acc;
}

// Writes the per-point density of the sorted points (rho, grad rho per padded
// point) to the caller's order; sigma = |grad rho|^2.  Any output but rho may
// be NULL.
__global__ void __launch_bounds__(NT) scatter_density(int64_t npad, const int *__restrict__ pidx,
                                                      const double *__restrict__ r4,
                                                      double *__restrict__ rho,
                                                      double *__restrict__ gx,
                                                      double *__restrict__ gy,
                                                      double *__restrict__ gz,
                                                      double *__restrict__ sigma) {
  const int64_t sp = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (sp >= npad) return;
  const int64_t g = pidx[sp];
  if (g < 0) return;
  const double r = r4[4 * sp], x = r4[4 * sp + 1], y = r4[4 * sp + 2], z = r4[4 * sp + 3];
  rho[g] = r;
  if (gx) gx[g] = x;
  if (gy) gy[g] = y;
  if (gz) gz[g] = z;
  if (sigma) sigma[g] = x * x + y * y + z * z;
}

unsigned nblocks(int64_t n) { return (unsigned)((n + NT - 1) / NT); }

// Device scratch of one set-up call, freed on every exit path.
struct Scratch {
  void *p = nullptr;
  ~Scratch() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

void sparse_bounding_boxes(hipStream_t s, int64_t npts, double *pts, double *w, int *owner,
                           int per, double *boxes) {
  if (npts <= 0) return;
  NKX_REQUIRE(npts < INT32_MAX && per > 0, NKX_ERR_INVALID_ARGUMENT,
              "bounding boxes: 0 < npts < 2^31 - 1 and max_points_per_box > 0");
  const int nparts = (int)std::min<int64_t>(1024, (npts + NT - 1) / NT);
  // scratch: partial boxes, the box, keys (2), indices (2), permuted copies
  const size_t b_part = (size_t)6 * nparts * sizeof(double), b_bb = 6 * sizeof(double);
  const size_t b_key = (size_t)npts * sizeof(uint64_t), b_idx = (size_t)npts * sizeof(int);
  const size_t b_pts = (size_t)3 * npts * sizeof(double), b_w = (size_t)npts * sizeof(double);
  size_t b_tmp = 0;
  NKX_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, b_tmp, (uint64_t *)nullptr,
                                                   (uint64_t *)nullptr, (int *)nullptr,
                                                   (int *)nullptr, (int)npts, 0, 63, s));
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t total = up(b_part) + up(b_bb) + 2 * up(b_key) + 2 * up(b_idx) + up(b_pts) +
                       up(b_w) + up(b_idx) + up(b_tmp);
  Scratch sc;
  NKX_HIP_CHECK(hipMalloc(&sc.p, total));
  char *q = static_cast<char *>(sc.p);
  auto take = [&](size_t x) {
    char *r = q;
    q += up(x);
    return r;
  };
  double *part = (double *)take(b_part), *bb = (double *)take(b_bb);
  uint64_t *k0 = (uint64_t *)take(b_key), *k1 = (uint64_t *)take(b_key);
  int *i0 = (int *)take(b_idx), *i1 = (int *)take(b_idx);
  double *pts2 = (double *)take(b_pts), *w2 = (double *)take(b_w);
  int *own2 = (int *)take(b_idx);
  void *tmp = take(b_tmp);
  hipLaunchKernelGGL(bbox_partial, dim3(nparts), dim3(NT), 0, s, npts, pts, part);
  hipLaunchKernelGGL(bbox_final, dim3(1), dim3(NT), 0, s, nparts, part, bb);
  hipLaunchKernelGGL(morton_keys, dim3(nblocks(npts)), dim3(NT), 0, s, npts, pts, bb, k0, i0);
  NKX_HIP_CHECK(
      hipcub::DeviceRadixSort::SortPairs(tmp, b_tmp, k0, k1, i0, i1, (int)npts, 0, 63, s));
  hipLaunchKernelGGL(permute_points, dim3(nblocks(npts)), dim3(NT), 0, s, npts, i1, pts, w,
                     owner, pts2, w2, own2);
  NKX_HIP_CHECK(hipMemcpyAsync(pts, pts2, b_pts, hipMemcpyDeviceToDevice, s));
  if (w) NKX_HIP_CHECK(hipMemcpyAsync(w, w2, b_w, hipMemcpyDeviceToDevice, s));
  if (owner) NKX_HIP_CHECK(hipMemcpyAsync(owner, own2, b_idx, hipMemcpyDeviceToDevice, s));
  const int64_t nbox = (npts + per - 1) / per;
  hipLaunchKernelGGL(tile_boxes, dim3(nblocks(nbox)), dim3(NT), 0, s, npts, per, nbox, pts,
                     boxes);
  NKX_HIP_CHECK(hipGetLastError());
  NKX_HIP_CHECK(hipStreamSynchronize(s));
}

int64_t sparse_neighbor_list(hipStream_t s, int64_t nbox, const double *boxes, int nbf,
                             const double *ctr_h, const double *rad_h, int *offsets,
                             int *neighbors) {
  NKX_REQUIRE(nbox >= 0 && nbox < INT32_MAX && nbf >= 0, NKX_ERR_INVALID_ARGUMENT,
              "neighbor list: 0 <= nboxes < 2^31 - 1, nbf >= 0");
  std::vector<int> cnt(nbox), off(nbox + 1, 0);
  Scratch sr, sc;
  NKX_HIP_CHECK(hipMalloc(&sr.p, (size_t)4 * std::max(nbf, 1) * sizeof(double)));
  NKX_HIP_CHECK(hipMalloc(&sc.p, (size_t)std::max<int64_t>(nbox, 1) * sizeof(int)));
  double *ctr = static_cast<double *>(sr.p), *rad = ctr + 3 * (size_t)std::max(nbf, 1);
  int *cdev = static_cast<int *>(sc.p);
  if (nbf > 0) {
    NKX_HIP_CHECK(hipMemcpyAsync(ctr, ctr_h, (size_t)3 * nbf * sizeof(double),
                                 hipMemcpyHostToDevice, s));
    NKX_HIP_CHECK(
        hipMemcpyAsync(rad, rad_h, (size_t)nbf * sizeof(double), hipMemcpyHostToDevice, s));
  }
  if (nbox > 0) {
    hipLaunchKernelGGL(nl_count, dim3((unsigned)nbox), dim3(NT), 0, s, boxes, nbf, ctr, rad,
                       cdev);
    NKX_HIP_CHECK(hipGetLastError());
    NKX_HIP_CHECK(hipMemcpyAsync(cnt.data(), cdev, (size_t)nbox * sizeof(int),
                                 hipMemcpyDeviceToHost, s));
    NKX_HIP_CHECK(hipStreamSynchronize(s));
  }
  int64_t tot = 0;
  for (int64_t b = 0; b < nbox; ++b) {
    tot += cnt[b];
    NKX_REQUIRE(tot <= INT32_MAX, NKX_ERR_UNSUPPORTED, "neighbor list longer than 2^31 - 1");
    off[b + 1] = (int)tot;
  }
  NKX_HIP_CHECK(hipMemcpyAsync(offsets, off.data(), (size_t)(nbox + 1) * sizeof(int),
                               hipMemcpyHostToDevice, s));
  if (neighbors && nbox > 0 && tot > 0) {
    hipLaunchKernelGGL(nl_fill, dim3((unsigned)nbox), dim3(NT), 0, s, boxes, nbf, ctr, rad,
                       offsets, neighbors);
    NKX_HIP_CHECK(hipGetLastError());
  }
  NKX_HIP_CHECK(hipStreamSynchronize(s));
  return tot;
}

void vk_density_matrix(hipStream_t s, int nbf, int nocc, const double *C, const double *f,
                       double *P) {
  if (nbf <= 0) return;
  const unsigned t = (unsigned)((nbf + 15) / 16);
  hipLaunchKernelGGL(density_matrix, dim3(t, t), dim3(NT), 0, s, nbf, nocc, C, f, P);
  NKX_HIP_CHECK(hipGetLastError());
}

void vk_scatter_density(hipStream_t s, const Plan &pl, const double *r4, double *rho, double *gx,
                        double *gy, double *gz, double *sigma) {
  const int64_t npad = pl.npts_pad;
  hipLaunchKernelGGL(scatter_density, dim3(nblocks(npad)), dim3(NT), 0, s, npad, pl.pidx.p, r4,
                     rho, gx, gy, gz, sigma);
  NKX_HIP_CHECK(hipGetLastError());
}

}  // namespace nkx
