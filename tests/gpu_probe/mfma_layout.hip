// Test-only probe of the MFMA operand / result lane layouts the kernels of
// libnukexc_hip.so rely on.  One wavefront, 64 lanes, all accesses within
// the 64/256-element arrays it is given.
#include <hip/hip_runtime.h>

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// a[lane], b[lane] as the kernels load them; d[lane * 4 + r] = result register r.
__global__ void probe_f64(const double *a, const double *b, double *d) {
  const int l = threadIdx.x;
  v4d acc = {0.0, 0.0, 0.0, 0.0};
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[l], b[l], acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) d[l * 4 + r] = acc[r];
}

// a, b: 8 bf16 per lane (lane-major); d as above.
__global__ void probe_bf16(const __bf16 *a, const __bf16 *b, float *d) {
  const int l = threadIdx.x;
  v8bf av, bv;
  for (int i = 0; i < 8; ++i) {
    av[i] = a[l * 8 + i];
    bv[i] = b[l * 8 + i];
  }
  v4f acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) d[l * 4 + r] = acc[r];
}

extern "C" int mfma_probe_f64(const double *a_h, const double *b_h, double *d_h) {
  double *a, *b, *d;
  if (hipMalloc(&a, 64 * 8) || hipMalloc(&b, 64 * 8) || hipMalloc(&d, 256 * 8)) return 1;
  (void)hipMemcpy(a, a_h, 64 * 8, hipMemcpyHostToDevice);
  (void)hipMemcpy(b, b_h, 64 * 8, hipMemcpyHostToDevice);
  hipLaunchKernelGGL(probe_f64, dim3(1), dim3(64), 0, 0, a, b, d);
  int rc = hipDeviceSynchronize() != hipSuccess;
  (void)hipMemcpy(d_h, d, 256 * 8, hipMemcpyDeviceToHost);
  (void)hipFree(a); (void)hipFree(b); (void)hipFree(d);
  return rc;
}

// a_h, b_h: 512 bf16 values each given as raw uint16 bits.
extern "C" int mfma_probe_bf16(const unsigned short *a_h, const unsigned short *b_h, float *d_h) {
  __bf16 *a, *b;
  float *d;
  if (hipMalloc(&a, 512 * 2) || hipMalloc(&b, 512 * 2) || hipMalloc(&d, 256 * 4)) return 1;
  (void)hipMemcpy(a, a_h, 512 * 2, hipMemcpyHostToDevice);
  (void)hipMemcpy(b, b_h, 512 * 2, hipMemcpyHostToDevice);
  hipLaunchKernelGGL(probe_bf16, dim3(1), dim3(64), 0, 0, a, b, d);
  int rc = hipDeviceSynchronize() != hipSuccess;
  (void)hipMemcpy(d_h, d, 256 * 4, hipMemcpyDeviceToHost);
  (void)hipFree(a); (void)hipFree(b); (void)hipFree(d);
  return rc;
}

// Accumulation rounding of v_mfma_f32_16x16x32_bf16: every lane's C is c_h[r];
// A has a single nonzero bf16, A[0][0] = a, and B a single nonzero, B[0][0] =
// b, so D[0][0] = C[0][0] + a * b and every other D element = C.
__global__ void probe_round(const float *c, __bf16 a, __bf16 b, float *d) {
  const int l = threadIdx.x;
  v8bf av, bv;
  for (int i = 0; i < 8; ++i) {
    av[i] = (__bf16)0.0f;
    bv[i] = (__bf16)0.0f;
  }
  if (l == 0) {  // A[0][0] = a (lane 0 holds row 0, k = 0..7), B[0][0] = b
    av[0] = a;
    bv[0] = b;
  }
  v4f acc = {c[l * 4 + 0], c[l * 4 + 1], c[l * 4 + 2], c[l * 4 + 3]};
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) d[l * 4 + r] = acc[r];
}

extern "C" int mfma_probe_round(const float *c_h, float a_f, float b_f, float *d_h) {
  float *c, *d;
  if (hipMalloc(&c, 256 * 4) || hipMalloc(&d, 256 * 4)) return 1;
  (void)hipMemcpy(c, c_h, 256 * 4, hipMemcpyHostToDevice);
  hipLaunchKernelGGL(probe_round, dim3(1), dim3(64), 0, 0, c, (__bf16)a_f, (__bf16)b_f, d);
  int rc = hipDeviceSynchronize() != hipSuccess;
  (void)hipMemcpy(d_h, d, 256 * 4, hipMemcpyDeviceToHost);
  (void)hipFree(c); (void)hipFree(d);
  return rc;
}
