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

// v_mfma_i32_16x16x64_i8: 16 int8 per lane for A and B (lane-major, as
// loaded by a kernel), d[lane * 4 + r] = result register r (int32).
typedef int v4i __attribute__((ext_vector_type(4)));
__global__ void probe_i8(const v4i *a, const v4i *b, int *d) {
  const int l = threadIdx.x;
  v4i acc = {0, 0, 0, 0};
  acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[l], b[l], acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) d[l * 4 + r] = acc[r];
}

extern "C" int mfma_probe_i8(const signed char *a_h, const signed char *b_h, int *d_h) {
  v4i *a, *b;
  int *d;
  if (hipMalloc(&a, 1024) || hipMalloc(&b, 1024) || hipMalloc(&d, 256 * 4)) return 1;
  (void)hipMemcpy(a, a_h, 1024, hipMemcpyHostToDevice);
  (void)hipMemcpy(b, b_h, 1024, hipMemcpyHostToDevice);
  hipLaunchKernelGGL(probe_i8, dim3(1), dim3(64), 0, 0, a, b, d);
  int rc = hipDeviceSynchronize() != hipSuccess;
  (void)hipMemcpy(d_h, d, 256 * 4, hipMemcpyDeviceToHost);
  (void)hipFree(a); (void)hipFree(b); (void)hipFree(d);
  return rc;
}
