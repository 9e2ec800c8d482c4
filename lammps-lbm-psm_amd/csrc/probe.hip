// Streaming copy kernel used as the measured HBM ceiling next to the LBM
// kernels (bench.py) and as the byte-count reference of the FETCH_SIZE /
// WRITE_SIZE counters.  Not part of the lbmpsm C ABI.
#include <hip/hip_runtime.h>
#include <stdint.h>

__global__ __launch_bounds__(256) void lbmpsm_copy_probe_kernel(const double* __restrict__ a,
                                                                double* __restrict__ b,
                                                                int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) __builtin_nontemporal_store(a[i], b + i);  // as the LBM kernels: plain loads, streaming stores
}

extern "C" int lbmpsm_copy_probe(int64_t bytes, int reps, double* ms_per_copy) {
  const int64_t n = bytes / 8;
  double *a = nullptr, *b = nullptr;
  if (n <= 0 || reps <= 0 || !ms_per_copy) return -1;
  if (hipMalloc(&a, n * 8) != hipSuccess) return -2;
  if (hipMalloc(&b, n * 8) != hipSuccess) {
    (void)hipFree(a);
    return -2;
  }
  int rc = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipMemset(a, 0, n * 8) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = -3;
  } else {
    const dim3 grid((unsigned)((n + 255) / 256));
    for (int r = 0; r < 2; ++r) hipLaunchKernelGGL(lbmpsm_copy_probe_kernel, grid, dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(lbmpsm_copy_probe_kernel, grid, dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1, 0);
    float ms = 0.f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      rc = -3;
    *ms_per_copy = ms / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}
