// Streaming copy kernels used as the measured HBM ceiling next to the LBM
// kernels (bench.py) and as the byte-count reference of the FETCH_SIZE /
// WRITE_SIZE counters.  Not part of the lbmpsm C ABI.
//
// Forms (lbmpsm_copy_probe `form`):
//   0  one double per thread, plain load, non-temporal store (the LBM
//      kernels' access pattern);
//   1  16 bytes per thread, plain load, non-temporal store;
//   2  16 bytes per thread, non-temporal load and store;
//   3  16 bytes x 4 per thread in a grid-stride loop over a fixed grid,
//      plain load, non-temporal store;
//   4  as 3 with non-temporal loads;
//   5  as 3 with plain loads and stores.
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef unsigned int v4u __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void lbmpsm_copy_probe_kernel(const double* __restrict__ a,
                                                                double* __restrict__ b,
                                                                int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) __builtin_nontemporal_store(a[i], b + i);
}

template <bool NTL, bool NTS>
__device__ __forceinline__ void copy16(const v4u* a, v4u* b, int64_t i) {
  v4u v;
  if constexpr (NTL) v = __builtin_nontemporal_load(a + i);
  else v = a[i];
  if constexpr (NTS) __builtin_nontemporal_store(v, b + i);
  else b[i] = v;
}

template <bool NTL>
__global__ __launch_bounds__(256) void k_copy16(const v4u* __restrict__ a, v4u* __restrict__ b,
                                                int64_t n4) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n4) copy16<NTL, true>(a, b, i);
}

// Four independent 16-byte loads in flight per thread, then four stores;
// the grid strides over the array.
template <bool NTL, bool NTS>
__global__ __launch_bounds__(256) void k_copy16x4(const v4u* __restrict__ a, v4u* __restrict__ b,
                                                  int64_t n4) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  for (; i + 3 * stride < n4; i += 4 * stride) {
    v4u v0, v1, v2, v3;
    if constexpr (NTL) {
      v0 = __builtin_nontemporal_load(a + i);
      v1 = __builtin_nontemporal_load(a + i + stride);
      v2 = __builtin_nontemporal_load(a + i + 2 * stride);
      v3 = __builtin_nontemporal_load(a + i + 3 * stride);
    } else {
      v0 = a[i];
      v1 = a[i + stride];
      v2 = a[i + 2 * stride];
      v3 = a[i + 3 * stride];
    }
    if constexpr (NTS) {
      __builtin_nontemporal_store(v0, b + i);
      __builtin_nontemporal_store(v1, b + i + stride);
      __builtin_nontemporal_store(v2, b + i + 2 * stride);
      __builtin_nontemporal_store(v3, b + i + 3 * stride);
    } else {
      b[i] = v0;
      b[i + stride] = v1;
      b[i + 2 * stride] = v2;
      b[i + 3 * stride] = v3;
    }
  }
  for (; i < n4; i += stride) copy16<NTL, NTS>(a, b, i);
}

static void launch_form(int form, const void* a, void* b, int64_t bytes) {
  const int64_t n = bytes / 8, n4 = bytes / 16;
  const dim3 blk(256);
  switch (form) {
    case 0:
      hipLaunchKernelGGL(lbmpsm_copy_probe_kernel, dim3((unsigned)((n + 255) / 256)), blk, 0, 0,
                         (const double*)a, (double*)b, n);
      break;
    case 1:
    case 2: {
      const dim3 grid((unsigned)((n4 + 255) / 256));
      if (form == 1)
        hipLaunchKernelGGL(k_copy16<false>, grid, blk, 0, 0, (const v4u*)a, (v4u*)b, n4);
      else
        hipLaunchKernelGGL(k_copy16<true>, grid, blk, 0, 0, (const v4u*)a, (v4u*)b, n4);
      break;
    }
    default: {
      // 256 compute units x 16 workgroups of 256 threads
      const dim3 grid(4096);
      if (form == 3)
        hipLaunchKernelGGL((k_copy16x4<false, true>), grid, blk, 0, 0, (const v4u*)a, (v4u*)b, n4);
      else if (form == 4)
        hipLaunchKernelGGL((k_copy16x4<true, true>), grid, blk, 0, 0, (const v4u*)a, (v4u*)b, n4);
      else
        hipLaunchKernelGGL((k_copy16x4<false, false>), grid, blk, 0, 0, (const v4u*)a, (v4u*)b,
                           n4);
      break;
    }
  }
}

// bytes: size of each of the two buffers (a multiple of 16); returns the
// average time of one copy over `reps` launches after two untimed ones.
extern "C" int lbmpsm_copy_probe_form(int form, int64_t bytes, int reps, double* ms_per_copy) {
  if (form < 0 || form > 5 || bytes <= 0 || bytes % 16 || reps <= 0 || !ms_per_copy) return -1;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess) return -2;
  if (hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipFree(a);
    return -2;
  }
  int rc = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipMemset(a, 0, bytes) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = -3;
  } else {
    for (int r = 0; r < 2; ++r) launch_form(form, a, b, bytes);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r) launch_form(form, a, b, bytes);
    (void)hipEventRecord(e1, 0);
    float ms = 0.f;
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      rc = -3;
    *ms_per_copy = ms / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

extern "C" int lbmpsm_copy_probe(int64_t bytes, int reps, double* ms_per_copy) {
  return lbmpsm_copy_probe_form(0, bytes - bytes % 16, reps, ms_per_copy);
}
