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
//
// lbmpsm_pull_probe times the D3Q19 collide-stream's memory access pattern
// alone: every stored node pulls its 19 populations from the neighbours
// (blocked layout, 64-node blocks of 19 runs; periodic wrap on all axes)
// and writes them to its own block of the other buffer, with no arithmetic.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lbm_common.h"

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

// One wave per 64-node chunk of a z-row (sz a multiple of 64).
template <typename T>
__global__ __launch_bounds__(64) void k_pull_pattern(const T* __restrict__ fin,
                                                     T* __restrict__ fout, int sx, int sy, int sz) {
  using V = lbmpsm::VelSet<19>;
  const int nch = sz >> 6;
  const int wave = (int)blockIdx.x;
  const int row = wave / nch;
  if (row >= sx * sy) return;
  const int c = (wave - row * nch) * 64 + (int)threadIdx.x;
  const int i = row / sy, j = row - i * sy;
  T f[19];
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    int ii = i - V::ex[q], jj = j - V::ey[q], cc = c - V::ez[q];
    ii += ii < 0 ? sx : 0;
    ii -= ii >= sx ? sx : 0;
    jj += jj < 0 ? sy : 0;
    jj -= jj >= sy ? sy : 0;
    cc += cc < 0 ? sz : 0;
    cc -= cc >= sz ? sz : 0;
    const size_t n = ((size_t)ii * sy + jj) * sz + cc;
    f[q] = __builtin_nontemporal_load(fin + (n >> 6) * (19 * 64) + (n & 63) + q * 64);
  }
  const size_t n = ((size_t)i * sy + j) * sz + c;
#pragma unroll
  for (int q = 0; q < 19; ++q)
    __builtin_nontemporal_store(f[q], fout + (n >> 6) * (19 * 64) + (n & 63) + q * 64);
}

template <typename T>
static void launch_pull(const void* a, void* b, int sx, int sy, int sz) {
  const int64_t waves = (int64_t)sx * sy * (sz >> 6);
  hipLaunchKernelGGL(k_pull_pattern<T>, dim3((unsigned)waves), dim3(64), 0, 0, (const T*)a,
                     (T*)b, sx, sy, sz);
}

// elem_bytes 8, 4 or 2 (the FP64, FP32 and FP16 storage forms); returns the
// average time of one pass over `reps` launches after two untimed ones.
extern "C" int lbmpsm_pull_probe(int sx, int sy, int sz, int elem_bytes, int reps,
                                 double* ms_per_pass) {
  if (sx <= 0 || sy <= 0 || sz <= 0 || sz % 64 || reps <= 0 || !ms_per_pass ||
      (elem_bytes != 8 && elem_bytes != 4 && elem_bytes != 2))
    return -1;
  if ((int64_t)sx * sy * (sz >> 6) >= ((int64_t)1 << 31)) return -1;
  const size_t bytes = (size_t)sx * sy * sz * 19 * elem_bytes;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess) return -2;
  if (hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipFree(a);
    return -2;
  }
  auto launch = [&](int r) {
    const void* s = (r & 1) ? b : a;
    void* d = (r & 1) ? a : b;
    if (elem_bytes == 8) launch_pull<double>(s, d, sx, sy, sz);
    else if (elem_bytes == 4) launch_pull<float>(s, d, sx, sy, sz);
    else launch_pull<unsigned short>(s, d, sx, sy, sz);
  };
  int rc = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipMemset(a, 0, bytes) != hipSuccess || hipMemset(b, 0, bytes) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = -3;
  } else {
    for (int r = 0; r < 2; ++r) launch(r);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r) launch(r);
    (void)hipEventRecord(e1, 0);
    float ms = 0.f;
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      rc = -3;
    *ms_per_pass = ms / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}
