// HBM streaming rates of the device this library runs on: a plain copy and a
// plain read of a buffer, each with every load and store a 16 B vector, 8
// independent vectors in flight per lane, over a grid of 16 blocks per CU.
// The rates are what the memory system delivers to a kernel that does nothing
// but stream, measured the same way as the step's kernels (HIP events around
// back-to-back launches on one stream); they put the step's bandwidth beside
// what is reachable, next to the published 8 TB/s peak.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "../../include/godunov.h"
#include "gd_internal.hpp"

namespace gd {
namespace {

typedef unsigned int v4u __attribute__((ext_vector_type(4)));
constexpr int kLanes = 256;
constexpr int kUnroll = 8;

template <bool NT>
__global__ __launch_bounds__(kLanes) void k_stream_copy(const v4u* __restrict__ a, v4u* __restrict__ b, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * kLanes * kUnroll;
  for (int64_t base = (int64_t)blockIdx.x * kLanes * kUnroll + threadIdx.x; base < n; base += stride) {
    v4u x[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + (int64_t)u * kLanes;
      if (i < n) x[u] = NT ? __builtin_nontemporal_load(a + i) : a[i];
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + (int64_t)u * kLanes;
      if (i < n) {
        if (NT) __builtin_nontemporal_store(x[u], b + i);
        else b[i] = x[u];
      }
    }
  }
}

// read only: each lane folds what it reads into one word, stored per lane so
// that the loads are live
template <bool NT>
__global__ __launch_bounds__(kLanes) void k_stream_read(const v4u* __restrict__ a, unsigned* __restrict__ sink,
                                                        int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * kLanes * kUnroll;
  unsigned acc = 0;
  for (int64_t base = (int64_t)blockIdx.x * kLanes * kUnroll + threadIdx.x; base < n; base += stride) {
    v4u x[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + (int64_t)u * kLanes;
      x[u] = i < n ? (NT ? __builtin_nontemporal_load(a + i) : a[i]) : v4u{0u, 0u, 0u, 0u};
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) acc ^= x[u].x ^ x[u].y ^ x[u].z ^ x[u].w;
  }
  sink[(int64_t)blockIdx.x * kLanes + threadIdx.x] = acc;
}

int hip_fail(hipError_t e, const char* what) {
  return set_error(GD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace
}  // namespace gd

int gd_hbm_stream_rates(int64_t bytes, int32_t iters, double* copy_gbs, double* read_gbs) {
  using namespace gd;
  if (bytes < 16 * 1024 || bytes % 16 != 0 || iters < 1 || !copy_gbs || !read_gbs)
    return set_error(GD_ERR_INVALID, "gd_hbm_stream_rates: bytes must be a multiple of 16 and at least 16 KiB, iters >= 1");
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return hip_fail(e, "device query");
  const int64_t n = bytes / 16;
  const int64_t per_block = (int64_t)kLanes * kUnroll;
  const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + per_block - 1) / per_block, 16LL * cus));
  v4u *a = nullptr, *b = nullptr;
  unsigned* sink = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = GD_OK;
  auto chk = [&](hipError_t x, const char* what) {
    if (x != hipSuccess && rc == GD_OK) rc = hip_fail(x, what);
    return rc == GD_OK;
  };
  if (chk(hipMalloc(&a, bytes), "hipMalloc") && chk(hipMalloc(&b, bytes), "hipMalloc") &&
      chk(hipMalloc(&sink, (size_t)blocks * kLanes * sizeof(unsigned)), "hipMalloc") &&
      chk(hipStreamCreate(&s), "hipStreamCreate") && chk(hipEventCreate(&e0), "hipEventCreate") &&
      chk(hipEventCreate(&e1), "hipEventCreate") && chk(hipMemsetAsync(a, 1, bytes, s), "hipMemsetAsync")) {
    // best of the cached and the non-temporal forms, each timed after one warm launch
    double best_copy = 0.0, best_read = 0.0;
    for (int nt = 0; nt < 2 && rc == GD_OK; ++nt) {
      for (int kind = 0; kind < 2 && rc == GD_OK; ++kind) {
        for (int it = 0; it <= iters && rc == GD_OK; ++it) {
          if (it == 1) chk(hipEventRecord(e0, s), "hipEventRecord");
          if (kind == 0) {
            if (nt) hipLaunchKernelGGL(k_stream_copy<true>, dim3(blocks), dim3(kLanes), 0, s, a, b, n);
            else hipLaunchKernelGGL(k_stream_copy<false>, dim3(blocks), dim3(kLanes), 0, s, a, b, n);
          } else {
            if (nt) hipLaunchKernelGGL(k_stream_read<true>, dim3(blocks), dim3(kLanes), 0, s, a, sink, n);
            else hipLaunchKernelGGL(k_stream_read<false>, dim3(blocks), dim3(kLanes), 0, s, a, sink, n);
          }
          chk(hipGetLastError(), "launch");
        }
        float ms = 0.0f;
        if (chk(hipEventRecord(e1, s), "hipEventRecord") && chk(hipEventSynchronize(e1), "hipEventSynchronize") &&
            chk(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime") && ms > 0.0f) {
          const double moved = (kind == 0 ? 2.0 : 1.0) * (double)bytes * iters;
          const double gbs = moved / (ms * 1e-3) / 1e9;
          if (kind == 0) best_copy = std::max(best_copy, gbs);
          else best_read = std::max(best_read, gbs);
        }
      }
    }
    *copy_gbs = best_copy;
    *read_gbs = best_read;
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (s) hipStreamDestroy(s);
  if (sink) hipFree(sink);
  if (b) hipFree(b);
  if (a) hipFree(a);
  return rc;
}
