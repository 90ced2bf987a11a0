// Single-block reductions shared by the reference and ns time loops: the fold
// of per-block dt partials into an Acc slot, and the wall-force sum.  Both run
// with any block size that is a multiple of 64 (up to 1024) and give the same
// result for the same block size on every call (fixed summation order).
#pragma once
#include <hip/hip_runtime.h>

#include "gd_internal.hpp"

namespace gd {

// Whole-wave max, min and count by xor butterflies (every lane of the wave
// active): max, min and an integer sum are exact, so every lane ends with the
// same value whatever the order.
__device__ __forceinline__ double wmax64(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wmin64(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ unsigned wsum_u32(unsigned v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// whole-wave floating-point sum in a fixed butterfly order
__device__ __forceinline__ double wsum64(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// part[3 b .. 3 b + 2] = (max |u| + a, min local dt, non-physical count) of
// block b; the slot gets their max / min / sum.  Loads are issued four
// partials at a time so one block drains eight thousand partials quickly.
__device__ __forceinline__ void fold_acc_block(const double* part, int64_t nparts, Acc* w) {
  __shared__ double s_v[16][3];
  const int nt = blockDim.x;
  double vmax = 0.0, hdt = __longlong_as_double(0x7ff0000000000000ll), bad = 0.0;
  int64_t b = threadIdx.x;
  for (; b + 3 * nt < nparts; b += 4 * nt) {
    double x[4][3];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int c = 0; c < 3; ++c) x[u][c] = part[3 * (b + (int64_t)u * nt) + c];
#pragma unroll
    for (int u = 0; u < 4; ++u) { vmax = fmax(vmax, x[u][0]); hdt = fmin(hdt, x[u][1]); bad += x[u][2]; }
  }
  for (; b < nparts; b += nt) {
    vmax = fmax(vmax, part[3 * b]);
    hdt = fmin(hdt, part[3 * b + 1]);
    bad += part[3 * b + 2];
  }
  vmax = wmax64(vmax);
  hdt = wmin64(hdt);
  bad = wsum64(bad);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { s_v[wid][0] = vmax; s_v[wid][1] = hdt; s_v[wid][2] = bad; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < nt / 64; ++q) { vmax = fmax(vmax, s_v[q][0]); hdt = fmin(hdt, s_v[q][1]); bad += s_v[q][2]; }
    w->vmax_bits = (unsigned long long)__double_as_longlong(vmax);
    w->hdt_bits = (unsigned long long)__double_as_longlong(hdt);
    w->nonphysical = (unsigned)bad;
  }
}

// out[0..1] = sum of the n wall-face forces (x, y); out[2..3] += the same
__device__ __forceinline__ void wall_sum_block(const double* wf, int32_t n, double* out) {
  __shared__ double s_f[16][2];
  const int nt = blockDim.x;
  double ax = 0.0, ay = 0.0;
  for (int b = threadIdx.x; b < n; b += nt) { ax += wf[2 * b]; ay += wf[2 * b + 1]; }
  ax = wsum64(ax);
  ay = wsum64(ay);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { s_f[wid][0] = ax; s_f[wid][1] = ay; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < nt / 64; ++q) { ax += s_f[q][0]; ay += s_f[q][1]; }
    out[0] = ax; out[1] = ay; out[2] += ax; out[3] += ay;
  }
}

}  // namespace gd
