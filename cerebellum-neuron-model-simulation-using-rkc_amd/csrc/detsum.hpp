// Deterministic block sums: every block reduces 256 consecutive values with
// a fixed stride-halving tree (s[i] += s[i + stride], stride 128 .. 1), so a
// sum is bit-reproducible run to run and independent of the launch order of
// blocks.  Partial sums are reduced again with the same tree until one value
// is left.
#pragma once
#include <hip/hip_runtime.h>

namespace cbrkc {

constexpr int RED_BLOCK = 256;

// Block-wide sum of one value per thread; blockDim.x must be 256.  The
// result is valid in thread 0.
__device__ inline double block_tree_sum(double x, double *s) {
  const int t = threadIdx.x;
  s[t] = x;
  __syncthreads();
  if (t < 128) s[t] = s[t] + s[t + 128];
  __syncthreads();
  if (t < 64) s[t] = s[t] + s[t + 64];
  __syncthreads();
  if (t < 64) {
    // wave 0 alone from here; LDS accesses of one wave complete in order
    volatile double *vs = s;
    if (t < 32) vs[t] = vs[t] + vs[t + 32];
    __builtin_amdgcn_wave_barrier();
    if (t < 16) vs[t] = vs[t] + vs[t + 16];
    __builtin_amdgcn_wave_barrier();
    if (t < 8) vs[t] = vs[t] + vs[t + 8];
    __builtin_amdgcn_wave_barrier();
    if (t < 4) vs[t] = vs[t] + vs[t + 4];
    __builtin_amdgcn_wave_barrier();
    if (t < 2) vs[t] = vs[t] + vs[t + 2];
    __builtin_amdgcn_wave_barrier();
    if (t < 1) vs[t] = vs[t] + vs[t + 1];
  }
  return s[0];
}

__global__ void __launch_bounds__(256) k_reduce256(const double *__restrict__ in, double *__restrict__ out, int n);

}  // namespace cbrkc
