// Host stand-in for the HIP runtime header, for the host build of the
// small-cell device body (tests/host_emu): just enough for that code to
// compile as host C++ with one host thread per lane.  Shuffles and wave
// barriers synchronise the lanes of one cell group (emu_main.cpp).
#pragma once
#include <math.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#define __host__
#define __device__
#define __global__
#define __forceinline__ inline
#define __shared__
#define __launch_bounds__(...)
struct emu_dim3 {
  unsigned x, y, z;
};
extern thread_local emu_dim3 threadIdx;
extern thread_local emu_dim3 blockIdx;
double emu_shfl_xor(double x, int off);
void emu_wave_barrier();
inline double __shfl_xor(double x, int off) { return emu_shfl_xor(x, off); }
#define __builtin_amdgcn_wave_barrier() emu_wave_barrier()
inline unsigned atomicOr(unsigned *p, unsigned v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
inline int atomicAdd(int *p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
typedef int hipError_t;
typedef void *hipStream_t;
constexpr hipError_t hipSuccess = 0;
inline const char *hipGetErrorString(hipError_t) { return "host stand-in"; }
