// Philox4x32-10 counter-based random numbers (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11), for the seeded
// synthetic inputs and the Poisson synaptic drive.  Counter-based, so any
// (stream, index) draw is available on host or device without state.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cbrkc {

struct u32x4 {
  uint32_t v[4];
};

__host__ __device__ inline u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                               uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  u32x4 o;
  o.v[0] = c0; o.v[1] = c1; o.v[2] = c2; o.v[3] = c3;
  return o;
}

// 53-bit uniform double in [0, 1) from two 32-bit words.
__host__ __device__ inline double u01_53(uint32_t hi, uint32_t lo) {
  return ((double)(hi >> 5) * 67108864.0 + (double)(lo >> 6)) * (1.0 / 9007199254740992.0);
}

// Sequential stream over counters (index, stream, 0, 0) with a 64-bit seed
// as the key; host-side convenience for generators.
struct PhiloxStream {
  uint32_t k0, k1, stream;
  uint32_t idx = 0;
  uint32_t buf[4];
  int left = 0;
  PhiloxStream(uint64_t seed, uint32_t stream_id)
      : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), stream(stream_id) {}
  uint32_t next_u32() {
    if (left == 0) {
      u32x4 r = philox4x32_10(idx++, stream, 0u, 0u, k0, k1);
      for (int i = 0; i < 4; i++) buf[i] = r.v[i];
      left = 4;
    }
    return buf[4 - left--];
  }
  double uniform() {
    uint32_t a = next_u32();
    uint32_t b = next_u32();
    return u01_53(a, b);
  }
};

}  // namespace cbrkc
