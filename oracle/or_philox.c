/*
 * Oracle: Philox4x32-10 counter-based generator (TEST INFRASTRUCTURE).
 *
 * Written from the published description: J. K. Salmon, M. A. Moraes,
 * R. O. Dror, D. E. Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11.  Ten rounds of the Philox-4x32 S-box with multipliers 0xD2511F53 and
 * 0xCD9E8D57 and Weyl key increments 0x9E3779B9 / 0xBB67AE85.
 *
 * Pinned by the known-answer vectors of the paper's reference
 * implementation (tests/test_oracle_rng.py).
 *
 * The reference itself draws its network noise from cuRAND XORWOW
 * (syn.cu:37-42); that closed library stream is not reproduced.  Synthetic
 * inputs here use this counter-based stream instead.
 */
#include "oracle.h"

static void mulhilo(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo)
{
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(0xD2511F53u, c0, &hi0, &lo0);
    mulhilo(0xCD9E8D57u, c2, &hi1, &lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
