// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11), usable on host and device so a
// draw is the same wherever it is made.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace upc {

struct Philox4x32 {
  uint32_t c[4];
  uint32_t k[2];

  __host__ __device__ __forceinline__ static uint32_t mulhi(uint32_t a, uint32_t b, uint32_t* lo)
  {
    const uint64_t p = (uint64_t)a * b;
    *lo = (uint32_t)p;
    return (uint32_t)(p >> 32);
  }

  // One 10-round block for (counter, key); returns the four output words in c[].
  __host__ __device__ __forceinline__ void gen()
  {
    constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t k0 = k[0], k1 = k[1];
    for (int r = 0; r < 10; ++r) {
      uint32_t lo0, lo1;
      const uint32_t hi0 = mulhi(M0, c[0], &lo0);
      const uint32_t hi1 = mulhi(M1, c[2], &lo1);
      const uint32_t n0 = hi1 ^ c[1] ^ k0;
      const uint32_t n2 = hi0 ^ c[3] ^ k1;
      c[0] = n0;
      c[1] = lo1;
      c[2] = n2;
      c[3] = lo0;
      k0 += W0;
      k1 += W1;
    }
  }

  // Block `sub` of the stream (seed, event): counter (event lo, event hi, sub, 0).
  __host__ __device__ __forceinline__ void block(unsigned long long seed, unsigned long long event, uint32_t sub)
  {
    c[0] = (uint32_t)event;
    c[1] = (uint32_t)(event >> 32);
    c[2] = sub;
    c[3] = 0;
    k[0] = (uint32_t)seed;
    k[1] = (uint32_t)(seed >> 32);
    gen();
  }
};

// 53-bit uniform in [0, 1) from two 32-bit words.
__host__ __device__ __forceinline__ double philox_u01(uint32_t hi, uint32_t lo)
{
  return ((double)(hi >> 5) * 67108864.0 + (double)(lo >> 6)) * (1.0 / 9007199254740992.0);
}

}  // namespace upc
