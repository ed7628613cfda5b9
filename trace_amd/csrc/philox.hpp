// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC 2011), written from the
// published round function.  Host and device share it so that a render and
// its reference restatement draw the same numbers for the same
// (seed, path, bounce) counter.
#pragma once

#include <cstdint>

#ifdef __HIPCC__
#define PHILOX_QUAL __host__ __device__ inline
#else
#define PHILOX_QUAL inline
#endif

struct Philox4 {
  uint32_t v[4];
};

PHILOX_QUAL uint32_t philox_mulhi(uint32_t a, uint32_t b) {
  return uint32_t((uint64_t(a) * uint64_t(b)) >> 32);
}

PHILOX_QUAL Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                  uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = philox_mulhi(M0, c0), lo0 = M0 * c0;
    const uint32_t hi1 = philox_mulhi(M1, c2), lo1 = M1 * c2;
    const uint32_t n0 = hi1 ^ c1 ^ k0;
    const uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += W0;
    k1 += W1;
  }
  Philox4 o;
  o.v[0] = c0;
  o.v[1] = c1;
  o.v[2] = c2;
  o.v[3] = c3;
  return o;
}

// 24-bit uniform in [0, 1): exact in FP32 and FP64.
PHILOX_QUAL float u01(uint32_t x) { return float(x >> 8) * (1.0f / 16777216.0f); }

// The draw of a path: counter = (path index lo, path index hi, draw, 0),
// key = seed.  Draw 0 jitters the camera sample; draw k + 1 scatters at path
// vertex k.
PHILOX_QUAL Philox4 path_draw(uint64_t seed, uint64_t path, uint32_t draw) {
  return philox4x32_10(uint32_t(path), uint32_t(path >> 32), draw, 0u, uint32_t(seed),
                       uint32_t(seed >> 32));
}
