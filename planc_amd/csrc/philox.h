// Philox4x64-10 counter-based generator (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11), written so that one source
// compiles for the host (g++) and for gfx950 (hipcc).  The stream layout is
// that of numpy.random.Philox(key=[seed, 0], counter=[0, stream, 0, 0]):
// 64-bit word e of a stream is word (e % 4) of the block whose counter is
// {e / 4 + 1, stream, 0, 0}.  Every generated input of this project is a
// pure function of (seed, stream, element index), so the host and the device
// produce bit-identical uniforms and any element can be produced in parallel.
#ifndef PLANC_PHILOX_H_
#define PLANC_PHILOX_H_

#include <stdint.h>

#if defined(__HIPCC__)
#define PLANC_HD __host__ __device__ __forceinline__
#else
#define PLANC_HD inline
#endif

namespace planc_rng {

PLANC_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * (unsigned __int128)b) >> 64);
#endif
}

struct u64x4 {
  uint64_t v[4];
};

PLANC_HD u64x4 philox4x64_10(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3,
                             uint64_t k0, uint64_t k1) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ULL;
  const uint64_t M1 = 0xCA5A826395121157ULL;
  const uint64_t W0 = 0x9E3779B97F4A7C15ULL;
  const uint64_t W1 = 0xBB67AE8584CAA73BULL;
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    uint64_t hi0 = mulhi64(M0, c0), lo0 = M0 * c0;
    uint64_t hi1 = mulhi64(M1, c2), lo1 = M1 * c2;
    uint64_t n0 = hi1 ^ c1 ^ k0;
    uint64_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  u64x4 out;
  out.v[0] = c0;
  out.v[1] = c1;
  out.v[2] = c2;
  out.v[3] = c3;
  return out;
}

// 64-bit word e of stream (seed, stream).
PLANC_HD uint64_t word(uint64_t seed, uint64_t stream, uint64_t e) {
  u64x4 b = philox4x64_10(e / 4 + 1, stream, 0, 0, seed, 0);
  return b.v[e % 4];
}

// Uniform double in [0,1) with 53 random bits: numpy's next_double.
PLANC_HD double u01(uint64_t w) { return (double)(w >> 11) * (1.0 / 9007199254740992.0); }

// Standard normal by Box-Muller from words 2e and 2e+1 of the stream; the
// first uniform is mapped to (0,1] so the logarithm is finite.
PLANC_HD double normal(uint64_t seed, uint64_t stream, uint64_t e) {
  double u1 = 1.0 - u01(word(seed, stream, 2 * e));
  double u2 = u01(word(seed, stream, 2 * e + 1));
  const double two_pi = 6.283185307179586476925286766559;
#if defined(__HIP_DEVICE_COMPILE__)
  return sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
#else
  return __builtin_sqrt(-2.0 * __builtin_log(u1)) * __builtin_cos(two_pi * u2);
#endif
}

}  // namespace planc_rng

#endif  // PLANC_PHILOX_H_
