// Row updates of MU and HALS with a quad of lanes per row (shared by the NMF
// and NTF kernels).  Lane q of the quad holds the Q = KM/4 consecutive row
// values t = q*Q .. q*Q+Q-1 of h and of its numerator row; St is the k x k
// FP32 Gram in LDS, transposed (St[x][t] = G(t, x)).  A column dot product
// h . G(:, x) is the quad sum of the four lanes' partial dots (lane_dot).
#ifndef PLANC_AMD_ROWUPD_H_
#define PLANC_AMD_ROWUPD_H_

#include "common.h"

namespace planc {

// Sum over the four lanes of a quad with DPP quad permutes (no LDS traffic);
// every lane of the quad gets ((v0 + v1) + (v2 + v3)).
__device__ inline float quad_sum(float v) {
  const int x1 = __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, false);  // lanes 1,0,3,2
  v += __builtin_bit_cast(float, x1);
  const int x2 = __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x4E, 0xF, 0xF, false);  // lanes 2,3,0,1
  return v + __builtin_bit_cast(float, x2);
}

// kind 1 HALS (ntf/ntfhals.hpp:13-29, nmf/hals.hpp:67-75): columns in order,
// each using the already-updated columns before it, floored at 1e-17.
// kind 0 MU (nmf/mu.hpp:58-75): h .* n ./ (h G + eps) from the old row.
// kind 2: no update.  q: the lane's position in its quad.
// ROLLED: the column loop stays a loop and the lane's values h[x % Q] are
// read and written through selects (small code, for kernels that run once
// per call and are bound by instruction fetch); otherwise the loop is fully
// unrolled with static register indexes and no branch per column: the Gram
// rows of the next columns are read from LDS while this one is computed, the
// columns go in groups of four, and a group's columns at or past k (zero
// Gram rows and numerators, as every caller pads them) are computed and then
// zeroed instead of branched around (fewer instructions per column, for
// kernels over many rows and for the sequential HALS sweep).
template <int KM>
__device__ inline void load_st_row(const float (&St)[KM][KM], int x, int q, float (&dst)[KM / 4]) {
  const float *src = &St[x][q * (KM / 4)];
#pragma unroll
  for (int u = 0; u < KM / 4; ++u) dst[u] = src[u];
}

// The lane's part of h . s: four interleaved chains (u mod 4) summed as
// ((c0 + c1) + (c2 + c3)), a dependent chain of Q/4 + 2 operations instead
// of Q (the HALS sweep waits on one such dot per column; MU's columns are
// independent and keep one chain each, fewer registers).
template <int Q>
__device__ inline float lane_dot(const float (&h)[Q], const float (&s)[Q]) {
  static_assert(Q % 4 == 0, "Q is KM / 4 with KM a multiple of 16");
  float c[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int u = 0; u < Q; ++u) c[u & 3] = fmaf(h[u], s[u], c[u & 3]);
  return (c[0] + c[1]) + (c[2] + c[3]);
}

template <int KM>
__device__ inline void zero_padding(float (&h)[KM / 4], int q, int k) {
#pragma unroll
  for (int u = 0; u < KM / 4; ++u) h[u] = q * (KM / 4) + u < k ? h[u] : 0.0f;
}

template <int KM, bool ROLLED>
__device__ inline void quad_row_update(float (&h)[KM / 4], const float (&nv)[KM / 4],
                                       const float (&St)[KM][KM], int q, int k, int kind, float eps) {
  constexpr int Q = KM / 4;
  if (kind == 1) {
    if constexpr (ROLLED) {
#pragma unroll 1
      for (int x = 0; x < k; ++x) {
        const float *sx = &St[x][q * Q];
        float ps = 0.0f, hx = 0.0f, mx = 0.0f;
        const int xo = x % Q;
#pragma unroll
        for (int u = 0; u < Q; ++u) {
          ps = fmaf(h[u], sx[u], ps);
          hx = (u == xo) ? h[u] : hx;
          mx = (u == xo) ? nv[u] : mx;
        }
        ps = quad_sum(ps);
        float v = hx + (mx - ps);
        v = v < 1e-17f ? 1e-17f : v;
        const bool own = q == x / Q;
#pragma unroll
        for (int u = 0; u < Q; ++u) h[u] = (own && u == xo) ? v : h[u];
      }
    } else {
      // Gram rows of the next PF columns in flight while column x is
      // updated; the scheduling barrier keeps their loads ahead of the
      // column's arithmetic (left alone, the compiler sinks each load to
      // just before its use and every column waits out the LDS latency).
      constexpr int PF = 2;
      float sx[PF + 1][Q];
#pragma unroll
      for (int j = 0; j < PF; ++j) load_st_row<KM>(St, j, q, sx[j]);
#pragma unroll
      for (int x = 0; x < KM; ++x) {
        if ((x & 3) == 0 && x >= k) break;
        if (x + PF < KM) load_st_row<KM>(St, x + PF, q, sx[(x + PF) % (PF + 1)]);
        __builtin_amdgcn_sched_barrier(0);
        const float ps = quad_sum(lane_dot<Q>(h, sx[x % (PF + 1)]));
        float v = h[x % Q] + (nv[x % Q] - ps);
        v = v < 1e-17f ? 1e-17f : v;
        h[x % Q] = q == x / Q ? v : h[x % Q];
      }
      zero_padding<KM>(h, q, k);
    }
  } else if (kind == 0) {
    float hn[Q];
#pragma unroll
    for (int u = 0; u < Q; ++u) hn[u] = h[u];
    if constexpr (ROLLED) {
#pragma unroll 1
      for (int c = 0; c < k; ++c) {
        const float *sc = &St[c][q * Q];
        float ps = 0.0f, hc = 0.0f, mc = 0.0f;
        const int co = c % Q;
#pragma unroll
        for (int u = 0; u < Q; ++u) {
          ps = fmaf(h[u], sc[u], ps);
          hc = (u == co) ? h[u] : hc;
          mc = (u == co) ? nv[u] : mc;
        }
        ps = quad_sum(ps);
        const float v = (hc * mc) / (ps + eps);
        const bool own = q == c / Q;
#pragma unroll
        for (int u = 0; u < Q; ++u) hn[u] = (own && u == co) ? v : hn[u];
      }
    } else {
      // Every lane of the quad gets each column's sum; a lane keeps the sums
      // of its own columns and divides only those (Q divisions, not KM).
      float sc[2][Q], den[Q];
#pragma unroll
      for (int u = 0; u < Q; ++u) den[u] = 1.0f;
      load_st_row<KM>(St, 0, q, sc[0]);
#pragma unroll
      for (int c = 0; c < KM; ++c) {
        if ((c & 3) == 0 && c >= k) break;
        if (c + 1 < KM) load_st_row<KM>(St, c + 1, q, sc[(c + 1) & 1]);
        float ps = 0.0f;  // columns are independent in MU: one chain per column
#pragma unroll
        for (int u = 0; u < Q; ++u) ps = fmaf(h[u], sc[c & 1][u], ps);
        ps = quad_sum(ps);
        den[c % Q] = q == c / Q ? ps : den[c % Q];
      }
#pragma unroll
      for (int u = 0; u < Q; ++u) hn[u] = (h[u] * nv[u]) / (den[u] + eps);
      zero_padding<KM>(hn, q, k);
    }
#pragma unroll
    for (int u = 0; u < Q; ++u) h[u] = hn[u];
  }
}

}  // namespace planc

#endif  // PLANC_AMD_ROWUPD_H_
