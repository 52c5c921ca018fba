// Thin-rank products on MFMA: the bulk arithmetic of NMF and NTF.
//
//   NN:  C (L x k)  = A (L x K) * B (K x k)     A*H, X_(LxR)*KRP_R, mode-0 MTTKRP
//   TN:  C (k x L)  = B^T (k x K) * A (K x L)   W^T*A, KRP_L^T*X_(LxR)
//
// A is the big column-major operand (FP32, BF16 or FP16) streamed from HBM
// exactly once per product; B is the thin K x k operand, either a factor
// matrix or a Khatri-Rao product of up to six factors generated on the fly
// (never stored in FP32).  k <= 64 per launch (wider k loops over 64-column
// slices).
//
// Streaming kernels (FP32 NN, FP32 TN, 16-bit TN).  The big operand goes
// straight from HBM into MFMA operand registers, 16 bytes per lane, with no
// LDS staging: each lane's 16-byte word holds 4 (FP32) or 8 (16-bit)
// consecutive elements of one column, and the MFMA row/k slots are
// permuted so that those elements are exactly what the instruction needs.
//   FP32 TN, v_mfma_f32_16x16x4f32: lane (lc = lane%16, lg = lane/16) holds
//     A(kb + 4lg + j, l0 + lc), j = 0..3; the j-th MFMA contracts
//     k = kb + 4g + j in k-slot g, so one load feeds 4 MFMAs per 16 columns
//     of B^T (B(kb + 4lg + j, c) comes from one 16-byte load of B's column).
//   FP32 NN: lane holds A(l0 + 4lc + j, kb + 4lg + u) for load u = 0..3; the
//     (u, j) MFMA has row slot i <-> row l0 + 4i + j, k-slot g <-> kb + 4g + u.
//     Output row l0 + 16lg + 4v + j then sits in lane (lc, lg) at D_j[v].
//   16-bit TN, v_mfma_f32_16x16x32_{bf16,f16}: lane holds A(kb + 8lg .. +7,
//     l0 + lc) -- exactly the MFMA B-operand layout; the thin operand is
//     pre-packed once per product into the A-operand layout (16-bit, 16 B
//     per lane per 16 columns per 32 k) by pack_thin16_kernel.
//   16-bit NN: lane reads the words A(l0 + 8lc .. +7, kb + 8lg + u), u =
//     0..7, and transposes that 8 x 8 block in registers into eight MFMA
//     A-operands (row slot lc <-> row l0 + 8lc + j); the packed thin operand
//     is the B-operand.
// A wave owns 16*CW columns (TN), 64 rows (FP32 NN) or 128 rows (16-bit NN).
// The four waves of a block either take different column/row groups over the
// same k range or split the k range of one group in an interleaved order and
// are summed in LDS in fixed order.  When that still leaves too few waves,
// blockIdx.y splits K across blocks: each block writes its partial tile and a
// second kernel sums the partials in split order (deterministic, no atomics).
// Variants chosen per shape by the tuner: 32-k FP32 / 64-k 16-bit TN steps
// (whole 128-byte lines per column per step) and the block-shared thin
// operand (an LDS ring filled once per block).
//
// 16-bit NN with rows not in 16-byte groups (L % 8 != 0 or unaligned) uses an
// LDS-transposing kernel with a separate split-K reduction.
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <vector>

#include "common.h"

namespace planc {

namespace {

constexpr int NTHR = 256;

// ---------------------------------------------------------------------------
// Thin operand: B(g, c) for g < K, c < k.
struct ThinDev {
  const float *B;
  long ldb;
  const float *rowscale;
  int nf;
  const float *f[6];
  unsigned dim[6];
  long ld[6];
  int k;
  int vec;  // 16-byte loads valid at 4-aligned g (alignment and shape checked on the host)
};

__device__ inline float thin_at(const ThinDev &t, unsigned g, int c) {
  float v;
  if (t.nf == 0) {
    v = t.B[g + t.ldb * c];
  } else {
    v = 1.0f;
    unsigned rem = g;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      if (q < t.nf) {
        unsigned nr = rem / t.dim[q];
        unsigned d = rem - nr * t.dim[q];
        rem = nr;
        v *= t.f[q][d + t.ld[q] * c];
      }
    }
  }
  if (t.rowscale) v *= t.rowscale[g];
  return v;
}

// B(kb + u, c), u = 0..3 (zero at g >= kend or c >= k).  kb % 4 == 0.
__device__ inline floatx4 thin4(const ThinDev &t, long kb, long kend, int c) {
  floatx4 v = {0.f, 0.f, 0.f, 0.f};
  if (c >= t.k) return v;
  if (t.vec && kb + 3 < kend) {
    if (t.nf == 0) {
      v = *(const floatx4 *)(t.B + kb + t.ldb * c);
    } else {
      unsigned rest = (unsigned)kb / t.dim[0];
      unsigned d0 = (unsigned)kb - rest * t.dim[0];
      float s = 1.0f;
#pragma unroll
      for (int q = 1; q < 6; ++q) {
        if (q < t.nf) {
          unsigned nr = rest / t.dim[q];
          unsigned d = rest - nr * t.dim[q];
          rest = nr;
          s *= t.f[q][d + t.ld[q] * c];
        }
      }
      v = *(const floatx4 *)(t.f[0] + d0 + t.ld[0] * c);
      v[0] *= s;
      v[1] *= s;
      v[2] *= s;
      v[3] *= s;
    }
    if (t.rowscale) {
      floatx4 r = *(const floatx4 *)(t.rowscale + kb);
      v[0] *= r[0];
      v[1] *= r[1];
      v[2] *= r[2];
      v[3] *= r[3];
    }
  } else {
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (kb + u < kend) v[u] = thin_at(t, (unsigned)(kb + u), c);
  }
  return v;
}

template <typename TA>
__device__ inline uint16_t to16(float x);
template <>
__device__ inline uint16_t to16<bf16_t>(float x) {
  return float_to_bf16(x);
}
template <>
__device__ inline uint16_t to16<f16_t>(float x) {
  return float_to_f16(x);
}

template <typename TA>
__device__ inline float from16(uint16_t h);
template <>
__device__ inline float from16<bf16_t>(uint16_t h) {
  return bf16_to_float(h);
}
template <>
__device__ inline float from16<f16_t>(uint16_t h) {
  return f16_to_float(h);
}

template <typename TA>
__device__ inline floatx4 mfma16(const uint4 &a, const uint4 &b, floatx4 c);
template <>
__device__ inline floatx4 mfma16<bf16_t>(const uint4 &a, const uint4 &b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                 __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}
template <>
__device__ inline floatx4 mfma16<f16_t>(const uint4 &a, const uint4 &b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(halfx8, a),
                                                __builtin_bit_cast(halfx8, b), c, 0, 0, 0);
}

__device__ inline floatx4 mfma4(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
struct StreamParams {
  const void *A;
  long lda, L, K;
  ThinDev th;
  const uint4 *thin16;  // packed 16-bit thin operand (16-bit TN / NN)
  long plane;           // words of one packed plane; plane 1 (NP == 2) holds the low halves
  int np;               // packed planes: 1, or 2 for the compensated thin operand
  const float *colscale;  // scale of output row l (TN: A's column l; 16-bit NN: A's row l), or null
  float *C;
  long ldc;
  int cmajor;
  int wgrp;   // column (TN) / row (NN) groups per block; wgrp * wk == 4
  int wk;     // k phases per group
  long kchunk;  // k range of one blockIdx.y (multiple of the k step)
  int nsplit;
  float *part;
  int avec;   // 16-byte loads of A valid (lda and base aligned)
  int cvec;   // 16-byte stores of C valid
  float oscale;  // every output element is multiplied by it (1: none; 16-bit products only)
};

// Sums the k-phase accumulators of a block in LDS (phase 0 .. wk-1 order).
// With a K split the phase-0 waves then store the block's partial tile
// (summed over splits by splitk_reduce_kernel) and false is returned;
// otherwise returns true in the waves that hold a final tile.
template <int NA>
__device__ inline bool finish_tile(const StreamParams &p, floatx4 (&acc)[NA], floatx4 (*red)[NA][64]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int grp = wave % p.wgrp, kph = wave / p.wgrp;
  if (p.wk > 1) {
    if (kph > 0) {
#pragma unroll
      for (int a = 0; a < NA; ++a) red[wave - p.wgrp][a][lane] = acc[a];
    }
    __syncthreads();
    if (kph == 0) {
      for (int q = 1; q < p.wk; ++q) {
#pragma unroll
        for (int a = 0; a < NA; ++a) {
          floatx4 r = red[grp + (q - 1) * p.wgrp][a][lane];
          acc[a] += r;
        }
      }
    }
  }
  if (kph != 0) return false;
  if (p.nsplit <= 1) return true;
  // Partial tile of (split, group): NA x 64 floatx4, lane fastest.
  floatx4 *dst = (floatx4 *)p.part +
                 (((long)blockIdx.y * gridDim.x + blockIdx.x) * p.wgrp + grp) * (long)NA * 64;
#pragma unroll
  for (int a = 0; a < NA; ++a) dst[a * 64 + lane] = acc[a];
  return false;
}

// Output element (c, l) of a TN product.
__device__ inline void store_tn(const StreamParams &p, long l, int c, float v) {
  if (p.cmajor)
    p.C[c + p.ldc * l] = v;
  else
    p.C[l + p.ldc * c] = v;
}

// TN accumulator a = cw*NT + t of a group starting at column l0:
// v[u] = C(16t + 4lg + u, l0 + 16cw + lc).
template <int NT>
__device__ inline void store_acc_tn(const StreamParams &p, long l0, int a, floatx4 v) {
  const int lane = threadIdx.x & 63, lc = lane & 15, lg = lane >> 4;
  const int cw = a / NT, t = a % NT;
  const long l = l0 + 16 * cw + lc;
  if (l >= p.L) return;
  const float sc = (p.colscale ? p.colscale[l] : 1.0f) * p.oscale;
  v[0] *= sc;
  v[1] *= sc;
  v[2] *= sc;
  v[3] *= sc;
  const int c = 16 * t + 4 * lg;
  if (p.cmajor && p.cvec && c + 3 < p.th.k) {
    *(floatx4 *)(p.C + c + p.ldc * l) = v;
  } else {
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (c + u < p.th.k) store_tn(p, l, c + u, v[u]);
  }
}

// NN accumulator a = t*4 + j of a group starting at row l0:
// v[u] = C(l0 + 16lg + 4u + j, 16t + lc).
__device__ inline void store_acc_nn(const StreamParams &p, long l0, int a, floatx4 v) {
  const int lane = threadIdx.x & 63, lc = lane & 15, lg = lane >> 4;
  const int t = a / 4, j = a % 4;
  const int c = 16 * t + lc;
  if (c >= p.th.k) return;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const long l = l0 + 16 * lg + 4 * u + j;
    if (l >= p.L) continue;
    if (p.cmajor)
      p.C[c + p.ldc * l] = v[u];
    else
      p.C[l + p.ldc * c] = v[u];
  }
}

// Element (l, c) of a 16-bit NN accumulator a = 8t + j of the group at row
// l0: v[u] = C(l0 + 32lg + 8u + j, 16t + lc).
__device__ inline void store_acc_nn16(const StreamParams &p, long l0, int a, floatx4 v) {
  const int lane = threadIdx.x & 63, lc = lane & 15, lg = lane >> 4;
  const int t = a / 8, j = a % 8;
  const int c = 16 * t + lc;
  if (c >= p.th.k) return;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const long l = l0 + 32 * lg + 8 * u + j;
    if (l >= p.L) continue;
    const float o = v[u] * ((p.colscale ? p.colscale[l] : 1.0f) * p.oscale);
    if (p.cmajor)
      p.C[c + p.ldc * l] = o;
    else
      p.C[l + p.ldc * c] = o;
  }
}

template <int NT, int CW>
__device__ inline void epilogue_tn(const StreamParams &p, floatx4 (&acc)[CW * NT], long l0) {
#pragma unroll
  for (int a = 0; a < CW * NT; ++a) store_acc_tn<NT>(p, l0, a, acc[a]);
}

// NN tile of a wave at row l0: lane (lc, lg) holds rows l0 + 16lg + 4v + j
// of column 16t + lc in acc[4t + j][v].  For a full 64-row tile of an
// L x k output with 16-byte stores, the 16 columns of each t go through the
// wave's LDS slab ep (16 columns x EP_LD floats, padded against bank
// conflicts) so that every store instruction writes four whole 256-byte
// column segments; otherwise four consecutive rows per 16-byte store (or
// single elements).  LDS operations of one wave complete in order, so the
// slab needs no barrier, only the compiler fences around it.
constexpr int EP_LD = 68;
template <int NT>
__device__ inline void epilogue_nn(const StreamParams &p, floatx4 (&acc)[4 * NT], long l0, float *ep) {
  const int lane = threadIdx.x & 63, lc = lane & 15, lg = lane >> 4;
  if (ep && !p.cmajor && p.cvec && l0 + 63 < p.L) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        floatx4 o = {acc[t * 4 + 0][v], acc[t * 4 + 1][v], acc[t * 4 + 2][v], acc[t * 4 + 3][v]};
        *(floatx4 *)(ep + lc * EP_LD + 16 * lg + 4 * v) = o;
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int cl = 4 * q + lg;  // column of this lane's 16 bytes
        const floatx4 o = *(const floatx4 *)(ep + cl * EP_LD + 4 * lc);
        const int c = 16 * t + cl;
        if (c < p.th.k) *(floatx4 *)(p.C + l0 + 4 * lc + p.ldc * c) = o;
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    return;
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int c = 16 * t + lc;
    if (c >= p.th.k) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const long l = l0 + 16 * lg + 4 * v;
      floatx4 o = {acc[t * 4 + 0][v], acc[t * 4 + 1][v], acc[t * 4 + 2][v], acc[t * 4 + 3][v]};
      if (!p.cmajor && p.cvec && l + 3 < p.L) {
        *(floatx4 *)(p.C + l + p.ldc * c) = o;
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (l + j >= p.L) continue;
          if (p.cmajor)
            p.C[c + p.ldc * (l + j)] = o[j];
          else
            p.C[l + j + p.ldc * c] = o[j];
        }
      }
    }
  }
}

// Sum over the K splits of accumulator slabs: wave w of block (g, y) takes
// slab a = 4 y + w of group g and adds its splits in split order, eight
// loads in flight at a time.  KIND 0 (the only one): the TN forms with NT,
// CW; the NN forms use splitk_reduce_nn_kernel below.
template <int KIND, int NT, int CW>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(StreamParams p, int NA, long ngroups) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long g = blockIdx.x;
  const int a = 4 * blockIdx.y + wave;
  if (a >= NA) return;
  const long split_stride = ngroups * (long)NA * 64;
  const floatx4 *src = (const floatx4 *)p.part + g * (long)NA * 64 + (long)a * 64 + lane;
  floatx4 t = {0.f, 0.f, 0.f, 0.f};
  int sp = 0;
  for (; sp + 8 <= p.nsplit; sp += 8) {
    floatx4 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = src[(long)(sp + u) * split_stride];
#pragma unroll
    for (int u = 0; u < 8; ++u) t += v[u];
  }
  for (; sp < p.nsplit; ++sp) t += src[(long)sp * split_stride];
  static_assert(KIND == 0, "TN forms only");
  store_acc_tn<NT>(p, g * (16 * CW), a, t);
}

// The same sum for the NN forms, written through LDS.  Block (g, t) owns the
// G x 16 output tile of rows l0 = G g .. l0 + G - 1 and columns 16t .. 16t +
// 15, held in J accumulator slabs a = J t + j (FP32 NN: G = 64, J = 4, slab
// j holds rows 16lg + 4u + j; 16-bit NN: G = 128, J = 8, rows 32lg + 8u + j);
// wave w adds the splits of slabs j = w and w + 4 in split order (the sums of
// splitk_reduce_kernel, bit for bit), the sums go to LDS, and each thread
// then stores four consecutive rows of one column with one 16-byte store, so
// every column segment of the tile is written whole by one wave instead of
// element by element from slabs rows apart.  Tiles that end past L, a
// row-major C or unaligned stores take the per-slab stores.
template <int KIND>
__global__ __launch_bounds__(256) void splitk_reduce_nn_kernel(StreamParams p, int NA, long ngroups) {
  constexpr int G = KIND == 1 ? 64 : 128;  // rows of a group
  constexpr int J = KIND == 1 ? 4 : 8;     // slabs of a 16-column tile
  constexpr int JW = J / 4;                // slabs per wave
  constexpr int LD = G + 4;                // LDS row pitch (16-byte aligned)
  __shared__ __attribute__((aligned(16))) float tile[16 * LD];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lc = lane & 15, lg = lane >> 4;
  const long g = blockIdx.x;
  const int t = blockIdx.y;
  const long l0 = g * G;
  const long split_stride = ngroups * (long)NA * 64;
  const floatx4 *src = (const floatx4 *)p.part + g * (long)NA * 64 + (long)(J * t + wave) * 64 + lane;
  floatx4 s[JW];
#pragma unroll
  for (int h = 0; h < JW; ++h) s[h] = floatx4{0.f, 0.f, 0.f, 0.f};
  int sp = 0;
  for (; sp + 8 <= p.nsplit; sp += 8) {
    floatx4 v[JW][8];
#pragma unroll
    for (int h = 0; h < JW; ++h)
#pragma unroll
      for (int u = 0; u < 8; ++u) v[h][u] = src[(long)(sp + u) * split_stride + 4 * 64 * h];
#pragma unroll
    for (int h = 0; h < JW; ++h)
#pragma unroll
      for (int u = 0; u < 8; ++u) s[h] += v[h][u];
  }
  for (; sp < p.nsplit; ++sp)
#pragma unroll
    for (int h = 0; h < JW; ++h) s[h] += src[(long)sp * split_stride + 4 * 64 * h];
  const bool fast = !p.cmajor && p.cvec && l0 + G <= p.L;  // uniform over the block
  if (!fast) {
#pragma unroll
    for (int h = 0; h < JW; ++h) {
      if (KIND == 1)
        store_acc_nn(p, l0, J * t + wave + 4 * h, s[h]);
      else
        store_acc_nn16(p, l0, J * t + wave + 4 * h, s[h]);
    }
    return;
  }
#pragma unroll
  for (int h = 0; h < JW; ++h) {
    const int j = wave + 4 * h;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int r = KIND == 1 ? 16 * lg + 4 * u + j : 32 * lg + 8 * u + j;
      tile[lc * LD + r] = s[h][u];
    }
  }
  __syncthreads();
  constexpr int RC = G / 4;  // 16-byte row chunks of a column
#pragma unroll
  for (int i = 0; i < JW; ++i) {
    const int q = threadIdx.x + 256 * i;
    const int cc = q / RC, rc = q % RC;
    const int c = 16 * t + cc;
    if (c >= p.th.k) continue;
    const long l = l0 + 4 * rc;
    floatx4 o = *(const floatx4 *)(tile + cc * LD + 4 * rc);
    if (KIND == 2) {
#pragma unroll
      for (int e = 0; e < 4; ++e) o[e] = o[e] * ((p.colscale ? p.colscale[l + e] : 1.0f) * p.oscale);
    }
    *(floatx4 *)(p.C + l + p.ldc * c) = o;
  }
}

// ---------------------------------------------------------------------------
// Software pipeline over this wave's full k-steps i = 0 .. nsw-1: a ring of
// PD register buffers, load(i, buf) issues the loads of step i, mma(buf)
// consumes them.  The steady-state body has no branches (prefetch indexes
// are clamped to the last step, a re-load that hits in cache), so each mma
// waits only for the oldest buffer's loads.
template <int PD, class Buf, class Load, class Mma>
__device__ inline void stream_steps(long nsw, Buf (&buf)[PD], Load load, Mma mma) {
  if (nsw <= 0) return;
#pragma unroll
  for (int u = 0; u < PD; ++u) load(u < nsw ? (long)u : nsw - 1, buf[u]);
  long i0 = 0;
  for (; i0 + PD <= nsw; i0 += PD) {
#pragma unroll
    for (int u = 0; u < PD; ++u) {
      mma(buf[u]);
      __builtin_amdgcn_sched_barrier(0);
      const long nx = i0 + u + PD;
      load(nx < nsw ? nx : nsw - 1, buf[u]);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
#pragma unroll
  for (int u = 0; u < PD - 1; ++u)
    if (i0 + u < nsw) mma(buf[u]);
}

// Thin-operand access modes of the FP32 kernels.
constexpr int THIN_PLAIN = 0;  // B column (or the single KRP factor), 16-byte loads
constexpr int THIN_KRP2 = 2;   // two-factor KRP: 16-byte load of f0 times one f1 value
constexpr int THIN_SAFE = 9;   // anything else: masked scalar loads, generic KRP, rowscale

// Per-lane column base pointers of the thin operand for c = 16t + lc
// (clamped to k - 1: the extra columns only feed outputs that are not stored).
template <int NT>
__device__ inline void thin_cols(const ThinDev &th, const float *(&tb)[NT], const float *(&tf1)[NT]) {
  const int lc = threadIdx.x & 15;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    int c = 16 * t + lc;
    c = c < th.k ? c : th.k - 1;
    if (th.nf == 0) {
      tb[t] = th.B + th.ldb * c;
      tf1[t] = nullptr;
    } else {
      tb[t] = th.f[0] + th.ld[0] * c;
      tf1[t] = th.nf > 1 ? th.f[1] + th.ld[1] * c : nullptr;
    }
  }
}

// B(kb + 0..3, 16t + lc) for a full step (kb + 3 < K), as raw loads: a[t]
// and, for THIN_KRP2, the second factor's value sc[t] that multiplies them
// (applied by thin_apply when the step is consumed, so the loads stay in
// flight).
template <int NT, int TM>
__device__ inline void thin_step(const ThinDev &th, long kb, const float *const (&tb)[NT],
                                 const float *const (&tf1)[NT], floatx4 (&a)[NT], float (&sc)[NT]) {
  if (TM == THIN_PLAIN) {
#pragma unroll
    for (int t = 0; t < NT; ++t) a[t] = *(const floatx4 *)(tb[t] + kb);
  } else if (TM == THIN_KRP2) {
    const unsigned r = (unsigned)kb / th.dim[0];
    const unsigned d0 = (unsigned)kb - r * th.dim[0];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      a[t] = *(const floatx4 *)(tb[t] + d0);
      sc[t] = tf1[t][r];
    }
  } else {
    const int lc = threadIdx.x & 15;
#pragma unroll
    for (int t = 0; t < NT; ++t) a[t] = thin4(th, kb, kb + 4, 16 * t + lc);
  }
}

template <int NT, int TM>
__device__ inline void thin_apply(floatx4 (&a)[NT], const float (&sc)[NT]) {
  if (TM == THIN_KRP2) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      a[t][0] *= sc[t];
      a[t][1] *= sc[t];
      a[t][2] *= sc[t];
      a[t][3] *= sc[t];
    }
  }
}

// ---------------------------------------------------------------------------
// FP32 TN.  Step = 16 k.
template <int NT, int CW>
struct BufTN {
  floatx4 x[CW];
  floatx4 a[NT];
  float sc[NT];
};

template <int NT, int CW, int TM>
__global__ __launch_bounds__(NTHR) void tn_f32_kernel(StreamParams p) {
  constexpr int STEP = 16, PD = 3, NA = CW * NT;
  __shared__ floatx4 red[3][NA][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lc = lane & 15, lg = lane >> 4;
  const int grp = wave % p.wgrp, kph = wave / p.wgrp;
  const long l0 = ((long)blockIdx.x * p.wgrp + grp) * (16 * CW);
  const long kbeg = (long)blockIdx.y * p.kchunk;
  const long kend = min(p.K, kbeg + p.kchunk);
  const long nfull = kend > kbeg ? (kend - kbeg) / STEP : 0;
  const float *A = (const float *)p.A;
  const float *col[CW];
#pragma unroll
  for (int cw = 0; cw < CW; ++cw) {
    const long l = l0 + 16 * cw + lc;
    col[cw] = A + p.lda * (l < p.L ? l : p.L - 1);
  }
  const float *tb[NT], *tf1[NT];
  thin_cols<NT>(p.th, tb, tf1);
  floatx4 acc[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) acc[a] = floatx4{0.f, 0.f, 0.f, 0.f};
  auto mma = [&](BufTN<NT, CW> &b) {
    thin_apply<NT, TM>(b.a, b.sc);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int cw = 0; cw < CW; ++cw) acc[cw * NT + t] = mfma4(b.a[t][j], b.x[cw][j], acc[cw * NT + t]);
  };
  auto load = [&](long i, BufTN<NT, CW> &b) {
    const long kb = kbeg + (kph + i * p.wk) * STEP + 4 * lg;
#pragma unroll
    for (int cw = 0; cw < CW; ++cw) {
      if (TM != THIN_SAFE) {
        b.x[cw] = *(const floatx4 *)(col[cw] + kb);
      } else {
#pragma unroll
        for (int u = 0; u < 4; ++u) b.x[cw][u] = col[cw][kb + u];
      }
    }
    thin_step<NT, TM>(p.th, kb, tb, tf1, b.a, b.sc);
  };
  BufTN<NT, CW> buf[PD];
  const long nsw = nfull > kph ? (nfull - kph + p.wk - 1) / p.wk : 0;
  stream_steps<PD>(nsw, buf, load, mma);
  if (kend - kbeg > nfull * STEP && nfull % p.wk == kph) {
    // The partial last step, masked.
    BufTN<NT, CW> b;
    const long kb = kbeg + nfull * STEP + 4 * lg;
#pragma unroll
    for (int cw = 0; cw < CW; ++cw)
#pragma unroll
      for (int u = 0; u < 4; ++u) b.x[cw][u] = kb + u < kend ? col[cw][kb + u] : 0.0f;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      b.a[t] = thin4(p.th, kb, kend, 16 * t + lc);
      b.sc[t] = 1.0f;
    }
    mma(b);
  }
  if (finish_tile<NA>(p, acc, red)) epilogue_tn<NT, CW>(p, acc, l0);
}

// FP32 TN with 32-k steps: lane (lc, lg) reads A(kb + 8lg .. + 7, l0 + 16cw +
// lc) as two 16-byte words, so the four lane groups cover a whole 128-byte
// line of each column per step (the 16-k kernel reads half lines, two steps
// apart).  Half h feeds the MFMAs with k-slot g <-> kb + 8g + 4h + j and the
// thin words B(kb + 8lg + 4h + 0..3, c).  Fast thin modes only.
template <int NT, int CW>
struct BufTN2 {
  floatx4 x[CW][2];
  floatx4 a[NT][2];
  float sc[NT][2];
};

template <int NT, int CW, int TM>
__global__ __launch_bounds__(NTHR) void tn_f32d_kernel(StreamParams p) {
  constexpr int STEP = 32, PD = 2, NA = CW * NT;
  __shared__ floatx4 red[3][NA][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lc = lane & 15, lg = lane >> 4;
  const int grp = wave % p.wgrp, kph = wave / p.wgrp;
  const long l0 = ((long)blockIdx.x * p.wgrp + grp) * (16 * CW);
  const long kbeg = (long)blockIdx.y * p.kchunk;
  const long kend = min(p.K, kbeg + p.kchunk);
  const long nfull = kend > kbeg ? (kend - kbeg) / STEP : 0;
  const float *A = (const float *)p.A;
  const float *col[CW];
#pragma unroll
  for (int cw = 0; cw < CW; ++cw) {
    const long l = l0 + 16 * cw + lc;
    col[cw] = A + p.lda * (l < p.L ? l : p.L - 1);
  }
  const float *tb[NT], *tf1[NT];
  thin_cols<NT>(p.th, tb, tf1);
  floatx4 acc[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) acc[a] = floatx4{0.f, 0.f, 0.f, 0.f};
  auto mma = [&](BufTN2<NT, CW> &b) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      floatx4 ah[NT];
      float sh[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        ah[t] = b.a[t][h];
        sh[t] = b.sc[t][h];
      }
      thin_apply<NT, TM>(ah, sh);
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
          for (int cw = 0; cw < CW; ++cw) acc[cw * NT + t] = mfma4(ah[t][j], b.x[cw][h][j], acc[cw * NT + t]);
    }
  };
  auto load = [&](long i, BufTN2<NT, CW> &b) {
    const long kb = kbeg + (kph + i * p.wk) * STEP + 8 * lg;
#pragma unroll
    for (int cw = 0; cw < CW; ++cw) {
      b.x[cw][0] = *(const floatx4 *)(col[cw] + kb);
      b.x[cw][1] = *(const floatx4 *)(col[cw] + kb + 4);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      floatx4 ah[NT];
      float sh[NT];
      thin_step<NT, TM>(p.th, kb + 4 * h, tb, tf1, ah, sh);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        b.a[t][h] = ah[t];
        b.sc[t][h] = sh[t];
      }
    }
  };
  BufTN2<NT, CW> buf[PD];
  const long nsw = nfull > kph ? (nfull - kph + p.wk - 1) / p.wk : 0;
  stream_steps<PD>(nsw, buf, load, mma);
  if (kend - kbeg > nfull * STEP && nfull % p.wk == kph) {
    // The partial last step, masked.
    BufTN2<NT, CW> b;
    const long kb = kbeg + nfull * STEP + 8 * lg;
#pragma unroll
    for (int cw = 0; cw < CW; ++cw)
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const long kk = kb + 4 * h + u;
          b.x[cw][h][u] = kk < kend ? col[cw][kk] : 0.0f;
        }
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        b.a[t][h] = thin4(p.th, kb + 4 * h, kend, 16 * t + lc);
        b.sc[t][h] = 1.0f;
      }
    // thin4 already applied the KRP scale: multiply by one.
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
          for (int cw = 0; cw < CW; ++cw) acc[cw * NT + t] = mfma4(b.a[t][h][j], b.x[cw][h][j], acc[cw * NT + t]);
  }
  if (finish_tile<NA>(p, acc, red)) epilogue_tn<NT, CW>(p, acc, l0);
}

// ---------------------------------------------------------------------------
// Packs the thin operand into the 16-bit MFMA A-operand layout, in 64-k
// super-steps S split into two halves h: for lane (lc, lg),
// out[((2S + h)*NT + t)*64 + lc + 16 lg] = B(64 S + 16 lg + 8 h + 0..7, 16 t + lc).
// pack_index() finds the word of any 8-aligned k block.
__device__ inline long pack_index(long kk, int t, int lc, int NT) {
  const long S = kk >> 6;
  const int rem = (int)(kk & 63);
  return ((2 * S + ((rem >> 3) & 1)) * NT + t) * 64 + lc + 16 * (rem >> 4);
}

// NP == 2: compensated operand.  Plane 0 holds hi = rn16(b), plane 1 (at
// word offset `plane`) lo = rn16(b - hi), so hi + lo carries b to about 16
// (BF16) or 22 (FP16) significant bits and the products, with both planes
// contracted against the same big-operand words, round only the big operand.
// The NP words of 8 consecutive values (lo: the first four, hi: the next).
template <typename TA, int NP>
__device__ inline void pack8(const floatx4 &lo, const floatx4 &hi, uint4 (&w)[NP]) {
  uint16_t h[8];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    h[u] = to16<TA>(lo[u]);
    h[4 + u] = to16<TA>(hi[u]);
  }
  w[0] = __builtin_bit_cast(uint4, h);
  if constexpr (NP == 2) {
    uint16_t r[8];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      r[u] = to16<TA>(lo[u] - from16<TA>(h[u]));
      r[4 + u] = to16<TA>(hi[u] - from16<TA>(h[4 + u]));
    }
    w[1] = __builtin_bit_cast(uint4, r);
  }
}

// One thread per (S, t, lane): both halves h of its column's 16 rows 64 S +
// 16 (lane >> 4) .. + 15 (64 contiguous bytes: a wave reads whole 256-byte
// column segments).
template <typename TA, int NP>
__global__ void pack_thin16_kernel(ThinDev th, long K, int NT, long nsuper, uint4 *out, long plane) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nsuper * NT * 64) return;
  const int lane = (int)(e & 63);
  const long st = e >> 6;
  const int t = (int)(st % NT);
  const long S = st / NT;
  const int c = 16 * t + (lane & 15);
  const long kb = 64 * S + 16 * (lane >> 4);
  const floatx4 v0 = thin4(th, kb, K, c), v1 = thin4(th, kb + 4, K, c);
  const floatx4 v2 = thin4(th, kb + 8, K, c), v3 = thin4(th, kb + 12, K, c);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    uint4 w[NP];
    pack8<TA, NP>(h ? v2 : v0, h ? v3 : v1, w);
    const long o = ((2 * S + h) * NT + t) * 64 + lane;
    out[o] = w[0];
    if constexpr (NP == 2) out[plane + o] = w[1];
  }
}

// On-the-fly 16-bit thin operand of the block-shared NN kernel (instead of
// the packed buffer): producer thread tid < NT * 64 owns position (t, ln) of
// a 32-k step, column c = 16 t + (ln & 15) (clamped to k - 1: columns past k
// only feed output columns that are never stored) and the 8 rows kb + 8 (ln
// >> 4) + 0..7.  load() issues the loads: B's rows, or the first KRP factor's
// rows and the second factor's entry (the 8 rows share it: dim0 % 8 == 0),
// and the row scales; store() multiplies, rounds to the NP planes and writes
// the LDS words (q NT + t) 64 + ln.
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
struct RawFly {
  floatx4 a0, a1, r0, r1;
  float sc;
};
struct ThinProducer16 {
  const float *tb, *tf1, *rs;
  unsigned dim0;
  int t, ln;
  bool on, krp2;
  __device__ ThinProducer16(const ThinDev &th, int NT) {
    const int tid = threadIdx.x;
    on = tid < NT * 64;
    t = on ? tid >> 6 : 0;
    ln = tid & 63;
    int c = 16 * t + (ln & 15);
    c = c < th.k ? c : th.k - 1;
    tb = th.nf == 0 ? th.B + th.ldb * c : th.f[0] + th.ld[0] * c;
    krp2 = th.nf == 2;
    tf1 = krp2 ? th.f[1] + th.ld[1] * c : nullptr;
    dim0 = krp2 ? th.dim[0] : 1u;
    rs = th.rowscale;
  }
  __device__ void load(long kb, RawFly &r) const {
    if (!on) return;
    kb += 8 * (ln >> 4);
    if (krp2) {
      const unsigned q = (unsigned)kb / dim0;
      const unsigned d0 = (unsigned)kb - q * dim0;
      r.a0 = *(const floatx4 *)(tb + d0);
      r.a1 = *(const floatx4 *)(tb + d0 + 4);
      r.sc = tf1[q];
    } else {
      r.a0 = *(const floatx4 *)(tb + kb);
      r.a1 = *(const floatx4 *)(tb + kb + 4);
      r.sc = 1.0f;
    }
    if (rs) {
      r.r0 = *(const floatx4 *)(rs + kb);
      r.r1 = *(const floatx4 *)(rs + kb + 4);
    }
  }
  template <typename TA, int NT, int NP>
  __device__ void store(const RawFly &r, u32x4 (*ts)[NP * NT * 64], int stage) const {
    if (!on) return;
    floatx4 lo = r.a0 * r.sc, hi = r.a1 * r.sc;
    if (rs) {
      lo *= r.r0;
      hi *= r.r1;
    }
    uint4 w[NP];
    pack8<TA, NP>(lo, hi, w);
#pragma unroll
    for (int q = 0; q < NP; ++q) ts[stage][(q * NT + t) * 64 + ln] = __builtin_bit_cast(u32x4, w[q]);
  }
};

// 16-bit TN.  DUAL = 0: step = 32 k, a lane reads 8 consecutive k of a
// column (16 bytes); DUAL = 1: step = 64 k, a lane reads 16 consecutive k
// (two 16-byte words, a full 128-byte line per column across the four lane
// groups) feeding two MFMAs.  SAFE: scalar loads of A (unaligned lda/base).
template <int NT, int CW, int NS, int NP>
struct BufH {
  uint4 x[CW][NS];
  uint4 a[NP][NT][NS];
};

// NP == 2: the compensated thin operand (two planes, both contracted with
// the same A words into one accumulator).
template <typename TA, int NT, int CW, int SAFE, int DUAL, int NP>
__global__ __launch_bounds__(NTHR) void tn_h_kernel(StreamParams p) {
  constexpr int NS = DUAL ? 2 : 1, STEP = 32 * NS, PD = DUAL ? 2 : 3, NA = CW * NT;
  constexpr int LW = 8 * NS;  // k per lane per step
  __shared__ floatx4 red[3][NA][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lc = lane & 15, lg = lane >> 4;
  const int grp = wave % p.wgrp, kph = wave / p.wgrp;
  const long l0 = ((long)blockIdx.x * p.wgrp + grp) * (16 * CW);
  const long kbeg = (long)blockIdx.y * p.kchunk;
  const long kend = min(p.K, kbeg + p.kchunk);
  const long nfull = kend > kbeg ? (kend - kbeg) / STEP : 0;
  const uint16_t *A = (const uint16_t *)p.A;
  const uint16_t *col[CW];
#pragma unroll
  for (int cw = 0; cw < CW; ++cw) {
    const long l = l0 + 16 * cw + lc;
    col[cw] = A + p.lda * (l < p.L ? l : p.L - 1);
  }
  floatx4 acc[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) acc[a] = floatx4{0.f, 0.f, 0.f, 0.f};
  auto mma = [&](BufH<NT, CW, NS, NP> &b) {
#pragma unroll
    for (int h = 0; h < NS; ++h)
#pragma unroll
      for (int q = 0; q < NP; ++q)
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
          for (int cw = 0; cw < CW; ++cw)
            acc[cw * NT + t] = mfma16<TA>(b.a[q][t][h], b.x[cw][h], acc[cw * NT + t]);
  };
  auto load_thin = [&](long kb, BufH<NT, CW, NS, NP> &b) {
#pragma unroll
    for (int q = 0; q < NP; ++q)
#pragma unroll
      for (int h = 0; h < NS; ++h)
#pragma unroll
        for (int t = 0; t < NT; ++t) b.a[q][t][h] = p.thin16[q * p.plane + pack_index(kb + 8 * h, t, lc, NT)];
  };
  auto load = [&](long i, BufH<NT, CW, NS, NP> &b) {
    const long s = kph + i * p.wk;
    const long kb = kbeg + s * STEP + LW * lg;
#pragma unroll
    for (int cw = 0; cw < CW; ++cw)
#pragma unroll
      for (int h = 0; h < NS; ++h) {
        if (!SAFE) {
          b.x[cw][h] = *(const uint4 *)(col[cw] + kb + 8 * h);
        } else {
          uint16_t v[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) v[u] = col[cw][kb + 8 * h + u];
          b.x[cw][h] = __builtin_bit_cast(uint4, v);
        }
      }
    load_thin(kb, b);
  };
  BufH<NT, CW, NS, NP> buf[PD];
  const long nsw = nfull > kph ? (nfull - kph + p.wk - 1) / p.wk : 0;
  stream_steps<PD>(nsw, buf, load, mma);
  if (kend - kbeg > nfull * STEP && nfull % p.wk == kph) {
    BufH<NT, CW, NS, NP> b;
    const long kb = kbeg + nfull * STEP + LW * lg;
#pragma unroll
    for (int cw = 0; cw < CW; ++cw)
#pragma unroll
      for (int h = 0; h < NS; ++h) {
        uint16_t v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const long kk = kb + 8 * h + u;
          v[u] = kk < kend ? col[cw][kk] : (uint16_t)0;
        }
        b.x[cw][h] = __builtin_bit_cast(uint4, v);
      }
    load_thin(kb, b);
    mma(b);
  }
  if (finish_tile<NA>(p, acc, red)) epilogue_tn<NT, CW>(p, acc, l0);
}

// ---------------------------------------------------------------------------
// FP32 NN.  Step = 16 k; a wave owns 64 rows.  acc[t*4 + j][v] holds
// C(l0 + 16lg + 4v + j, 16t + lc).  Fast modes need L % 4 == 0 (a lane's
// four rows are in or out together; rows past L re-read rows L-4..L-1).
template <int NT>
struct BufNN {
  floatx4 x[4];
  floatx4 a[NT];
  float sc[NT];
};

template <int NT, int TM>
__global__ __launch_bounds__(NTHR) void nn_f32_kernel(StreamParams p) {
  // Three steps in flight: at K = 256 (sixteen steps per wave) a third
  // step measured 5 % faster than two, a fourth no faster.
  constexpr int STEP = 16, PD = 3, NA = 4 * NT;
  __shared__ floatx4 red[3][NA][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lc = lane & 15, lg = lane >> 4;
  const int grp = wave % p.wgrp, kph = wave / p.wgrp;
  const long l0 = ((long)blockIdx.x * p.wgrp + grp) * 64;
  const long kbeg = (long)blockIdx.y * p.kchunk;
  const long kend = min(p.K, kbeg + p.kchunk);
  const long nfull = kend > kbeg ? (kend - kbeg) / STEP : 0;
  const float *A = (const float *)p.A;
  const long lr = l0 + 4 * lc;  // first of this lane's 4 rows
  const float *base = A + (TM != THIN_SAFE ? (lr < p.L - 4 ? lr : p.L - 4) : 0);
  long rows[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) rows[j] = lr + j < p.L ? lr + j : p.L - 1;
  const float *tb[NT], *tf1[NT];
  thin_cols<NT>(p.th, tb, tf1);
  floatx4 acc[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) acc[a] = floatx4{0.f, 0.f, 0.f, 0.f};
  auto mma = [&](BufNN<NT> &b) {
    thin_apply<NT, TM>(b.a, b.sc);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t * 4 + j] = mfma4(b.x[u][j], b.a[t][u], acc[t * 4 + j]);
  };
  auto load = [&](long i, BufNN<NT> &b) {
    const long kb = kbeg + (kph + i * p.wk) * STEP + 4 * lg;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (TM != THIN_SAFE) {
        b.x[u] = *(const floatx4 *)(base + p.lda * (kb + u));
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) b.x[u][j] = A[rows[j] + p.lda * (kb + u)];
      }
    }
    thin_step<NT, TM>(p.th, kb, tb, tf1, b.a, b.sc);
  };
  BufNN<NT> buf[PD];
  const long nsw = nfull > kph ? (nfull - kph + p.wk - 1) / p.wk : 0;
  stream_steps<PD>(nsw, buf, load, mma);
  if (kend - kbeg > nfull * STEP && nfull % p.wk == kph) {
    BufNN<NT> b;
    const long kb = kbeg + nfull * STEP + 4 * lg;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int j = 0; j < 4; ++j) b.x[u][j] = kb + u < kend ? A[rows[j] + p.lda * (kb + u)] : 0.0f;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      b.a[t] = thin4(p.th, kb, kend, 16 * t + lc);
      b.sc[t] = 1.0f;
    }
    mma(b);
  }
  if (finish_tile<NA>(p, acc, red)) {
    __shared__ __attribute__((aligned(16))) float ep[4][16 * EP_LD];
    epilogue_nn<NT>(p, acc, l0, ep[wave]);
  }
}

// ---------------------------------------------------------------------------
// Block-shared thin operand.  With four waves on four row/column groups over
// the same k range, every wave needs the same thin-operand words each step;
// loading them per wave costs L1/L2 traffic close to that of A itself.  The
// shared kernels instead stage each k step of the thin operand once per
// block in an LDS ring of TS stages, in the MFMA lane layout (W 16-byte words
// per step, written by the block's first W threads), and the waves read
// their operand words from LDS right before the MFMAs.  Only A stays in the
// per-wave register ring.
//
// Iteration i (all 256 threads, one barrier per step):
//   mma(A buf i % PD, stage i % TS); load A step i + PD into that buffer;
//   store thin step i + TS - 1 (raw words loaded PD iterations earlier into
//   slot (i + TS - 1) % PD) into stage (i + TS - 1) % TS, which iteration
//   i - 1 finished reading before the last barrier; load the raw words of
//   step i + TS - 1 + PD into the freed slot; barrier.
// Prefetch indexes past the last step are clamped to it (re-loads that hit in
// cache), so the body has no branches on the step count.  The trip count is
// uniform over the block (wk == 1), as the barriers need.
template <int PD, int TS, class BufA, class Raw, class LoadA, class RawLoad, class RawStore, class Mma>
__device__ inline void stream_steps_shared(long nsw, BufA (&buf)[PD], Raw (&raw)[PD], LoadA loadA,
                                           RawLoad rload, RawStore rstore, Mma mma) {
  if (nsw <= 0) return;
  auto clamp = [&](long s) { return s < nsw ? s : nsw - 1; };
#pragma unroll
  for (int j = 0; j < TS - 1; ++j) {
    rload(clamp(j), raw[0]);
    rstore(raw[0], j);
  }
#pragma unroll
  for (int q = 0; q < PD; ++q) rload(clamp(TS - 1 + q), raw[(TS - 1 + q) % PD]);
#pragma unroll
  for (int u = 0; u < PD; ++u) loadA(clamp(u), buf[u]);
  __syncthreads();
  auto body = [&](long i, BufA &b, Raw &r) __attribute__((always_inline)) {
    mma(b, (int)(i % TS));
    __builtin_amdgcn_sched_barrier(0);
    loadA(clamp(i + PD), b);
    __builtin_amdgcn_sched_barrier(0);
    rstore(r, (int)((i + TS - 1) % TS));
    rload(clamp(i + TS - 1 + PD), r);
    __builtin_amdgcn_sched_barrier(0);
    __syncthreads();
  };
  long i = 0;
  for (; i + PD <= nsw; i += PD) {
#pragma unroll
    for (int u = 0; u < PD; ++u) body(i + u, buf[u], raw[(u + TS - 1) % PD]);
  }
#pragma unroll
  for (int u = 0; u < PD - 1; ++u)
    if (i + u < nsw) body(i + u, buf[u], raw[(u + TS - 1) % PD]);
}

// FP32 thin-operand words of one producer thread: B(kb + 0..3, c) for its
// column c = 16t + lc (clamped to k - 1) and kb = kbeg + 16 s + 4 lg.
struct RawF32 {
  floatx4 a;
  float sc;
};

template <int TM>
struct ThinProducerF32 {
  const float *tb, *tf1;
  unsigned dim0;
  int t, ln;
  bool on;
  __device__ ThinProducerF32(const ThinDev &th, int NT) {
    const int tid = threadIdx.x;
    on = tid < NT * 64;
    t = on ? tid >> 6 : 0;
    ln = tid & 63;
    int c = 16 * t + (ln & 15);
    c = c < th.k ? c : th.k - 1;
    tb = th.nf == 0 ? th.B + th.ldb * c : th.f[0] + th.ld[0] * c;
    tf1 = th.nf > 1 ? th.f[1] + th.ld[1] * c : nullptr;
    dim0 = th.nf > 0 ? th.dim[0] : 1u;
  }
  __device__ void load(long kb, RawF32 &r) const {
    if (!on) return;
    kb += 4 * (ln >> 4);
    if (TM == THIN_KRP2) {
      const unsigned q = (unsigned)kb / dim0;
      const unsigned d0 = (unsigned)kb - q * dim0;
      r.a = *(const floatx4 *)(tb + d0);
      r.sc = tf1[q];
    } else {
      r.a = *(const floatx4 *)(tb + kb);
      r.sc = 1.0f;
    }
  }
  template <int NT>
  __device__ void store(const RawF32 &r, floatx4 (*ts)[NT][64], int stage) const {
    if (!on) return;
    floatx4 v = r.a;
    if (TM == THIN_KRP2) {
      v[0] *= r.sc;
      v[1] *= r.sc;
      v[2] *= r.sc;
      v[3] *= r.sc;
    }
    ts[stage][t][ln] = v;
  }
};

constexpr int SH_TS = 3;  // LDS stages of the shared thin operand

// FP32 NN with the block-shared thin operand (wk == 1: wave w owns rows
// l0 = (4 blockIdx.x + w) * 64).  Layout and epilogue as nn_f32_kernel.
template <int NT, int TM>
__global__ __launch_bounds__(NTHR) void nn_f32_sh_kernel(StreamParams p) {
  constexpr int STEP = 16, PD = 2, NA = 4 * NT;
  __shared__ floatx4 ts[SH_TS][NT][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lc = lane & 15, lg = lane >> 4;
  const long l0 = ((long)blockIdx.x * 4 + wave) * 64;
  const long kbeg = (long)blockIdx.y * p.kchunk;
  const long kend = min(p.K, kbeg + p.kchunk);
  const long nfull = kend > kbeg ? (kend - kbeg) / STEP : 0;
  const float *A = (const float *)p.A;
  const long lr = l0 + 4 * lc;
  const float *base = A + (lr < p.L - 4 ? lr : p.L - 4);
  floatx4 acc[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) acc[a] = floatx4{0.f, 0.f, 0.f, 0.f};
  struct BufA {
    floatx4 x[4];
  };
  const ThinProducerF32<TM> prod(p.th, NT);
  auto loadA = [&](long s, BufA &b) {
    const long kb = kbeg + s * STEP + 4 * lg;
#pragma unroll
    for (int u = 0; u < 4; ++u) b.x[u] = *(const floatx4 *)(base + p.lda * (kb + u));
  };
  auto rload = [&](long s, RawF32 &r) { prod.load(kbeg + s * STEP, r); };
  auto rstore = [&](const RawF32 &r, int stage) { prod.template store<NT>(r, ts, stage); };
  auto mma = [&](BufA &b, int stage) {
    floatx4 a[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) a[t] = ts[stage][t][lane];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t * 4 + j] = mfma4(b.x[u][j], a[t][u], acc[t * 4 + j]);
  };
  BufA buf[PD];
  RawF32 raw[PD];
  stream_steps_shared<PD, SH_TS>(nfull, buf, raw, loadA, rload, rstore, mma);
  if (kend - kbeg > nfull * STEP) {
    // The partial last step, masked.
    long rows[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) rows[j] = lr + j < p.L ? lr + j : p.L - 1;
    BufA b;
    const long kb = kbeg + nfull * STEP + 4 * lg;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int j = 0; j < 4; ++j) b.x[u][j] = kb + u < kend ? A[rows[j] + p.lda * (kb + u)] : 0.0f;
    floatx4 a[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) a[t] = thin4(p.th, kb, kend, 16 * t + lc);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t * 4 + j] = mfma4(b.x[u][j], a[t][u], acc[t * 4 + j]);
  }
  if (p.nsplit > 1) {
    floatx4 *dst = (floatx4 *)p.part + (((long)blockIdx.y * gridDim.x + blockIdx.x) * 4 + wave) * (long)NA * 64;
#pragma unroll
    for (int a = 0; a < NA; ++a) dst[a * 64 + lane] = acc[a];
    return;
  }
  __shared__ __attribute__((aligned(16))) float ep[4][16 * EP_LD];
  epilogue_nn<NT>(p, acc, l0, ep[wave]);
}

// FP32 TN with the block-shared thin operand (wave w owns columns
// l0 = (4 blockIdx.x + w) * 16 CW).  Layout and epilogue as tn_f32_kernel.
template <int NT, int CW, int TM>
__global__ __launch_bounds__(NTHR) void tn_f32_sh_kernel(StreamParams p) {
  constexpr int STEP = 16, PD = 3, NA = CW * NT;
  __shared__ floatx4 ts[SH_TS][NT][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lc = lane & 15, lg = lane >> 4;
  const long l0 = ((long)blockIdx.x * 4 + wave) * (16 * CW);
  const long kbeg = (long)blockIdx.y * p.kchunk;
  const long kend = min(p.K, kbeg + p.kchunk);
  const long nfull = kend > kbeg ? (kend - kbeg) / STEP : 0;
  const float *A = (const float *)p.A;
  const float *col[CW];
#pragma unroll
  for (int cw = 0; cw < CW; ++cw) {
    const long l = l0 + 16 * cw + lc;
    col[cw] = A + p.lda * (l < p.L ? l : p.L - 1);
  }
  floatx4 acc[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) acc[a] = floatx4{0.f, 0.f, 0.f, 0.f};
  struct BufA {
    floatx4 x[CW];
  };
  const ThinProducerF32<TM> prod(p.th, NT);
  auto loadA = [&](long s, BufA &b) {
    const long kb = kbeg + s * STEP + 4 * lg;
#pragma unroll
    for (int cw = 0; cw < CW; ++cw) b.x[cw] = *(const floatx4 *)(col[cw] + kb);
  };
  auto rload = [&](long s, RawF32 &r) { prod.load(kbeg + s * STEP, r); };
  auto rstore = [&](const RawF32 &r, int stage) { prod.template store<NT>(r, ts, stage); };
  auto mma = [&](BufA &b, int stage) {
    floatx4 a[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) a[t] = ts[stage][t][lane];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int cw = 0; cw < CW; ++cw) acc[cw * NT + t] = mfma4(a[t][j], b.x[cw][j], acc[cw * NT + t]);
  };
  BufA buf[PD];
  RawF32 raw[PD];
  stream_steps_shared<PD, SH_TS>(nfull, buf, raw, loadA, rload, rstore, mma);
  if (kend - kbeg > nfull * STEP) {
    BufA b;
    const long kb = kbeg + nfull * STEP + 4 * lg;
#pragma unroll
    for (int cw = 0; cw < CW; ++cw)
#pragma unroll
      for (int u = 0; u < 4; ++u) b.x[cw][u] = kb + u < kend ? col[cw][kb + u] : 0.0f;
    floatx4 a[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) a[t] = thin4(p.th, kb, kend, 16 * t + lc);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int cw = 0; cw < CW; ++cw) acc[cw * NT + t] = mfma4(a[t][j], b.x[cw][j], acc[cw * NT + t]);
  }
  if (p.nsplit > 1) {
    floatx4 *dst = (floatx4 *)p.part + (((long)blockIdx.y * gridDim.x + blockIdx.x) * 4 + wave) * (long)NA * 64;
#pragma unroll
    for (int a = 0; a < NA; ++a) dst[a * 64 + lane] = acc[a];
    return;
  }
  epilogue_tn<NT, CW>(p, acc, l0);
}

// 16-bit TN (64-k steps, two 16-byte words per lane and column as the DUAL
// tn_h_kernel) with the block-shared thin operand: the packed words of one
// step (2 halves x NT tiles x 64 lanes, contiguous in the packed buffer) are
// copied into the LDS stage by the block's threads (up to four words per
// thread: two planes of the compensated operand at NT = 4).
template <typename TA, int NT, int CW, int NP>
__global__ __launch_bounds__(NTHR) void tn_h_sh_kernel(StreamParams p) {
  constexpr int NS = 2, STEP = 64, PD = 2, NA = CW * NT;
  constexpr int LW = 8 * NS;
  constexpr int WORDS1 = NS * NT * 64;            // words of one plane per step
  constexpr int WORDS = NP * WORDS1;              // <= 4 * NTHR
  constexpr int RW = (WORDS + NTHR - 1) / NTHR;  // words per thread
  __shared__ uint4 ts[SH_TS][WORDS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lc = lane & 15, lg = lane >> 4;
  const long l0 = ((long)blockIdx.x * 4 + wave) * (16 * CW);
  const long kbeg = (long)blockIdx.y * p.kchunk;
  const long kend = min(p.K, kbeg + p.kchunk);
  const long nfull = kend > kbeg ? (kend - kbeg) / STEP : 0;
  const uint16_t *A = (const uint16_t *)p.A;
  const uint16_t *col[CW];
#pragma unroll
  for (int cw = 0; cw < CW; ++cw) {
    const long l = l0 + 16 * cw + lc;
    col[cw] = A + p.lda * (l < p.L ? l : p.L - 1);
  }
  floatx4 acc[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) acc[a] = floatx4{0.f, 0.f, 0.f, 0.f};
  struct BufA {
    uint4 x[CW][NS];
  };
  // Words threadIdx.x + j NTHR (j < RW) of a step; LDS word q WORDS1 + i
  // holds word i of plane q.
  struct RawH {
    uint4 w[RW];
  };
  // Step s starts at word (kbeg / STEP + s) * WORDS1 of each packed plane.
  const uint4 *pk = p.thin16 + (kbeg / STEP) * (long)WORDS1;
  auto loadA = [&](long s, BufA &b) {
    const long kb = kbeg + s * STEP + LW * lg;
#pragma unroll
    for (int cw = 0; cw < CW; ++cw)
#pragma unroll
      for (int h = 0; h < NS; ++h) b.x[cw][h] = *(const uint4 *)(col[cw] + kb + 8 * h);
  };
  auto rload = [&](long s, RawH &r) {
#pragma unroll
    for (int j = 0; j < RW; ++j) {
      const int w = threadIdx.x + j * NTHR;
      r.w[j] = uint4{0u, 0u, 0u, 0u};
      if (w < WORDS) {
        const int q = w >= WORDS1 ? 1 : 0;
        r.w[j] = pk[q * p.plane + s * WORDS1 + (w - q * WORDS1)];
      }
    }
  };
  auto rstore = [&](const RawH &r, int stage) {
#pragma unroll
    for (int j = 0; j < RW; ++j) {
      const int w = threadIdx.x + j * NTHR;
      if (w < WORDS) ts[stage][w] = r.w[j];
    }
  };
  auto mma = [&](BufA &b, int stage) {
    const uint4 *st = ts[stage] + lane;
#pragma unroll
    for (int h = 0; h < NS; ++h)
#pragma unroll
      for (int q = 0; q < NP; ++q)
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const uint4 a = st[q * WORDS1 + (h * NT + t) * 64];
#pragma unroll
          for (int cw = 0; cw < CW; ++cw) acc[cw * NT + t] = mfma16<TA>(a, b.x[cw][h], acc[cw * NT + t]);
        }
  };
  BufA buf[PD];
  RawH raw[PD];
  stream_steps_shared<PD, SH_TS>(nfull, buf, raw, loadA, rload, rstore, mma);
  if (kend - kbeg > nfull * STEP) {
    BufA b;
    const long kb = kbeg + nfull * STEP + LW * lg;
#pragma unroll
    for (int cw = 0; cw < CW; ++cw)
#pragma unroll
      for (int h = 0; h < NS; ++h) {
        uint16_t v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const long kk = kb + 8 * h + u;
          v[u] = kk < kend ? col[cw][kk] : (uint16_t)0;
        }
        b.x[cw][h] = __builtin_bit_cast(uint4, v);
      }
#pragma unroll
    for (int h = 0; h < NS; ++h)
#pragma unroll
      for (int q = 0; q < NP; ++q)
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const uint4 a = p.thin16[q * p.plane + pack_index(kb + 8 * h, t, lc, NT)];
#pragma unroll
          for (int cw = 0; cw < CW; ++cw) acc[cw * NT + t] = mfma16<TA>(a, b.x[cw][h], acc[cw * NT + t]);
        }
  }
  if (p.nsplit > 1) {
    floatx4 *dst = (floatx4 *)p.part + (((long)blockIdx.y * gridDim.x + blockIdx.x) * 4 + wave) * (long)NA * 64;
#pragma unroll
    for (int a = 0; a < NA; ++a) dst[a * 64 + lane] = acc[a];
    return;
  }
  epilogue_tn<NT, CW>(p, acc, l0);
}

// ---------------------------------------------------------------------------
// Streaming 16-bit NN (A(L x K) L-contiguous, L % 8 == 0, 16-byte aligned
// columns).  A wave owns 128 rows: lane (lc, lg) reads the 16-byte words
// A(l0 + 8lc .. + 7, kb + 8lg + u), u = 0..7 (per load instruction four
// 256-byte column segments), and transposes its 8 x 8 block of 16-bit values
// in registers, so word j holds A(l0 + 8lc + j, kb + 8lg .. + 7): the MFMA
// A-operand with row slot lc <-> row l0 + 8lc + j.  The thin operand is the
// packed 16-bit B-operand (pack_thin16_kernel, as for 16-bit TN).  acc[8t +
// j][v] = C(l0 + 32lg + 8v + j, 16t + lc): every lane ends with 32
// consecutive rows of its column.  One k phase per group (wk == 1).
template <int NT, int NP>
struct BufNN16 {
  uint4 w[8];
  uint4 b[NP][NT];
};

__device__ inline uint4 transpose8_word(const uint4 (&w)[8], int j) {
  // halves j of w[0..7] -> one word (element u from w[u])
  const int sh = 16 * (j & 1);
  auto pick = [&](int u) -> unsigned {
    const unsigned d = (j >> 1) == 0 ? w[u].x : (j >> 1) == 1 ? w[u].y : (j >> 1) == 2 ? w[u].z : w[u].w;
    return (d >> sh) & 0xFFFFu;
  };
  uint4 o;
  o.x = pick(0) | (pick(1) << 16);
  o.y = pick(2) | (pick(3) << 16);
  o.z = pick(4) | (pick(5) << 16);
  o.w = pick(6) | (pick(7) << 16);
  return o;
}

// SH: the thin words of each step staged once per block in an LDS ring
// (stream_steps_shared; the block's four waves share the k range); SH == 2:
// the staged words are formed from the FP32 thin operand on the fly
// (ThinProducer16) rather than read from the packed buffer.  NP == 2: the
// compensated thin operand (both planes against the same A words).
template <typename TA, int NT, int SH, int NP>
__global__ __launch_bounds__(NTHR) void nn16_kernel(StreamParams p) {
  constexpr int STEP = 32, PD = 3, NA = 8 * NT;  // three steps of A in flight per wave (measured: 2 -> 3 cut the 128^4 FP16 contraction 4-6 %)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lc = lane & 15, lg = lane >> 4;
  const long l0 = ((long)blockIdx.x * 4 + wave) * 128;
  const long kbeg = (long)blockIdx.y * p.kchunk;
  const long kend = min(p.K, kbeg + p.kchunk);
  const long nfull = kend > kbeg ? (kend - kbeg) / STEP : 0;
  const uint16_t *A = (const uint16_t *)p.A;
  const long r0 = l0 + 8 * lc;  // first of this lane's 8 rows (L % 8 == 0)
  const uint16_t *base = A + (r0 < p.L ? r0 : p.L - 8);
  floatx4 acc[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) acc[a] = floatx4{0.f, 0.f, 0.f, 0.f};
  auto mma = [&](BufNN16<NT, NP> &b) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint4 o = transpose8_word(b.w, j);
#pragma unroll
      for (int q = 0; q < NP; ++q)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[8 * t + j] = mfma16<TA>(o, b.b[q][t], acc[8 * t + j]);
    }
  };
  auto load = [&](long i, BufNN16<NT, NP> &b) {
    const long kb = kbeg + i * STEP + 8 * lg;
#pragma unroll
    for (int u = 0; u < 8; ++u) b.w[u] = *(const uint4 *)(base + p.lda * (kb + u));
#pragma unroll
    for (int q = 0; q < NP; ++q)
#pragma unroll
      for (int t = 0; t < NT; ++t) b.b[q][t] = p.thin16[q * p.plane + pack_index(kb, t, lc, NT)];
  };
  if constexpr (SH != 0) {
    // LDS word (q NT + t) 64 + lane: plane q, tile t of the step; producer
    // thread tid + 256 j (j < RW) fills word tid + 256 j.
    constexpr int WORDS = NP * NT * 64, RW = (WORDS + NTHR - 1) / NTHR;
    __shared__ u32x4 ts[SH_TS][WORDS];
    struct BufA {
      uint4 w[8];
    };
    struct RawN {
      u32x4 w[RW];
    };
    const int tid = threadIdx.x;
    auto loadA = [&](long i, BufA &b) {
      const long kb = kbeg + i * STEP + 8 * lg;
#pragma unroll
      for (int u = 0; u < 8; ++u) b.w[u] = *(const uint4 *)(base + p.lda * (kb + u));
    };
    auto mma_st = [&](BufA &b, int stage) {
      uint4 bt[NP][NT];
#pragma unroll
      for (int q = 0; q < NP; ++q)
#pragma unroll
        for (int t = 0; t < NT; ++t) bt[q][t] = __builtin_bit_cast(uint4, ts[stage][(q * NT + t) * 64 + lane]);
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const uint4 o = transpose8_word(b.w, j);
#pragma unroll
        for (int q = 0; q < NP; ++q)
#pragma unroll
          for (int t = 0; t < NT; ++t) acc[8 * t + j] = mfma16<TA>(o, bt[q][t], acc[8 * t + j]);
      }
    };
    BufA buf[PD];
    if constexpr (SH == 2) {
      const ThinProducer16 prod(p.th, NT);
      RawFly raw[PD];
      auto rload = [&](long i, RawFly &r) { prod.load(kbeg + i * STEP, r); };
      auto rstore = [&](const RawFly &r, int stage) { prod.template store<TA, NT, NP>(r, ts, stage); };
      stream_steps_shared<PD, SH_TS>(nfull, buf, raw, loadA, rload, rstore, mma_st);
    } else {
    auto rload = [&](long i, RawN &r) {
#pragma unroll
      for (int j = 0; j < RW; ++j) {
        const int w = tid + j * NTHR;
        if (w < WORDS) {
          const int qt = w >> 6, pl = w & 63;
          const int q = qt / NT, pt = qt - q * NT;
          r.w[j] = __builtin_bit_cast(
              u32x4, p.thin16[q * p.plane + pack_index(kbeg + i * STEP + 8 * (pl >> 4), pt, pl & 15, NT)]);
        }
      }
    };
    auto rstore = [&](const RawN &r, int stage) {
#pragma unroll
      for (int j = 0; j < RW; ++j) {
        const int w = tid + j * NTHR;
        if (w < WORDS) ts[stage][w] = r.w[j];
      }
    };
    RawN raw[PD];
    stream_steps_shared<PD, SH_TS>(nfull, buf, raw, loadA, rload, rstore, mma_st);
    }
  } else {
    BufNN16<NT, NP> buf[PD];
    stream_steps<PD>(nfull, buf, load, mma);
  }
  if (kend - kbeg > nfull * STEP) {
    // The partial last step, masked (element loads; the packed thin operand
    // is zero past K).
    BufNN16<NT, NP> b;
    const long kb = kbeg + nfull * STEP + 8 * lg;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      uint16_t v[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = kb + u < kend ? base[e + p.lda * (kb + u)] : (uint16_t)0;
      b.w[u] = __builtin_bit_cast(uint4, v);
    }
    if constexpr (SH == 2) {
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        uint4 w[NP];
        pack8<TA, NP>(thin4(p.th, kb, kend, 16 * t + lc), thin4(p.th, kb + 4, kend, 16 * t + lc), w);
#pragma unroll
        for (int q = 0; q < NP; ++q) b.b[q][t] = w[q];
      }
    } else {
#pragma unroll
      for (int q = 0; q < NP; ++q)
#pragma unroll
        for (int t = 0; t < NT; ++t) b.b[q][t] = p.thin16[q * p.plane + pack_index(kb, t, lc, NT)];
    }
    mma(b);
  }
  if (p.nsplit > 1) {
    floatx4 *dst = (floatx4 *)p.part + (((long)blockIdx.y * gridDim.x + blockIdx.x) * 4 + wave) * (long)NA * 64;
#pragma unroll
    for (int a = 0; a < NA; ++a) dst[a * 64 + lane] = acc[a];
    return;
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int c = 16 * t + lc;
    if (c >= p.th.k) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int hf = 0; hf < 2; ++hf) {
        const long l = l0 + 32 * lg + 8 * v + 4 * hf;
        floatx4 o = floatx4{acc[8 * t + 4 * hf + 0][v], acc[8 * t + 4 * hf + 1][v],
                            acc[8 * t + 4 * hf + 2][v], acc[8 * t + 4 * hf + 3][v]} *
                    p.oscale;
        if (p.colscale) {
#pragma unroll
          for (int e = 0; e < 4; ++e) o[e] *= l + e < p.L ? p.colscale[l + e] : 1.0f;
        }
        if (!p.cmajor && p.cvec && l + 3 < p.L) {
          *(floatx4 *)(p.C + l + p.ldc * c) = o;
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            if (l + e >= p.L) continue;
            if (p.cmajor)
              p.C[c + p.ldc * (l + e)] = o[e];
            else
              p.C[l + e + p.ldc * c] = o[e];
          }
        }
      }
  }
}


// ---------------------------------------------------------------------------
// 16-bit NN: LDS-transposing kernel (A(L x K) is L-contiguous, the 16-bit
// MFMA wants 8 consecutive k per lane).  Block = 128 rows x all k columns;
// K-chunks of 64 staged through LDS (double buffered), 8x4 register blocks
// transposed on the way in.
constexpr int BL = 128;

struct GemmParams {
  const void *A;
  long lda, L, K;
  ThinDev th;
  float *C;
  long ldc;
  int cmajor;
  float *partial;
  int splits;
  long kchunk;
  float oscale;  // output multiplier (1: none)
  const float *colscale;  // scale of output row l (A's row l), or null
};

// NP == 2: the compensated thin operand (pack_thin16_kernel), formed here
// while staging B.
template <typename TA, int NT, bool VEC, int NP>
__global__ __launch_bounds__(NTHR) void nn_h_kernel(GemmParams p) {
  constexpr int BK = 64, SA = BK + 8, KP = 16 * NT;
  constexpr int BGROUPS = (BK / 8) * KP;  // groups of 8 consecutive kk of one column
  constexpr int BITER = (BGROUPS + NTHR - 1) / NTHR;
  __shared__ __attribute__((aligned(16))) uint16_t As[2][BL][SA];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[NP][2][KP][SA];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long l0 = (long)blockIdx.x * BL;
  const int split = blockIdx.y;
  const long kbeg = (long)split * p.kchunk;
  const long kend = min(p.K, kbeg + p.kchunk);
  const int nst = (int)((kend - kbeg + BK - 1) / BK);
  const uint16_t *A = (const uint16_t *)p.A;

  uint4 ra[4];
  uint4 rb[NP][BITER];

  auto load_A = [&](long k0) {
    // 8 rows x 4 kk per thread: 4 loads of 8 consecutive rows of one column.
    int rg = tid & 15, kg = tid >> 4;
    long gl = l0 + 8 * rg;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      long gk = k0 + 4 * kg + u;
      if (VEC && gl + 7 < p.L && gk < kend) {
        ra[u] = *(const uint4 *)(A + gl + p.lda * gk);
      } else {
        uint16_t t[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) t[e] = (gl + e < p.L && gk < kend) ? A[gl + e + p.lda * gk] : 0;
        ra[u] = __builtin_bit_cast(uint4, t);
      }
    }
  };
  auto store_A = [&](int buf) {
    int rg = tid & 15, kg = tid >> 4;
    uint16_t t[4][8];
#pragma unroll
    for (int u = 0; u < 4; ++u) __builtin_memcpy(t[u], &ra[u], 16);
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      uint16_t w[4] = {t[0][r], t[1][r], t[2][r], t[3][r]};
      *(uint2 *)&As[buf][8 * rg + r][4 * kg] = __builtin_bit_cast(uint2, w);
    }
  };
  auto load_B = [&](long k0) {
#pragma unroll
    for (int v = 0; v < BITER; ++v) {
      int g = tid + NTHR * v;
      if (g < BGROUPS) {
        int c = g / (BK / 8), kk = (g % (BK / 8)) * 8;
        floatx4 lo = thin4(p.th, k0 + kk, kend, c), hi = thin4(p.th, k0 + kk + 4, kend, c);
        uint16_t h[8];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          h[e] = to16<TA>(lo[e]);
          h[4 + e] = to16<TA>(hi[e]);
        }
        rb[0][v] = __builtin_bit_cast(uint4, h);
        if constexpr (NP == 2) {
          uint16_t r[8];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            r[e] = to16<TA>(lo[e] - from16<TA>(h[e]));
            r[4 + e] = to16<TA>(hi[e] - from16<TA>(h[4 + e]));
          }
          rb[NP - 1][v] = __builtin_bit_cast(uint4, r);
        }
      }
    }
  };
  auto store_B = [&](int buf) {
#pragma unroll
    for (int v = 0; v < BITER; ++v) {
      int g = tid + NTHR * v;
      if (g < BGROUPS) {
        int c = g / (BK / 8), kk = (g % (BK / 8)) * 8;
#pragma unroll
        for (int q = 0; q < NP; ++q) *(uint4 *)&Bs[q][buf][c][kk] = rb[q][v];
      }
    }
  };

  floatx4 acc[2][NT];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[r][t] = floatx4{0.f, 0.f, 0.f, 0.f};

  if (nst > 0) {
    load_A(kbeg);
    load_B(kbeg);
    store_A(0);
    store_B(0);
    __syncthreads();
  }
  for (int s = 0; s < nst; ++s) {
    const int buf = s & 1;
    const bool more = s + 1 < nst;
    if (more) {
      load_A(kbeg + (long)(s + 1) * BK);
      load_B(kbeg + (long)(s + 1) * BK);
    }
#pragma unroll
    for (int ks = 0; ks < BK / 32; ++ks) {
      int kk = ks * 32 + 8 * (lane >> 4);
      uint4 a0 = *(const uint4 *)&As[buf][wave * 32 + (lane & 15)][kk];
      uint4 a1 = *(const uint4 *)&As[buf][wave * 32 + 16 + (lane & 15)][kk];
#pragma unroll
      for (int q = 0; q < NP; ++q)
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          uint4 b = *(const uint4 *)&Bs[q][buf][16 * t + (lane & 15)][kk];
          acc[0][t] = mfma16<TA>(a0, b, acc[0][t]);
          acc[1][t] = mfma16<TA>(a1, b, acc[1][t]);
        }
    }
    if (more) {
      store_A(buf ^ 1);
      store_B(buf ^ 1);
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        long gl = l0 + wave * 32 + 16 * r + 4 * (lane >> 4) + i;
        int c = 16 * t + (lane & 15);
        if (gl >= p.L) continue;
        float v = acc[r][t][i] * (p.splits > 1 ? 1.0f : p.oscale * (p.colscale ? p.colscale[gl] : 1.0f));
        if (p.splits > 1) {
          p.partial[((long)split * p.L + gl) * KP + c] = v;
        } else if (c < p.th.k) {
          if (p.cmajor)
            p.C[c + p.ldc * gl] = v;
          else
            p.C[gl + p.ldc * c] = v;
        }
      }
}

// Sum of the 16-bit NN split-K partials (fixed order) into the final layout.
__global__ void reduce_splits_kernel(const float *partial, int splits, long L, int KP, int k,
                                     float *C, long ldc, int cmajor, float oscale, const float *rowsc) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= L * k) return;
  long gl = idx / k;
  int c = (int)(idx % k);
  float s = 0.0f;
  for (int sp = 0; sp < splits; ++sp) s += partial[((long)sp * L + gl) * KP + c];
  s *= oscale * (rowsc ? rowsc[gl] : 1.0f);
  if (cmajor)
    C[c + ldc * gl] = s;
  else
    C[gl + ldc * c] = s;
}

// ---------------------------------------------------------------------------
// Launch geometry of a streaming product: columns per wave (16*cw, TN), k
// phases per group inside a block (wk; wgrp = 4 / wk groups per block) and
// K splits across blocks.
struct Geometry {
  int cw, wgrp, wk, nsplit;
  int dual;    // 16-bit TN: 64-k steps with two 16-byte loads per lane and column
  int shared;  // thin operand staged once per block in LDS (wk == 1; 16-bit: dual)
  long kchunk, blocks_x, groups;
  long tile_bytes;  // partial tile of one group
};

Geometry make_geometry(long L, long K, int trans, int dtype, int NT, int cw, int wk, long ns,
                       int dual = 0, int shared = 0) {
  Geometry g{};
  g.shared = shared;
  if (shared) {
    wk = 1;
    dual = (trans && dtype != PLANC_F32) ? 1 : 0;
  }
  const bool nn16 = !trans && dtype != PLANC_F32;  // streaming 16-bit NN: 128 rows per wave, wk 1
  if (nn16) wk = 1;
  g.dual = trans ? dual : 0;
  const int step = dtype == PLANC_F32 ? (g.dual ? 32 : 16) : (g.dual ? 64 : 32);
  g.cw = trans ? cw : 1;
  g.groups = trans ? (L + 16 * cw - 1) / (16 * cw) : (nn16 ? (L + 127) / 128 : (L + 63) / 64);
  g.tile_bytes = trans ? (long)cw * NT * 64 * 16 : (nn16 ? 8L : 4L) * NT * 64 * 16;
  g.wk = wk;
  g.wgrp = 4 / wk;
  g.blocks_x = (g.groups + g.wgrp - 1) / g.wgrp;
  const long ksteps = (K + step - 1) / step;
  if (ns > ksteps) ns = ksteps > 0 ? ksteps : 1;
  long kchunk = (ksteps + ns - 1) / ns * step;
  g.kchunk = kchunk > 0 ? kchunk : step;
  g.nsplit = (int)((K + g.kchunk - 1) / g.kchunk);
  if (g.nsplit < 1) g.nsplit = 1;
  return g;
}

// Candidate geometries: every (cw, wk) with K splits 1, 2, 4, ... while the
// partial traffic stays under 10% of A, every wave keeps at least 8 k steps
// and the grid stays under 16k waves.  The first entry is the static choice
// used when tuning is off or impossible (graph capture).
std::vector<Geometry> candidate_geometries(const planc_ctx *ctx, long L, long K, int trans, int dtype,
                                           int NT, bool shared_ok, bool f32dual_ok, bool fly_ok = false) {
  std::vector<Geometry> out;
  const int step = dtype == PLANC_F32 ? 16 : 32;
  const long ksteps = (K + step - 1) / step;
  const int esz = dtype == PLANC_F32 ? 4 : 2;
  const double abytes = (double)L * (double)K * esz;
  if (!trans && dtype != PLANC_F32) {
    // Streaming 16-bit NN: one k phase, K splits only; per-wave or
    // block-shared thin operand.
    // With few row groups the partial-traffic bound can stop the splits
    // before the grid fills the device; the largest split count reached is
    // then a candidate too.
    for (int sh = 0; sh < (shared_ok ? (fly_ok ? 3 : 2) : 1); ++sh) {
      Geometry last{};
      bool skipped = false, split_added = false;
      for (long ns = 1; ns <= 4096; ns *= 2) {
        Geometry g = make_geometry(L, K, trans, dtype, NT, 1, 1, ns, 0, sh);
        if (ns > 1 && (g.nsplit < ns || ksteps / g.nsplit < 8)) break;
        if (ns > 1 && 2.0 * g.nsplit * g.groups * g.tile_bytes > 0.10 * abytes) break;
        const long waves = g.blocks_x * 4 * g.nsplit;
        if (ns > 1 && waves > 16L * 1024) break;
        if (!out.empty() && waves < 2L * ctx->num_cu && ns < 4096) {
          last = g;
          skipped = true;
          continue;
        }
        out.push_back(g);
        split_added = split_added || ns > 1;
      }
      if (skipped && !split_added) out.push_back(last);
    }
    return out;
  }
  std::vector<int> cws;
  if (!trans)
    cws = {1};
  else if (dtype == PLANC_F32)
    cws = {2, 4};
  else
    cws = {4};
  // Static choice: one wave per SIMD or more, fewest partials.
  {
    const int cw = cws[0];
    int wk = 1;
    Geometry g = make_geometry(L, K, trans, dtype, NT, cw, 1, 1);
    while (wk < 4 && g.blocks_x * 4 < 4L * ctx->num_cu && ksteps >= 64L * wk) {
      wk *= 2;
      g = make_geometry(L, K, trans, dtype, NT, cw, wk, 1);
    }
    long ns = 1;
    while (g.blocks_x * 4 * ns < 4L * ctx->num_cu) {
      Geometry h = make_geometry(L, K, trans, dtype, NT, cw, wk, ns * 2);
      if (2.0 * h.nsplit * h.groups * h.tile_bytes > 0.06 * abytes || ksteps / ((long)h.nsplit * wk) < 8) break;
      ns *= 2;
      g = h;
    }
    out.push_back(g);
  }
  // dual: 64-k (16-bit) or 32-k (FP32, fast thin modes) steps, whole 128-byte lines per column.
  const int ndual = (trans && (dtype != PLANC_F32 || f32dual_ok)) ? 2 : 1;
  for (int dual = 0; dual < ndual; ++dual)
  for (int cw : cws)
    for (int wk = 1; wk <= 4; wk *= 2) {
      if (wk > 1 && ksteps < 8L * wk * (dual ? 2 : 1)) break;
      for (long ns = 1; ns <= 4096; ns *= 2) {
        Geometry g = make_geometry(L, K, trans, dtype, NT, cw, wk, ns, dual);
        if (ns > 1 && (g.nsplit < ns || ksteps / ((long)g.nsplit * wk * (dual ? 2 : 1)) < 8)) break;
        if (ns > 1 && 2.0 * g.nsplit * g.groups * g.tile_bytes > 0.10 * abytes) break;
        const long waves = g.blocks_x * 4 * g.nsplit;
        if (ns > 1 && waves > 16L * 1024) break;
        if (waves < 2L * ctx->num_cu && ns < 4096) continue;  // too few waves: try more splits
        out.push_back(g);
      }
    }
  // Block-shared thin operand: one k phase, four groups per block.
  const std::vector<int> scws = trans ? std::vector<int>{2, 4} : std::vector<int>{1};
  if (shared_ok)
    for (int cw : scws) {
      const int sstep = dtype == PLANC_F32 ? 16 : 64;
      const long sk = (K + sstep - 1) / sstep;
      for (long ns = 1; ns <= 4096; ns *= 2) {
        Geometry g = make_geometry(L, K, trans, dtype, NT, cw, 1, ns, 1, 1);
        if (ns > 1 && (g.nsplit < ns || sk / g.nsplit < 8)) break;
        if (ns > 1 && 2.0 * g.nsplit * g.groups * g.tile_bytes > 0.10 * abytes) break;
        const long waves = g.blocks_x * 4 * g.nsplit;
        if (ns > 1 && waves > 16L * 1024) break;
        if (waves < 2L * ctx->num_cu && ns < 4096) continue;
        out.push_back(g);
      }
    }
  return out;
}

int thin_mode(const ThinDev &th, bool avec, const float *rowscale) {
  if (!avec || rowscale || !th.vec) return THIN_SAFE;
  if (th.nf <= 1) return THIN_PLAIN;
  if (th.nf == 2) return THIN_KRP2;
  return THIN_SAFE;
}

template <int NT, int CW>
void launch_tn_f32(int tm, int dual, dim3 grid, hipStream_t st, const StreamParams &p) {
  if (dual && tm == THIN_PLAIN)
    tn_f32d_kernel<NT, CW, THIN_PLAIN><<<grid, NTHR, 0, st>>>(p);
  else if (dual)
    tn_f32d_kernel<NT, CW, THIN_KRP2><<<grid, NTHR, 0, st>>>(p);
  else if (tm == THIN_PLAIN)
    tn_f32_kernel<NT, CW, THIN_PLAIN><<<grid, NTHR, 0, st>>>(p);
  else if (tm == THIN_KRP2)
    tn_f32_kernel<NT, CW, THIN_KRP2><<<grid, NTHR, 0, st>>>(p);
  else
    tn_f32_kernel<NT, CW, THIN_SAFE><<<grid, NTHR, 0, st>>>(p);
}
template <int NT>
void launch_nn_f32(int tm, dim3 grid, hipStream_t st, const StreamParams &p) {
  if (tm == THIN_PLAIN)
    nn_f32_kernel<NT, THIN_PLAIN><<<grid, NTHR, 0, st>>>(p);
  else if (tm == THIN_KRP2)
    nn_f32_kernel<NT, THIN_KRP2><<<grid, NTHR, 0, st>>>(p);
  else
    nn_f32_kernel<NT, THIN_SAFE><<<grid, NTHR, 0, st>>>(p);
}
template <typename TA, int NT, int NP>
void launch_tn_h(bool safe, int dual, dim3 grid, hipStream_t st, const StreamParams &p) {
  if (safe)
    tn_h_kernel<TA, NT, 4, 1, 0, NP><<<grid, NTHR, 0, st>>>(p);
  else if (dual)
    tn_h_kernel<TA, NT, 4, 0, 1, NP><<<grid, NTHR, 0, st>>>(p);
  else
    tn_h_kernel<TA, NT, 4, 0, 0, NP><<<grid, NTHR, 0, st>>>(p);
}
template <int NT, int CW>
void launch_tn_f32_sh(int tm, dim3 grid, hipStream_t st, const StreamParams &p) {
  if (tm == THIN_PLAIN)
    tn_f32_sh_kernel<NT, CW, THIN_PLAIN><<<grid, NTHR, 0, st>>>(p);
  else
    tn_f32_sh_kernel<NT, CW, THIN_KRP2><<<grid, NTHR, 0, st>>>(p);
}
template <int NT>
void launch_nn_f32_sh(int tm, dim3 grid, hipStream_t st, const StreamParams &p) {
  if (tm == THIN_PLAIN)
    nn_f32_sh_kernel<NT, THIN_PLAIN><<<grid, NTHR, 0, st>>>(p);
  else
    nn_f32_sh_kernel<NT, THIN_KRP2><<<grid, NTHR, 0, st>>>(p);
}
template <typename TA, int CW, int NP>
void launch_tn_h_sh(int NT, dim3 grid, hipStream_t st, const StreamParams &p) {
  if (NT == 1) tn_h_sh_kernel<TA, 1, CW, NP><<<grid, NTHR, 0, st>>>(p);
  else if (NT == 2) tn_h_sh_kernel<TA, 2, CW, NP><<<grid, NTHR, 0, st>>>(p);
  else if (NT == 3) tn_h_sh_kernel<TA, 3, CW, NP><<<grid, NTHR, 0, st>>>(p);
  else tn_h_sh_kernel<TA, 4, CW, NP><<<grid, NTHR, 0, st>>>(p);
}
template <int NP>
void launch_nn16(int dtype, int NT, int shared, dim3 grid, hipStream_t st, const StreamParams &p) {
#define NN16(TA, SHV)                                                        \
  do {                                                                       \
    if (NT == 1) nn16_kernel<TA, 1, SHV, NP><<<grid, NTHR, 0, st>>>(p);      \
    else if (NT == 2) nn16_kernel<TA, 2, SHV, NP><<<grid, NTHR, 0, st>>>(p); \
    else if (NT == 3) nn16_kernel<TA, 3, SHV, NP><<<grid, NTHR, 0, st>>>(p); \
    else nn16_kernel<TA, 4, SHV, NP><<<grid, NTHR, 0, st>>>(p);              \
  } while (0)
  if (dtype == PLANC_BF16) {
    if (shared == 2) NN16(bf16_t, 2);
    else if (shared) NN16(bf16_t, 1);
    else NN16(bf16_t, 0);
  } else {
    if (shared == 2) NN16(f16_t, 2);
    else if (shared) NN16(f16_t, 1);
    else NN16(f16_t, 0);
  }
#undef NN16
}
template <int NP>
void launch_tn16(int dtype, int NT, const Geometry &g, dim3 grid, hipStream_t st, const StreamParams &p) {
  if (g.shared) {
    if (dtype == PLANC_BF16) {
      if (g.cw == 4) launch_tn_h_sh<bf16_t, 4, NP>(NT, grid, st, p);
      else launch_tn_h_sh<bf16_t, 2, NP>(NT, grid, st, p);
    } else {
      if (g.cw == 4) launch_tn_h_sh<f16_t, 4, NP>(NT, grid, st, p);
      else launch_tn_h_sh<f16_t, 2, NP>(NT, grid, st, p);
    }
    return;
  }
  const bool safe = !p.avec;
  if (dtype == PLANC_BF16) {
    if (NT == 1) launch_tn_h<bf16_t, 1, NP>(safe, g.dual, grid, st, p);
    else if (NT == 2) launch_tn_h<bf16_t, 2, NP>(safe, g.dual, grid, st, p);
    else if (NT == 3) launch_tn_h<bf16_t, 3, NP>(safe, g.dual, grid, st, p);
    else launch_tn_h<bf16_t, 4, NP>(safe, g.dual, grid, st, p);
  } else {
    if (NT == 1) launch_tn_h<f16_t, 1, NP>(safe, g.dual, grid, st, p);
    else if (NT == 2) launch_tn_h<f16_t, 2, NP>(safe, g.dual, grid, st, p);
    else if (NT == 3) launch_tn_h<f16_t, 3, NP>(safe, g.dual, grid, st, p);
    else launch_tn_h<f16_t, 4, NP>(safe, g.dual, grid, st, p);
  }
}
template <int NT, int CW>
void launch_reduce_tn(dim3 g, hipStream_t st, const StreamParams &p, int NA, long ngroups) {
  splitk_reduce_kernel<0, NT, CW><<<dim3(g.x, (unsigned)((NA + 3) / 4)), 256, 0, st>>>(p, NA, ngroups);
}

// Launches one streaming product (and its split-K reduction) with geometry g.
int launch_stream(planc_ctx *ctx, int dtype, int trans, int NT, int tm, StreamParams p, const Geometry &g) {
  p.wgrp = g.wgrp;
  p.wk = g.wk;
  p.kchunk = g.kchunk;
  p.nsplit = g.nsplit;
  // Partial tiles are indexed by padded group (blocks_x * wgrp >= groups).
  const long gpad = g.blocks_x * g.wgrp;
  if (g.nsplit > 1) {
    p.part = (float *)ws_get(ctx, 0, (size_t)g.nsplit * gpad * g.tile_bytes);
    if (!p.part) return set_error(ctx, PLANC_ENOMEM, "thin_gemm: workspace");
  }
  dim3 grid((unsigned)g.blocks_x, (unsigned)g.nsplit);
  hipStream_t st = ctx->stream;
  if (!trans && dtype != PLANC_F32) {
    if (p.np == 2) launch_nn16<2>(dtype, NT, g.shared, grid, st, p);
    else launch_nn16<1>(dtype, NT, g.shared, grid, st, p);
    PLANC_LAUNCH_CHECK(ctx);
    if (g.nsplit > 1) {
      const int NA = 8 * NT;
      splitk_reduce_nn_kernel<2><<<dim3((unsigned)gpad, (unsigned)NT), 256, 0, st>>>(p, NA, gpad);
      PLANC_LAUNCH_CHECK(ctx);
    }
    return PLANC_OK;
  }
  if (g.shared && trans && dtype == PLANC_F32) {
#define TNS(NTV)                                        \
  do {                                                  \
    if (g.cw == 4)                                      \
      launch_tn_f32_sh<NTV, 4>(tm, grid, st, p);        \
    else                                                \
      launch_tn_f32_sh<NTV, 2>(tm, grid, st, p);        \
  } while (0)
    if (NT == 1) TNS(1);
    else if (NT == 2) TNS(2);
    else if (NT == 3) TNS(3);
    else TNS(4);
#undef TNS
  } else if (trans && dtype != PLANC_F32) {
    if (p.np == 2) launch_tn16<2>(dtype, NT, g, grid, st, p);
    else launch_tn16<1>(dtype, NT, g, grid, st, p);
  } else if (g.shared) {
    if (NT == 1) launch_nn_f32_sh<1>(tm, grid, st, p);
    else if (NT == 2) launch_nn_f32_sh<2>(tm, grid, st, p);
    else if (NT == 3) launch_nn_f32_sh<3>(tm, grid, st, p);
    else launch_nn_f32_sh<4>(tm, grid, st, p);
  } else if (trans && dtype == PLANC_F32) {
#define TNF(NTV)                                        \
  do {                                                  \
    if (g.cw == 4)                                      \
      launch_tn_f32<NTV, 4>(tm, g.dual, grid, st, p);   \
    else                                                \
      launch_tn_f32<NTV, 2>(tm, g.dual, grid, st, p);   \
  } while (0)
    if (NT == 1) TNF(1);
    else if (NT == 2) TNF(2);
    else if (NT == 3) TNF(3);
    else TNF(4);
#undef TNF
  } else {
    if (NT == 1) launch_nn_f32<1>(tm, grid, st, p);
    else if (NT == 2) launch_nn_f32<2>(tm, grid, st, p);
    else if (NT == 3) launch_nn_f32<3>(tm, grid, st, p);
    else launch_nn_f32<4>(tm, grid, st, p);
  }
  PLANC_LAUNCH_CHECK(ctx);
  if (g.nsplit > 1) {
    const int NA = trans ? g.cw * NT : 4 * NT;
    dim3 rgrid((unsigned)gpad, (unsigned)NA);
    if (!trans) {
      splitk_reduce_nn_kernel<1><<<dim3(rgrid.x, (unsigned)NT), 256, 0, st>>>(p, NA, gpad);
    } else {
#define RED(NTV)                                                 \
  do {                                                           \
    if (g.cw == 4)                                               \
      launch_reduce_tn<NTV, 4>(rgrid, st, p, NA, gpad);          \
    else                                                         \
      launch_reduce_tn<NTV, 2>(rgrid, st, p, NA, gpad);          \
  } while (0)
      if (NT == 1) RED(1);
      else if (NT == 2) RED(2);
      else if (NT == 3) RED(3);
      else RED(4);
#undef RED
    }
    PLANC_LAUNCH_CHECK(ctx);
  }
  return PLANC_OK;
}

bool stream_capturing(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess) return true;
  return cs != hipStreamCaptureStatusNone;
}

// Picks the geometry of a product shape: cached per context; otherwise
// every candidate is timed on the live operands (one warm-up launch and two
// timed launches each; the products are idempotent) unless tuning is off
// (PLANC_AUTOTUNE=0) or the stream is being captured into a graph, where the
// static choice is used.  PLANC_TUNE_LOG=1 prints each candidate's time.
// fly_ok: the shared 16-bit NN kernel may form its thin operand on the fly
// (no packed buffer).  pack() launches the packing of the thin operand that
// the other 16-bit candidates read: it runs before the first candidate is
// timed, and with fly_ok it is timed too and its time charged to every
// candidate that needs it.
int pick_geometry(planc_ctx *ctx, int dtype, int trans, int NT, int tm, const StreamParams &p, bool fly_ok,
                  const std::function<int()> &pack, Geometry *out) {
  // PLANC_STREAM_VARIANT restricts the tuner to one kernel family (used to
  // test each path): "plain" (per-wave thin operand, 16/32-k steps), "dual"
  // (per-wave, 32-k FP32 / 64-k 16-bit TN steps), "shared" (block-shared
  // LDS thin operand) or "fly" (block-shared, formed on the fly: 16-bit NN),
  // each where it applies; unset: all.
  const char *venv = getenv("PLANC_STREAM_VARIANT");
  const std::string vs = venv ? venv : "";
  const int smode = vs == "plain" ? 0 : (vs == "shared" ? 1 : (vs == "dual" ? 3 : (vs == "fly" ? 4 : 2)));
  const std::vector<long> key = {dtype, trans, NT, tm, p.L, p.K, p.lda, p.th.k, p.avec, p.cmajor, smode, p.np,
                                 fly_ok ? 1 : 0};
  for (const auto &e : ctx->gemm_tuned)
    if (e.first == key) {
      const auto &v = e.second;
      *out = make_geometry(p.L, p.K, trans, dtype, NT, (int)v[0], (int)v[1], v[2], (int)v[3], (int)v[4]);
      return PLANC_OK;
    }
  const bool shared_ok = (dtype == PLANC_F32 ? tm != THIN_SAFE : p.avec != 0) && smode != 0 && smode != 3;
  std::vector<Geometry> cands = candidate_geometries(ctx, p.L, p.K, trans, dtype, NT, shared_ok,
                                                     dtype == PLANC_F32 && tm != THIN_SAFE, fly_ok && shared_ok);
  if (smode != 2) {
    std::vector<Geometry> sel;
    for (const Geometry &g : cands)
      if (smode == 4 ? g.shared == 2
                     : smode == 1 ? g.shared != 0 : (g.shared == 0 && (smode == 3) == (g.dual != 0)))
        sel.push_back(g);
    if (!sel.empty()) cands.swap(sel);
  }
  const char *env = getenv("PLANC_AUTOTUNE");
  const bool tune = !(env && env[0] == '0') && !stream_capturing(ctx->stream) && cands.size() > 1;
  const char *lenv = getenv("PLANC_TUNE_LOG");  // 1: print every candidate's time to stderr
  const bool log = lenv && lenv[0] == '1';
  Geometry best = cands[0];
  if (tune) {
    hipEvent_t e0, e1;
    PLANC_HIP_CHECK(ctx, hipEventCreate(&e0));
    PLANC_HIP_CHECK(ctx, hipEventCreate(&e1));
    float pack_ms = 0.0f;  // two pack launches, as each candidate is timed over two launches
    if (dtype != PLANC_F32) {
      int prc = pack();
      if (prc) return prc;
      bool any_fly = false;
      for (const Geometry &g : cands) any_fly = any_fly || g.shared == 2;
      if (any_fly) {
        PLANC_HIP_CHECK(ctx, hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < 2 && !prc; ++r) prc = pack();
        if (prc) return prc;
        PLANC_HIP_CHECK(ctx, hipEventRecord(e1, ctx->stream));
        PLANC_HIP_CHECK(ctx, hipEventSynchronize(e1));
        PLANC_HIP_CHECK(ctx, hipEventElapsedTime(&pack_ms, e0, e1));
      }
    }
    float best_ms = 0.0f;
    bool have = false;
    for (const Geometry &g : cands) {
      int rc = launch_stream(ctx, dtype, trans, NT, tm, p, g);  // warm-up (and workspace growth)
      if (rc) return rc;
      PLANC_HIP_CHECK(ctx, hipEventRecord(e0, ctx->stream));
      for (int r = 0; r < 2; ++r) {
        rc = launch_stream(ctx, dtype, trans, NT, tm, p, g);
        if (rc) return rc;
      }
      PLANC_HIP_CHECK(ctx, hipEventRecord(e1, ctx->stream));
      PLANC_HIP_CHECK(ctx, hipEventSynchronize(e1));
      float ms = 0.0f;
      PLANC_HIP_CHECK(ctx, hipEventElapsedTime(&ms, e0, e1));
      if (g.shared != 2) ms += pack_ms;
      if (log)
        fprintf(stderr,
                "planc tune dtype %d trans %d L %ld K %ld k %d: cw %d wk %d nsplit %ld dual %d shared %d -> %.2f us\n",
                dtype, trans, p.L, p.K, p.th.k, g.cw, g.wk, (long)g.nsplit, g.dual, g.shared, ms * 500.0f);
      if (!have || ms < best_ms) {
        best_ms = ms;
        best = g;
        have = true;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  if (!stream_capturing(ctx->stream))
    ctx->gemm_tuned.push_back({key, {best.cw, best.wk, (long)best.nsplit, best.dual, best.shared}});
  *out = best;
  return PLANC_OK;
}

int fill_thin(planc_ctx *ctx, ThinDev &th, const float *B, long ldb, const KrpDesc *krp,
              const float *rowscale, long K, int c0, int kc) {
  th = ThinDev{};
  th.k = kc;
  th.rowscale = rowscale;
  bool vec = rowscale == nullptr || ((uintptr_t)rowscale % 16 == 0);
  if (krp && krp->nf > 0) {
    if (krp->nf > 6) return set_error(ctx, PLANC_EINVAL, "thin_gemm: more than 6 KRP factors");
    long prod = 1;
    for (int q = 0; q < krp->nf; ++q) prod *= krp->dim[q];
    if (prod != K) return set_error(ctx, PLANC_EINVAL, "thin_gemm: KRP rows != K");
    if (K >= (1L << 31)) return set_error(ctx, PLANC_EINVAL, "thin_gemm: KRP with 2^31 rows or more");
    th.nf = krp->nf;
    for (int q = 0; q < krp->nf; ++q) {
      th.f[q] = krp->f[q] + krp->ld[q] * c0;
      th.dim[q] = (unsigned)krp->dim[q];
      th.ld[q] = krp->ld[q];
    }
    vec = vec && krp->dim[0] % 4 == 0 && krp->ld[0] % 4 == 0 && ((uintptr_t)krp->f[0] % 16 == 0);
  } else {
    if (!B) return set_error(ctx, PLANC_EINVAL, "thin_gemm: null B");
    if (K >= (1L << 32)) return set_error(ctx, PLANC_EINVAL, "thin_gemm: K too large");
    th.nf = 0;
    th.B = B + ldb * c0;
    th.ldb = ldb;
    vec = vec && ldb % 4 == 0 && ((uintptr_t)B % 16 == 0);
  }
  th.vec = vec ? 1 : 0;
  return PLANC_OK;
}

}  // namespace

int thin_gemm(planc_ctx *ctx, int dtype, int trans, const void *A, long lda, long L, long K,
              const float *B, long ldb, const KrpDesc *krp, const float *rowscale,
              const float *colscale, int k, float *C, long ldc, int cmajor, int splits_hint, float oscale) {
  if (L < 0 || K < 0 || k <= 0 || !A || !C) return set_error(ctx, PLANC_EINVAL, "thin_gemm: bad shape/pointer");
  if (dtype == PLANC_F32 && oscale != 1.0f) return set_error(ctx, PLANC_EINVAL, "thin_gemm: output scale with FP32");
  if (dtype == PLANC_F32 && !trans && colscale)
    return set_error(ctx, PLANC_EINVAL, "thin_gemm: output row scales with an FP32 NN product");
  if (dtype != PLANC_F32 && dtype != PLANC_BF16 && dtype != PLANC_F16)
    return set_error(ctx, PLANC_EINVAL, "thin_gemm: bad dtype");
  if (L == 0) return PLANC_OK;
  const int esz = dtype == PLANC_F32 ? 4 : 2;
  const int vecw = 16 / esz;
  const bool avec = (lda % vecw == 0) && (((uintptr_t)A) % 16 == 0);
  const int np = dtype != PLANC_F32 && ctx->thin_planes == 2 ? 2 : 1;
  for (int c0 = 0; c0 < k; c0 += 64) {
    const int kc = (k - c0 < 64) ? (k - c0) : 64;
    const int NT = (kc + 15) / 16;
    ThinDev th;
    int rc = fill_thin(ctx, th, B, ldb, krp, rowscale, K, c0, kc);
    if (rc) return rc;
    float *Cs = cmajor ? C + c0 : C + ldc * c0;
    if (K == 0) {
      // Empty contraction: C = 0.
      PLANC_HIP_CHECK(ctx, hipMemset2DAsync(Cs, sizeof(float) * ldc, 0,
                                            sizeof(float) * (cmajor ? kc : L), cmajor ? L : kc,
                                            ctx->stream));
      continue;
    }
    if (dtype != PLANC_F32 && !trans && !(avec && L % 8 == 0)) {
      // 16-bit NN with rows that do not come in 16-byte groups: the
      // LDS-transposing kernel.
      GemmParams p{};
      p.A = A;
      p.lda = lda;
      p.L = L;
      p.K = K;
      p.th = th;
      p.C = Cs;
      p.ldc = ldc;
      p.cmajor = cmajor;
      p.oscale = oscale;
      p.colscale = colscale;
      const int BK = 64;
      long lblocks = (L + BL - 1) / BL;
      long want = (4L * ctx->num_cu + lblocks - 1) / lblocks;
      long maxs = (K + 4L * BK - 1) / (4L * BK);
      if (want > maxs) want = maxs;
      if (want < 1) want = 1;
      if (splits_hint > 0) want = splits_hint;
      if (want > 4096) want = 4096;
      long kchunk = (K + want - 1) / want;
      kchunk = ((kchunk + BK - 1) / BK) * BK;
      int splits = (int)((K + kchunk - 1) / kchunk);
      p.splits = splits;
      p.kchunk = kchunk;
      const int KP = 16 * NT;
      if (splits > 1) {
        p.partial = (float *)ws_get(ctx, 0, sizeof(float) * (size_t)splits * L * KP);
        if (!p.partial) return set_error(ctx, PLANC_ENOMEM, "thin_gemm: workspace");
      }
      dim3 grid((unsigned)lblocks, (unsigned)splits);
      const bool vec = avec;
#define NNH(TA, NTV)                                                                   \
  do {                                                                                 \
    if (np == 2 && vec)                                                                \
      nn_h_kernel<TA, NTV, true, 2><<<grid, NTHR, 0, ctx->stream>>>(p);                \
    else if (np == 2)                                                                  \
      nn_h_kernel<TA, NTV, false, 2><<<grid, NTHR, 0, ctx->stream>>>(p);               \
    else if (vec)                                                                      \
      nn_h_kernel<TA, NTV, true, 1><<<grid, NTHR, 0, ctx->stream>>>(p);                \
    else                                                                               \
      nn_h_kernel<TA, NTV, false, 1><<<grid, NTHR, 0, ctx->stream>>>(p);               \
  } while (0)
      if (dtype == PLANC_BF16) {
        if (NT == 1) NNH(bf16_t, 1);
        else if (NT == 2) NNH(bf16_t, 2);
        else if (NT == 3) NNH(bf16_t, 3);
        else NNH(bf16_t, 4);
      } else {
        if (NT == 1) NNH(f16_t, 1);
        else if (NT == 2) NNH(f16_t, 2);
        else if (NT == 3) NNH(f16_t, 3);
        else NNH(f16_t, 4);
      }
#undef NNH
      PLANC_LAUNCH_CHECK(ctx);
      if (splits > 1) {
        long tot = L * kc;
        unsigned nb = (unsigned)((tot + 255) / 256);
        reduce_splits_kernel<<<nb, 256, 0, ctx->stream>>>(p.partial, splits, L, KP, kc, Cs, ldc, cmajor,
                                                          oscale, colscale);
        PLANC_LAUNCH_CHECK(ctx);
      }
      continue;
    }

    // Streaming kernels.
    StreamParams p{};
    p.A = A;
    p.lda = lda;
    p.L = L;
    p.K = K;
    p.th = th;
    p.colscale = colscale;
    p.C = Cs;
    p.ldc = ldc;
    p.cmajor = cmajor;
    p.avec = avec ? 1 : 0;
    p.cvec = (ldc % 4 == 0 && ((uintptr_t)Cs % 16 == 0)) ? 1 : 0;
    p.oscale = oscale;
    int tm = thin_mode(th, avec, rowscale);
    if (!trans && L % 4 != 0) tm = THIN_SAFE;
    // 16-bit TN / NN: the thin operand packed into the 16-bit MFMA operand
    // layout (8 consecutive k of one column per lane) once per product,
    // unless the chosen kernel forms it on the fly.
    bool packed = false;
    auto pack = [&]() -> int {
      if (dtype == PLANC_F32) return PLANC_OK;
      const long nsuper = (K + 63) / 64;
      const long tot = nsuper * 2 * NT * 64;  // words of one plane
      uint4 *pk = (uint4 *)ws_get(ctx, 6, (size_t)np * tot * 16);
      if (!pk) return set_error(ctx, PLANC_ENOMEM, "thin_gemm: packed operand");
      const unsigned nb = (unsigned)((tot / 2 + 255) / 256);
      if (dtype == PLANC_BF16 && np == 2)
        pack_thin16_kernel<bf16_t, 2><<<nb, 256, 0, ctx->stream>>>(th, K, NT, nsuper, pk, tot);
      else if (dtype == PLANC_BF16)
        pack_thin16_kernel<bf16_t, 1><<<nb, 256, 0, ctx->stream>>>(th, K, NT, nsuper, pk, tot);
      else if (np == 2)
        pack_thin16_kernel<f16_t, 2><<<nb, 256, 0, ctx->stream>>>(th, K, NT, nsuper, pk, tot);
      else
        pack_thin16_kernel<f16_t, 1><<<nb, 256, 0, ctx->stream>>>(th, K, NT, nsuper, pk, tot);
      PLANC_LAUNCH_CHECK(ctx);
      p.thin16 = pk;
      p.plane = tot;
      packed = true;
      return PLANC_OK;
    };
    p.np = np;
    // The on-the-fly operand reads B (or the first KRP factor, with the
    // second's entry shared by 8 consecutive rows) with 16-byte loads.
    const bool fly_ok = dtype != PLANC_F32 && !trans && th.vec && th.nf <= 2 &&
                        (th.nf < 2 || th.dim[0] % 8 == 0);
    Geometry g;
    if (splits_hint > 0) {
      g = make_geometry(L, K, trans, dtype, NT, trans && dtype == PLANC_F32 ? 2 : 4, 1, splits_hint);
    } else {
      rc = pick_geometry(ctx, dtype, trans, NT, tm, p, fly_ok, pack, &g);
      if (rc) return rc;
    }
    if (!packed && !(g.shared == 2 && !trans && dtype != PLANC_F32)) {
      rc = pack();
      if (rc) return rc;
    }
    rc = launch_stream(ctx, dtype, trans, NT, tm, p, g);
    if (rc) return rc;
  }
  return PLANC_OK;
}

}  // namespace planc
