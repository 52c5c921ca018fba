// Thin-rank products on MFMA: the bulk arithmetic of NMF and NTF.
//
//   NN:  C (L x k)  = A (L x K) * B (K x k)     A*H, X_(LxR)*KRP_R, mode-0 MTTKRP
//   TN:  C (k x L)  = B^T (k x K) * A (K x L)   W^T*A, KRP_L^T*X_(LxR)
//
// A is the big column-major operand (FP32, BF16 or FP16) streamed from HBM
// exactly once per product; B is the thin K x k operand, either a factor
// matrix or a Khatri-Rao product of up to six factors generated on the fly
// (never stored).  k <= 64 per launch (wider k loops over 64-column slices).
//
// Block = 4 wavefronts, tile = 128 rows of the MFMA A-operand x all k columns.
// Each K-chunk is staged through LDS (double buffered, one barrier per
// chunk): the A tile with 16-byte coalesced loads, the B tile computed by the
// block.  FP32 uses v_mfma_f32_16x16x4f32, 16-bit inputs
// v_mfma_f32_16x16x32_{bf16,f16}, all accumulating in FP32.  When the long
// dimension alone gives too few blocks, K is split across blocks and the
// partials are summed by a second kernel in fixed split order (deterministic,
// no atomics).
#include "common.h"

namespace planc {

namespace {

constexpr int BL = 128;
constexpr int NTHR = 256;

struct KrpDev {
  int nf;
  const float *f[6];
  int dim[6];
  long ld[6];
};

struct GemmParams {
  const void *A;
  long lda, L, K;
  const float *B;
  long ldb;
  KrpDev krp;
  const float *rowscale;
  const float *colscale;
  int k;
  float *C;
  long ldc;
  int cmajor;  // 1: C is k x L (C[c + ldc*l]); 0: C is L x k (C[l + ldc*c])
  float *partial;
  int splits;
  long kchunk;
};

__device__ inline void krp_digits(const KrpDev &kd, long g, int *d) {
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    if (q < kd.nf) {
      d[q] = (int)(g % kd.dim[q]);
      g /= kd.dim[q];
    }
  }
}

__device__ inline float krp_value(const KrpDev &kd, const int *d, int c) {
  float v = 1.0f;
#pragma unroll
  for (int q = 5; q >= 0; --q)
    if (q < kd.nf) v *= kd.f[q][d[q] + kd.ld[q] * c];
  return v;
}

__device__ inline void krp_next(const KrpDev &kd, int *d) {
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    if (q < kd.nf) {
      if (++d[q] < kd.dim[q]) return;
      d[q] = 0;
    }
  }
}

// Fills v[0..n) with B(gk0 + u, c) for u < n (0 outside K / k).
template <int n>
__device__ inline void b_values(const GemmParams &p, long gk0, long kend, int c, float *v) {
  if (c >= p.k) {
#pragma unroll
    for (int u = 0; u < n; ++u) v[u] = 0.0f;
    return;
  }
  if (p.krp.nf == 0) {
#pragma unroll
    for (int u = 0; u < n; ++u) {
      long gk = gk0 + u;
      float x = 0.0f;
      if (gk < kend) {
        x = p.B[gk + p.ldb * c];
        if (p.rowscale) x *= p.rowscale[gk];
      }
      v[u] = x;
    }
  } else {
    int d[6];
    krp_digits(p.krp, gk0 < kend ? gk0 : 0, d);
#pragma unroll
    for (int u = 0; u < n; ++u) {
      long gk = gk0 + u;
      float x = 0.0f;
      if (gk < kend) {
        x = krp_value(p.krp, d, c);
        if (p.rowscale) x *= p.rowscale[gk];
        krp_next(p.krp, d);
      }
      v[u] = x;
    }
  }
}

template <int NT, int TRANS>
__device__ inline void epilogue(const GemmParams &p, floatx4 (&acc)[2][NT], long l0, int split) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int KP = 16 * NT;
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        long gl = l0 + wave * 32 + 16 * r + 4 * (lane >> 4) + i;
        int c = 16 * t + (lane & 15);
        if (gl >= p.L) continue;
        float v = acc[r][t][i];
        if (p.splits > 1) {
          p.partial[((long)split * p.L + gl) * KP + c] = v;
        } else if (c < p.k) {
          if (TRANS && p.colscale) v *= p.colscale[gl];
          if (p.cmajor)
            p.C[c + p.ldc * gl] = v;
          else
            p.C[gl + p.ldc * c] = v;
        }
      }
}

// ---------------------------------------------------------------------------
// FP32 operands: v_mfma_f32_16x16x4f32.  LDS: As[kk][row], Bs[kk][col].
template <int NT, int TRANS, bool VEC>
__global__ __launch_bounds__(NTHR) void gemm_f32_kernel(GemmParams p) {
  constexpr int BK = 32, SA = BL + 16, KP = 16 * NT, SB = ((KP + 15) / 32) * 32 + 16;
  constexpr int BGROUPS = (BK / 4) * KP;  // groups of 4 consecutive kk of one column
  constexpr int BITER = (BGROUPS + NTHR - 1) / NTHR;
  __shared__ float As[2][BK][SA];
  __shared__ float Bs[2][BK][SB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long l0 = (long)blockIdx.x * BL;
  const int split = blockIdx.y;
  const long kbeg = (long)split * p.kchunk;
  const long kend = min(p.K, kbeg + p.kchunk);
  const int nst = (int)((kend - kbeg + BK - 1) / BK);
  const float *A = (const float *)p.A;

  floatx4 ra[4];
  float rb[BITER][4];

  auto load_A = [&](long k0) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      int q = tid + NTHR * v;
      if (!TRANS) {
        int kk = q >> 5, l = (q & 31) * 4;
        long gl = l0 + l, gk = k0 + kk;
        if (VEC && gl + 3 < p.L && gk < kend) {
          ra[v] = *(const floatx4 *)(A + gl + p.lda * gk);
        } else {
#pragma unroll
          for (int u = 0; u < 4; ++u) ra[v][u] = (gl + u < p.L && gk < kend) ? A[gl + u + p.lda * gk] : 0.0f;
        }
      } else {
        int j = q >> 3, kk = (q & 7) * 4;
        long gj = l0 + j, gk = k0 + kk;
        if (VEC && gj < p.L && gk + 3 < kend) {
          ra[v] = *(const floatx4 *)(A + gk + p.lda * gj);
        } else {
#pragma unroll
          for (int u = 0; u < 4; ++u) ra[v][u] = (gj < p.L && gk + u < kend) ? A[gk + u + p.lda * gj] : 0.0f;
        }
      }
    }
  };
  auto store_A = [&](int buf) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      int q = tid + NTHR * v;
      if (!TRANS) {
        int kk = q >> 5, l = (q & 31) * 4;
        *(floatx4 *)&As[buf][kk][l] = ra[v];
      } else {
        int j = q >> 3, kk = (q & 7) * 4;
#pragma unroll
        for (int u = 0; u < 4; ++u) As[buf][kk + u][j] = ra[v][u];
      }
    }
  };
  auto load_B = [&](long k0) {
#pragma unroll
    for (int v = 0; v < BITER; ++v) {
      int g = tid + NTHR * v;
      if (g < BGROUPS) {
        int c = g / (BK / 4), kk = (g % (BK / 4)) * 4;
        b_values<4>(p, k0 + kk, kend, c, rb[v]);
      }
    }
  };
  auto store_B = [&](int buf) {
#pragma unroll
    for (int v = 0; v < BITER; ++v) {
      int g = tid + NTHR * v;
      if (g < BGROUPS) {
        int c = g / (BK / 4), kk = (g % (BK / 4)) * 4;
#pragma unroll
        for (int u = 0; u < 4; ++u) Bs[buf][kk + u][c] = rb[v][u];
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
    for (int ks = 0; ks < BK / 4; ++ks) {
      int kk = ks * 4 + (lane >> 4);
      float a0 = As[buf][kk][wave * 32 + (lane & 15)];
      float a1 = As[buf][kk][wave * 32 + 16 + (lane & 15)];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        float b = Bs[buf][kk][16 * t + (lane & 15)];
        acc[0][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b, acc[0][t], 0, 0, 0);
        acc[1][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b, acc[1][t], 0, 0, 0);
      }
    }
    if (more) {
      store_A(buf ^ 1);
      store_B(buf ^ 1);
    }
    __syncthreads();
  }
  epilogue<NT, TRANS>(p, acc, l0, split);
}

// ---------------------------------------------------------------------------
// 16-bit operands: v_mfma_f32_16x16x32_{bf16,f16}.  LDS: As[row][kk],
// Bs[col][kk], 8 consecutive kk per lane read as one 16-byte word.
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

template <typename TA, int NT, int TRANS, bool VEC>
__global__ __launch_bounds__(NTHR) void gemm_h_kernel(GemmParams p) {
  constexpr int BK = 64, SA = BK + 8, KP = 16 * NT;
  constexpr int BGROUPS = (BK / 8) * KP;  // groups of 8 consecutive kk of one column
  constexpr int BITER = (BGROUPS + NTHR - 1) / NTHR;
  __shared__ __attribute__((aligned(16))) uint16_t As[2][BL][SA];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][KP][SA];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long l0 = (long)blockIdx.x * BL;
  const int split = blockIdx.y;
  const long kbeg = (long)split * p.kchunk;
  const long kend = min(p.K, kbeg + p.kchunk);
  const int nst = (int)((kend - kbeg + BK - 1) / BK);
  const uint16_t *A = (const uint16_t *)p.A;

  uint4 ra[4];
  uint4 rb[BITER];

  auto load_A = [&](long k0) {
    if (!TRANS) {
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
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int q = tid + NTHR * v;
        int j = q >> 3, kk = (q & 7) * 8;
        long gj = l0 + j, gk = k0 + kk;
        if (VEC && gj < p.L && gk + 7 < kend) {
          ra[v] = *(const uint4 *)(A + gk + p.lda * gj);
        } else {
          uint16_t t[8];
#pragma unroll
          for (int e = 0; e < 8; ++e) t[e] = (gj < p.L && gk + e < kend) ? A[gk + e + p.lda * gj] : 0;
          ra[v] = __builtin_bit_cast(uint4, t);
        }
      }
    }
  };
  auto store_A = [&](int buf) {
    if (!TRANS) {
      int rg = tid & 15, kg = tid >> 4;
      uint16_t t[4][8];
#pragma unroll
      for (int u = 0; u < 4; ++u) __builtin_memcpy(t[u], &ra[u], 16);
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        uint16_t w[4] = {t[0][r], t[1][r], t[2][r], t[3][r]};
        *(uint2 *)&As[buf][8 * rg + r][4 * kg] = __builtin_bit_cast(uint2, w);
      }
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int q = tid + NTHR * v;
        int j = q >> 3, kk = (q & 7) * 8;
        *(uint4 *)&As[buf][j][kk] = ra[v];
      }
    }
  };
  auto load_B = [&](long k0) {
#pragma unroll
    for (int v = 0; v < BITER; ++v) {
      int g = tid + NTHR * v;
      if (g < BGROUPS) {
        int c = g / (BK / 8), kk = (g % (BK / 8)) * 8;
        float x[8];
        b_values<8>(p, k0 + kk, kend, c, x);
        uint16_t h[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) h[e] = to16<TA>(x[e]);
        rb[v] = __builtin_bit_cast(uint4, h);
      }
    }
  };
  auto store_B = [&](int buf) {
#pragma unroll
    for (int v = 0; v < BITER; ++v) {
      int g = tid + NTHR * v;
      if (g < BGROUPS) {
        int c = g / (BK / 8), kk = (g % (BK / 8)) * 8;
        *(uint4 *)&Bs[buf][c][kk] = rb[v];
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
      for (int t = 0; t < NT; ++t) {
        uint4 b = *(const uint4 *)&Bs[buf][16 * t + (lane & 15)][kk];
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
  epilogue<NT, TRANS>(p, acc, l0, split);
}

// Sum of split-K partials (fixed order) into the final layout.
__global__ void reduce_splits_kernel(const float *partial, int splits, long L, int KP, int k,
                                     float *C, long ldc, int cmajor, const float *colscale) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= L * k) return;
  long gl = idx / k;
  int c = (int)(idx % k);
  float s = 0.0f;
  for (int sp = 0; sp < splits; ++sp) s += partial[((long)sp * L + gl) * KP + c];
  if (colscale) s *= colscale[gl];
  if (cmajor)
    C[c + ldc * gl] = s;
  else
    C[gl + ldc * c] = s;
}

template <int NT, int TRANS>
void launch_nt(int dtype, bool vec, dim3 grid, hipStream_t st, const GemmParams &p) {
  if (dtype == PLANC_F32) {
    if (vec)
      gemm_f32_kernel<NT, TRANS, true><<<grid, NTHR, 0, st>>>(p);
    else
      gemm_f32_kernel<NT, TRANS, false><<<grid, NTHR, 0, st>>>(p);
  } else if (dtype == PLANC_BF16) {
    if (vec)
      gemm_h_kernel<bf16_t, NT, TRANS, true><<<grid, NTHR, 0, st>>>(p);
    else
      gemm_h_kernel<bf16_t, NT, TRANS, false><<<grid, NTHR, 0, st>>>(p);
  } else {
    if (vec)
      gemm_h_kernel<f16_t, NT, TRANS, true><<<grid, NTHR, 0, st>>>(p);
    else
      gemm_h_kernel<f16_t, NT, TRANS, false><<<grid, NTHR, 0, st>>>(p);
  }
}

template <int TRANS>
void launch(int NT, int dtype, bool vec, dim3 grid, hipStream_t st, const GemmParams &p) {
  switch (NT) {
    case 1:
      launch_nt<1, TRANS>(dtype, vec, grid, st, p);
      break;
    case 2:
      launch_nt<2, TRANS>(dtype, vec, grid, st, p);
      break;
    case 3:
      launch_nt<3, TRANS>(dtype, vec, grid, st, p);
      break;
    default:
      launch_nt<4, TRANS>(dtype, vec, grid, st, p);
      break;
  }
}

}  // namespace

int thin_gemm(planc_ctx *ctx, int dtype, int trans, const void *A, long lda, long L, long K,
              const float *B, long ldb, const KrpDesc *krp, const float *rowscale,
              const float *colscale, int k, float *C, long ldc, int cmajor, int splits_hint) {
  if (L < 0 || K < 0 || k <= 0 || !A || !C) return set_error(ctx, PLANC_EINVAL, "thin_gemm: bad shape/pointer");
  if (dtype != PLANC_F32 && dtype != PLANC_BF16 && dtype != PLANC_F16)
    return set_error(ctx, PLANC_EINVAL, "thin_gemm: bad dtype");
  if (L == 0) return PLANC_OK;
  const int esz = dtype == PLANC_F32 ? 4 : 2;
  const int vecw = 16 / esz;
  const bool vec = (lda % vecw == 0) && (((uintptr_t)A) % 16 == 0);
  const int BK = dtype == PLANC_F32 ? 32 : 64;
  for (int c0 = 0; c0 < k; c0 += 64) {
    int kc = (k - c0 < 64) ? (k - c0) : 64;
    int NT = (kc + 15) / 16;
    GemmParams p{};
    p.A = A;
    p.lda = lda;
    p.L = L;
    p.K = K;
    p.rowscale = rowscale;
    p.colscale = colscale;
    p.k = kc;
    p.krp.nf = 0;
    if (krp && krp->nf > 0) {
      if (krp->nf > 6) return set_error(ctx, PLANC_EINVAL, "thin_gemm: more than 6 KRP factors");
      long prod = 1;
      for (int q = 0; q < krp->nf; ++q) prod *= krp->dim[q];
      if (prod != K) return set_error(ctx, PLANC_EINVAL, "thin_gemm: KRP rows != K");
      p.krp.nf = krp->nf;
      for (int q = 0; q < krp->nf; ++q) {
        p.krp.f[q] = krp->f[q] + krp->ld[q] * c0;
        p.krp.dim[q] = (int)krp->dim[q];
        p.krp.ld[q] = krp->ld[q];
      }
    } else {
      if (!B) return set_error(ctx, PLANC_EINVAL, "thin_gemm: null B");
      p.B = B + ldb * c0;
      p.ldb = ldb;
    }
    p.C = cmajor ? C + c0 : C + ldc * c0;
    p.ldc = ldc;
    p.cmajor = cmajor;
    long lblocks = (L + BL - 1) / BL;
    int splits = 1;
    if (K > 0) {
      long target = 4L * ctx->num_cu;
      long want = (target + lblocks - 1) / lblocks;
      long maxs = (K + 4L * BK - 1) / (4L * BK);  // each split >= 4 chunks
      if (want > maxs) want = maxs;
      if (want < 1) want = 1;
      if (splits_hint > 0) want = splits_hint;
      if (want > 4096) want = 4096;
      splits = (int)want;
    }
    long kchunk = (K + splits - 1) / splits;
    kchunk = ((kchunk + BK - 1) / BK) * BK;
    if (kchunk == 0) kchunk = BK;
    splits = (int)((K + kchunk - 1) / kchunk);
    if (splits < 1) splits = 1;
    p.splits = splits;
    p.kchunk = kchunk;
    int KP = 16 * NT;
    if (splits > 1) {
      p.partial = (float *)ws_get(ctx, 0, sizeof(float) * (size_t)splits * L * KP);
      if (!p.partial) return set_error(ctx, PLANC_ENOMEM, "thin_gemm: workspace");
    }
    dim3 grid((unsigned)lblocks, (unsigned)splits);
    if (trans)
      launch<1>(NT, dtype, vec, grid, ctx->stream, p);
    else
      launch<0>(NT, dtype, vec, grid, ctx->stream, p);
    PLANC_LAUNCH_CHECK(ctx);
    if (splits > 1) {
      long tot = L * kc;
      unsigned nb = (unsigned)((tot + 255) / 256);
      reduce_splits_kernel<<<nb, 256, 0, ctx->stream>>>(p.partial, splits, L, KP, kc, p.C, ldc,
                                                        cmajor, trans ? colscale : nullptr);
      PLANC_LAUNCH_CHECK(ctx);
    }
  }
  return PLANC_OK;
}

}  // namespace planc
