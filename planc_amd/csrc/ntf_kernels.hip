// NTF kernels besides the bulk contractions: projection contraction
// (multi-TTV of the dimension tree), Hadamard of Grams, and the FP64 terms
// of the CP relative error.
#include "common.h"

namespace planc {

namespace {

constexpr int NTHR = 256;

__device__ inline double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

struct ContractParams {
  const float *P;  // E x k, column-major (ld ldp); E = prod d[0..nb)
  long ldp;
  int nb;          // number of modes in the block
  int tn;          // position of the target mode inside the block
  int d[8];
  const float *F[8];
  long ldf[8];
  int k;
  float *M;        // k x d[tn], column-major (ld ldo)
  long ldo;
  long inner, In, outer;
};

// Product of F_t(digit_t(idx), c) over block positions [t0, t1) with idx the
// mixed-radix index of those modes (lowest fastest).  idx < 2^32.
__device__ inline float weight(const ContractParams &p, unsigned idx, int t0, int t1, int c) {
  float w = 1.0f;
  for (int t = t0; t < t1; ++t) {
    unsigned nx = idx / (unsigned)p.d[t];
    unsigned dg = idx - nx * (unsigned)p.d[t];
    idx = nx;
    w *= p.F[t][dg + p.ldf[t] * c];
  }
  return w;
}

constexpr int WMAX = 4096;  // weights held in LDS per block

// inner == 1 (target mode first in the block): lanes walk consecutive output
// indices i (coalesced), the 16 wavefronts split the outer index b
// (b = wave, wave + 16, ...), summed in wave order.
__global__ __launch_bounds__(1024) void contract_outer_kernel(ContractParams p) {
  __shared__ float wo[WMAX];
  __shared__ float red[16][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = blockIdx.y;
  const long i = (long)blockIdx.x * 64 + lane;
  const bool cached = p.outer <= WMAX;
  if (cached)
    for (long b = threadIdx.x; b < p.outer; b += blockDim.x) wo[b] = weight(p, (unsigned)b, p.tn + 1, p.nb, c);
  __syncthreads();
  float acc = 0.0f;
  if (i < p.In) {
    const float *src = p.P + p.ldp * c + i;
    const long st = p.In * 16;
    long b = wave;
    if (cached) {
      for (; b + 16 * 7 < p.outer; b += 16 * 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = src[p.In * b + st * u];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = fmaf(v[u], wo[b + 16 * u], acc);
      }
      for (; b < p.outer; b += 16) acc = fmaf(src[p.In * b], wo[b], acc);
    } else {
      for (; b < p.outer; b += 16) acc = fmaf(src[p.In * b], weight(p, (unsigned)b, p.tn + 1, p.nb, c), acc);
    }
  }
  red[wave][lane] = acc;
  __syncthreads();
  if (wave == 0 && i < p.In) {
    float t = 0.0f;
#pragma unroll
    for (int w = 0; w < 16; ++w) t += red[w][lane];
    p.M[c + p.ldo * i] = t;
  }
}

// inner > 1: one block per (tile of TI output indices, c); threads walk the
// contiguous inner index a, loop over b; each output is a deterministic
// block reduction.
constexpr int TI = 8;
__global__ __launch_bounds__(256) void contract_inner_kernel(ContractParams p) {
  __shared__ float wi[WMAX];
  __shared__ float wo[WMAX];
  __shared__ double red[TI][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = blockIdx.y;
  const long i0 = (long)blockIdx.x * TI;
  const bool ci = p.inner <= WMAX, co = p.outer <= WMAX;
  if (ci)
    for (long a = threadIdx.x; a < p.inner; a += blockDim.x) wi[a] = weight(p, (unsigned)a, 0, p.tn, c);
  if (co)
    for (long b = threadIdx.x; b < p.outer; b += blockDim.x) wo[b] = weight(p, (unsigned)b, p.tn + 1, p.nb, c);
  __syncthreads();
  float acc[TI];
#pragma unroll
  for (int t = 0; t < TI; ++t) acc[t] = 0.0f;
  const float *Pc = p.P + p.ldp * c;
  for (long b = 0; b < p.outer; ++b) {
    const float wb = co ? wo[b] : weight(p, (unsigned)b, p.tn + 1, p.nb, c);
    for (long a = threadIdx.x; a < p.inner; a += blockDim.x) {
      const float w = (ci ? wi[a] : weight(p, (unsigned)a, 0, p.tn, c)) * wb;
#pragma unroll
      for (int t = 0; t < TI; ++t) {
        const long i = i0 + t;
        if (i < p.In) acc[t] = fmaf(Pc[a + p.inner * (i + p.In * b)], w, acc[t]);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < TI; ++t) {
    double v = wave_sum_d((double)acc[t]);
    if (lane == 0) red[t][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < TI) {
    const long i = i0 + threadIdx.x;
    if (i < p.In) {
      double s = red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] + red[threadIdx.x][3];
      p.M[c + p.ldo * i] = (float)s;
    }
  }
}

struct GramSet {
  const double *g[8];
  int n;
};

__global__ void hadamard_kernel(GramSet gs, int skip, int kk, double *out) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kk) return;
  double v = 1.0;
  for (int i = 0; i < gs.n; ++i)
    if (i != skip) v *= gs.g[i][e];
  out[e] = v;
}

// sum_c lambda_c sum_i F(i, c) * M(c, i), two stages.
__global__ __launch_bounds__(NTHR) void lambda_dot_partial_kernel(const float *F, long I, long ldf,
                                                                  const float *M, long ldm, int k,
                                                                  const double *lambda,
                                                                  double *part) {
  __shared__ double red[NTHR / 64];
  double s = 0.0;
  long tot = I * k;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    long i = e % I;
    int c = (int)(e / I);
    s += lambda[c] * (double)F[i + ldf * c] * (double)M[c + ldm * i];
  }
  s = wave_sum_d(s);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < NTHR / 64; ++w) t += red[w];
    part[blockIdx.x] = t;
  }
}

// rel = sqrt(max(||X||^2 - 2 <X, Xhat> + ||Xhat||^2, 0) / ||X||^2) with
// ||Xhat||^2 = lambda^T (Hadamard of all Grams) lambda.
__global__ void ntf_relerr_kernel(const double *normX2, const double *part, long nparts, GramSet gs,
                                  int k, const double *lambda, double *rel_out) {
  if (threadIdx.x >= 64) return;
  double t = 0.0;
  for (long b = threadIdx.x; b < nparts; b += 64) t += part[b];
  t = wave_sum_d(t);
  double xh = 0.0;
  for (int e = threadIdx.x; e < k * k; e += 64) {
    int a = e % k, b = e / k;
    double v = lambda[a] * lambda[b];
    for (int i = 0; i < gs.n; ++i) v *= gs.g[i][e];
    xh += v;
  }
  xh = wave_sum_d(xh);
  if (threadIdx.x == 0) {
    double e2 = *normX2 - 2.0 * t + xh;
    if (e2 < 0) e2 = 0;
    *rel_out = sqrt(e2 / *normX2);
  }
}

// lambda <- lambda .* norms (NCPFactors::normalize(), common/ncpfactors.hpp:329-339).
__global__ void lambda_mul_kernel(double *lambda, const double *norms, int k, int init) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k) return;
  lambda[c] = (init ? 1.0 : lambda[c]) * norms[c];
}

// out (k x I, ld ldo) = in^T for in I x k (ld ldi).
__global__ void transpose_kernel(const float *in, long I, long ldi, int k, float *out, long ldo) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= I * k) return;
  int c = (int)(e % k);
  long i = e / k;
  out[c + ldo * i] = in[i + ldi * c];
}

// ---------------------------------------------------------------------------
// One mode update of AUNTF for a factor that fits one block's LDS
// (k * (I + 1) <= MODE_LDS_FLOATS): S = Hadamard of the other Grams, the
// HALS (ntf/ntfhals.hpp:13-29) or MU row update from this mode's MTTKRP,
// normalize(mode) (ntf/auntf.hpp:66-79: lambda = column 2-norms, columns
// divided by them), the FP64 Gram of the new factor, and optionally the CP
// relative error of ntf/auntf.hpp:224-245 from this mode's MTTKRP.  One
// launch replaces six.  kind: 0 MU, 1 HALS, 2 no row update (the factor was
// solved already, e.g. by BPP).
struct ModeParams {
  float *F;
  long I, ldf;
  int k;
  const float *M;  // MTTKRP: element (x, i) at M[i * mrs + x * mcs]
  long mrs, mcs;
  const double *g[8];
  int ng, skip;
  double *gout, *lambda;
  int kind;
  float eps;
  const double *normX2;
  double *rel_out;
};

constexpr int MODE_LDS_FLOATS = 24576;

// Column stride of the factor tile in LDS: a multiple of 4 (16-byte column
// reads in the Gram) plus 4 (columns start in different banks).
__host__ __device__ inline long mode_ld(long I) { return (I + 3) / 4 * 4 + 4; }
constexpr int MODE_THREADS = 512;

__device__ inline double block_sum_mode(double v, double *red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  v = wave_sum_d(v);
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < MODE_THREADS / 64; ++w) s += red[w];
  return s;
}

template <int KM>
__global__ __launch_bounds__(MODE_THREADS) void mode_update_kernel(ModeParams p) {
  __shared__ __attribute__((aligned(16))) float Sf[KM][KM];  // Sf[x][t] = S(t, x)
  __shared__ __attribute__((aligned(16))) float Fs[MODE_LDS_FLOATS];  // column c at Fs[c * mode_ld(I) + i]
  __shared__ double nrm[KM];
  __shared__ double red[MODE_THREADS / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int k = p.k;
  const long I = p.I, ldl = mode_ld(I);
  for (int e = tid; e < KM * KM; e += MODE_THREADS) {
    const int a = e % KM, b = e / KM;
    float v = 0.0f;
    if (a < k && b < k) {
      double d = 1.0;
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (q < p.ng && q != p.skip) d *= p.g[q][a + (long)k * b];
      v = (float)d;
    }
    Sf[b][a] = v;
  }
  __syncthreads();
  for (long i = tid; i < I; i += MODE_THREADS) {
    // The row of F and this row's MTTKRP values, loaded up front (independent
    // loads in flight together).
    float h[KM], mv[KM];
#pragma unroll
    for (int t = 0; t < KM; ++t) {
      h[t] = (t < k) ? p.F[i + p.ldf * t] : 0.0f;
      mv[t] = (t < k) ? p.M[i * p.mrs + t * p.mcs] : 0.0f;
    }
    // The column loops are not unrolled; h[] is only indexed by unrolled t
    // loops (selects), so it stays in registers.
    if (p.kind == 1) {
      const float floor_v = 1e-17f;
#pragma unroll 1
      for (int x = 0; x < k; ++x) {
        float s = 0.0f, hx = 0.0f, mx = 0.0f;
        const floatx4 *sx = (const floatx4 *)&Sf[x][0];
#pragma unroll
        for (int q = 0; q < KM / 4; ++q) {
          const floatx4 g = sx[q];
#pragma unroll
          for (int u = 0; u < 4; ++u) s = fmaf(h[4 * q + u], g[u], s);
        }
#pragma unroll
        for (int t = 0; t < KM; ++t) {
          hx = (t == x) ? h[t] : hx;
          mx = (t == x) ? mv[t] : mx;
        }
        float v = hx + (mx - s);
        v = (v < floor_v) ? floor_v : v;
#pragma unroll
        for (int t = 0; t < KM; ++t) h[t] = (t == x) ? v : h[t];
      }
#pragma unroll
      for (int t = 0; t < KM; ++t)
        if (t < k) Fs[t * ldl + i] = h[t];
    } else if (p.kind == 0) {
#pragma unroll 1
      for (int c = 0; c < k; ++c) {
        float s = 0.0f, hc = 0.0f, mc = 0.0f;
        const floatx4 *sc = (const floatx4 *)&Sf[c][0];
#pragma unroll
        for (int q = 0; q < KM / 4; ++q) {
          const floatx4 g = sc[q];
#pragma unroll
          for (int u = 0; u < 4; ++u) s = fmaf(h[4 * q + u], g[u], s);
        }
#pragma unroll
        for (int t = 0; t < KM; ++t) {
          hc = (t == c) ? h[t] : hc;
          mc = (t == c) ? mv[t] : mc;
        }
        Fs[c * ldl + i] = (hc * mc) / (s + p.eps);
      }
    } else {
#pragma unroll
      for (int t = 0; t < KM; ++t)
        if (t < k) Fs[t * ldl + i] = h[t];
    }
  }
  __syncthreads();
  for (int c = wave; c < k; c += MODE_THREADS / 64) {
    double s = 0.0;
    for (long i = lane; i < I; i += 64) {
      const double v = Fs[c * ldl + i];
      s += v * v;
    }
    s = wave_sum_d(s);
    if (lane == 0) {
      const double nv = sqrt(s);
      nrm[c] = nv;
      p.lambda[c] = nv;
    }
  }
  __syncthreads();
  for (long e = tid; e < I * k; e += MODE_THREADS) {
    const int c = (int)(e / I);
    const long i = e - (long)c * I;
    float v = Fs[c * ldl + i];
    const double nv = nrm[c];
    if (nv > 0) v = (float)((double)v / nv);
    Fs[c * ldl + i] = v;
    p.F[i + p.ldf * c] = v;
  }
  __syncthreads();
  for (int e = tid; e < k * k; e += MODE_THREADS) {
    const int a = e % k, b = e / k;
    if (a > b) continue;
    double s = 0.0;
    const float *fa = Fs + a * ldl, *fb = Fs + b * ldl;
    long i = 0;
    for (; i + 4 <= I; i += 4) {
      const floatx4 va = *(const floatx4 *)(fa + i), vb = *(const floatx4 *)(fb + i);
      s += (double)va[0] * (double)vb[0];
      s += (double)va[1] * (double)vb[1];
      s += (double)va[2] * (double)vb[2];
      s += (double)va[3] * (double)vb[3];
    }
    for (; i < I; ++i) s += (double)fa[i] * (double)fb[i];
    p.gout[a + (long)k * b] = s;
    p.gout[b + (long)k * a] = s;
  }
  if (!p.rel_out) return;
  __syncthreads();
  double t1 = 0.0;
  for (long e = tid; e < I * k; e += MODE_THREADS) {
    const int c = (int)(e / I);
    const long i = e - (long)c * I;
    t1 += nrm[c] * (double)Fs[c * ldl + i] * (double)p.M[i * p.mrs + c * p.mcs];
  }
  double xh = 0.0;
  for (int e = tid; e < k * k; e += MODE_THREADS) {
    const int a = e % k, b = e / k;
    double v = nrm[a] * nrm[b];
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (q < p.ng) v *= (q == p.skip) ? p.gout[e] : p.g[q][e];
    xh += v;
  }
  t1 = block_sum_mode(t1, red);
  xh = block_sum_mode(xh, red);
  if (tid == 0) {
    double e2 = *p.normX2 - 2.0 * t1 + xh;
    if (e2 < 0) e2 = 0;
    *p.rel_out = sqrt(e2 / *p.normX2);
  }
}

}  // namespace

// M (k x I_n) from a projection P (E x k) over a block of modes with dims d,
// factors F, target position tn inside the block.
int contract_projection(planc_ctx *ctx, const float *P, long ldp, int nb, const long *d,
                        const float *const *F, const long *ldf, int tn, int k, float *M, long ldo) {
  if (nb < 1 || nb > 8 || tn < 0 || tn >= nb) return set_error(ctx, PLANC_EINVAL, "contract: bad block");
  ContractParams p{};
  p.P = P;
  p.ldp = ldp;
  p.nb = nb;
  p.tn = tn;
  p.k = k;
  p.M = M;
  p.ldo = ldo;
  p.inner = 1;
  p.outer = 1;
  for (int t = 0; t < nb; ++t) {
    p.d[t] = (int)d[t];
    p.F[t] = F[t];
    p.ldf[t] = ldf[t];
    if (t < tn) p.inner *= d[t];
    if (t > tn) p.outer *= d[t];
  }
  p.In = d[tn];
  if (p.In == 0) return PLANC_OK;
  long E = p.inner * p.In * p.outer;
  if (E >= (1L << 32)) return set_error(ctx, PLANC_EINVAL, "contract: block of 2^32 rows or more");
  if (p.inner == 1) {
    dim3 grid((unsigned)((p.In + 63) / 64), (unsigned)k);
    contract_outer_kernel<<<grid, 1024, 0, ctx->stream>>>(p);
  } else {
    dim3 grid((unsigned)((p.In + TI - 1) / TI), (unsigned)k);
    contract_inner_kernel<<<grid, 256, 0, ctx->stream>>>(p);
  }
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int hadamard_grams(planc_ctx *ctx, const double *const *grams, int n, int skip, int k, double *out) {
  if (n < 1 || n > 8) return set_error(ctx, PLANC_EINVAL, "hadamard: 1..8 grams");
  GramSet gs{};
  gs.n = n;
  for (int i = 0; i < n; ++i) gs.g[i] = grams[i];
  hadamard_kernel<<<(k * k + 255) / 256, 256, 0, ctx->stream>>>(gs, skip, k * k, out);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int ntf_relerr(planc_ctx *ctx, const double *normX2, const float *F, long I, long ldf, const float *M,
               long ldm, int k, const double *lambda, const double *const *grams, int n,
               double *rel_out) {
  long nb = (I * k + NTHR * 4 - 1) / (NTHR * 4);
  if (nb > 1024) nb = 1024;
  if (nb < 1) nb = 1;
  double *part = (double *)ws_get(ctx, 3, sizeof(double) * (size_t)nb);
  if (!part) return set_error(ctx, PLANC_ENOMEM, "ntf_relerr: workspace");
  lambda_dot_partial_kernel<<<(unsigned)nb, NTHR, 0, ctx->stream>>>(F, I, ldf, M, ldm, k, lambda, part);
  PLANC_LAUNCH_CHECK(ctx);
  GramSet gs{};
  gs.n = n;
  for (int i = 0; i < n; ++i) gs.g[i] = grams[i];
  ntf_relerr_kernel<<<1, 64, 0, ctx->stream>>>(normX2, part, nb, gs, k, lambda, rel_out);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int lambda_mul(planc_ctx *ctx, double *lambda, const double *norms, int k, int init) {
  lambda_mul_kernel<<<(k + 63) / 64, 64, 0, ctx->stream>>>(lambda, norms, k, init);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int transpose_to_k_major(planc_ctx *ctx, const float *in, long I, long ldi, int k, float *out, long ldo) {
  long tot = I * k;
  if (tot == 0) return PLANC_OK;
  transpose_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, ctx->stream>>>(in, I, ldi, k, out, ldo);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

}  // namespace planc

namespace planc {

bool mode_update_fits(long I, int k) { return k >= 1 && k <= 64 && (long)k * mode_ld(I) <= MODE_LDS_FLOATS; }

int mode_update(planc_ctx *ctx, int kind, float eps, float *F, long I, long ldf, int k, const float *M,
                long mrs, long mcs, const double *const *grams, int ng, int skip, double *gout,
                double *lambda, const double *normX2, double *rel_out) {
  if (!mode_update_fits(I, k) || ng < 1 || ng > 8 || skip < 0 || skip >= ng)
    return set_error(ctx, PLANC_EINVAL, "mode_update: factor does not fit one block");
  ModeParams p{};
  p.F = F;
  p.I = I;
  p.ldf = ldf;
  p.k = k;
  p.M = M;
  p.mrs = mrs;
  p.mcs = mcs;
  for (int q = 0; q < ng; ++q) p.g[q] = grams[q];
  p.ng = ng;
  p.skip = skip;
  p.gout = gout;
  p.lambda = lambda;
  p.kind = kind;
  p.eps = eps;
  p.normX2 = normX2;
  p.rel_out = rel_out;
  if (k <= 16)
    mode_update_kernel<16><<<1, MODE_THREADS, 0, ctx->stream>>>(p);
  else if (k <= 32)
    mode_update_kernel<32><<<1, MODE_THREADS, 0, ctx->stream>>>(p);
  else if (k <= 48)
    mode_update_kernel<48><<<1, MODE_THREADS, 0, ctx->stream>>>(p);
  else
    mode_update_kernel<64><<<1, MODE_THREADS, 0, ctx->stream>>>(p);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

}  // namespace planc
