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
// mixed-radix index of those modes (lowest fastest).
__device__ inline float weight(const ContractParams &p, long idx, int t0, int t1, int c) {
  float w = 1.0f;
  for (int t = t0; t < t1; ++t) {
    long dg = idx % p.d[t];
    idx /= p.d[t];
    w *= p.F[t][dg + p.ldf[t] * c];
  }
  return w;
}

// Small inner extent: lanes walk consecutive output indices i, the four
// wavefronts split the outer index b, summed in wave order.
__global__ __launch_bounds__(NTHR) void contract_small_inner_kernel(ContractParams p) {
  __shared__ float red[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = blockIdx.y;
  const long i = (long)blockIdx.x * 64 + lane;
  float acc = 0.0f;
  if (i < p.In) {
    for (long b = wave; b < p.outer; b += 4) {
      float wo = weight(p, b, p.tn + 1, p.nb, c);
      float s = 0.0f;
      for (long a = 0; a < p.inner; ++a) {
        float wi = weight(p, a, 0, p.tn, c);
        s = fmaf(p.P[(a + p.inner * (i + p.In * b)) + p.ldp * c], wi, s);
      }
      acc = fmaf(s, wo, acc);
    }
  }
  red[wave][lane] = acc;
  __syncthreads();
  if (wave == 0 && i < p.In) {
    float t = red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane];
    p.M[c + p.ldo * i] = t;
  }
}

// Large inner extent: one block per (i, c), threads over (a, b) with a
// fastest (contiguous reads), deterministic block reduction.
__global__ __launch_bounds__(NTHR) void contract_large_inner_kernel(ContractParams p) {
  __shared__ double red[NTHR / 64];
  const long i = blockIdx.x;
  const int c = blockIdx.y;
  const long tot = p.inner * p.outer;
  float acc = 0.0f;
  for (long f = threadIdx.x; f < tot; f += NTHR) {
    long a = f % p.inner, b = f / p.inner;
    float w = weight(p, a, 0, p.tn, c) * weight(p, b, p.tn + 1, p.nb, c);
    acc = fmaf(p.P[(a + p.inner * (i + p.In * b)) + p.ldp * c], w, acc);
  }
  double v = wave_sum_d((double)acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < NTHR / 64; ++w) s += red[w];
    p.M[c + p.ldo * i] = (float)s;
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
  if (p.inner < 32) {
    dim3 grid((unsigned)((p.In + 63) / 64), (unsigned)k);
    contract_small_inner_kernel<<<grid, NTHR, 0, ctx->stream>>>(p);
  } else {
    dim3 grid((unsigned)p.In, (unsigned)k);
    contract_large_inner_kernel<<<grid, NTHR, 0, ctx->stream>>>(p);
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
