// Nesterov-accelerated NTF (NTFNES, ntf/ntfnes.hpp:9-188).
//
// Mode update (NTFNES::update, :119-166): with L and mu the extreme
// eigenvalues of S (the Hadamard product of the other Grams), lambda from
// get_lambda (:27-38) and q = (mu + lambda)/(L + lambda), a projected
// Nesterov iteration on the proximal problem
//   min 0.5 tr(Y^T (S + lambda I) Y) - tr((M + lambda H*)^T Y),  Y >= 0,
// H* the mode's proximal centre, from Y = H*; it stops when
// max |grad o Y| <= 1e-2 and min grad >= -1e-2 (stop_iter, :56-66).  One
// workgroup runs the whole loop (block reductions for the stop test), the
// factor and the iterates in FP64 in global scratch; the loop is bounded
// (the reference's is not) and a run past the bound sets a flag.
//
// The extreme eigenvalues come from a Householder tridiagonalisation and
// Sturm-sequence bisection on the host (S is k x k, k <= 64).
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"

namespace planc {

namespace {

constexpr int NES_THR = 1024;

__device__ inline double nes_block_reduce(double v, double *red, bool is_max) {
  // max (is_max) or min over the block; red holds NES_THR / 64 doubles.
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(v, off, 64);
    v = is_max ? fmax(v, o) : fmin(v, o);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double r = red[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = is_max ? fmax(r, red[w]) : fmin(r, red[w]);
  return r;
}

// work: Ht, Htprev, acc, mod (each I x k doubles).  M(c, r) at
// M[r * mrs + c * mcs].  prox (I x k FP32) is the centre H*; F receives the
// new factor; status[0] |= 8 when the loop bound is reached.
__global__ __launch_bounds__(NES_THR) void nes_loop_kernel(const double *S, int k, double lambda, double Lval,
                                                           double q, const float *M, long mrs, long mcs,
                                                           const float *prox, long I, double *work, float *F,
                                                           long ldf, long cap, int *status) {
  __shared__ double G[64 * 64];
  __shared__ double red[NES_THR / 64];
  __shared__ double coef[2];
  const long tot = I * k;
  double *Ht = work, *Hp = work + tot, *acc = work + 2 * tot, *mod = work + 3 * tot;
  for (int e = threadIdx.x; e < k * k; e += blockDim.x) G[e] = S[e] + ((e % k) == (e / k) ? lambda : 0.0);
  for (long e = threadIdx.x; e < tot; e += blockDim.x) {
    const long r = e % I;
    const int c = (int)(e / I);
    const double h = (double)prox[e];
    Ht[e] = h;
    acc[e] = h;
    mod[e] = -lambda * h - (double)M[r * mrs + (long)c * mcs];
  }
  __syncthreads();
  double alpha = 1.0;
  const double step = 1.0 / (Lval + lambda);
  for (long it = 0;; ++it) {
    // grad = mod + (S + lambda I) Y, the stop test on grad o Y and min grad.
    double amax = 0.0, gmin = 1e300;
    for (long e = threadIdx.x; e < tot; e += blockDim.x) {
      const long r = e % I;
      const int c = (int)(e / I);
      double g = mod[e];
      for (int t = 0; t < k; ++t) g += G[c + k * t] * acc[r + I * t];
      amax = fmax(amax, fabs(g * acc[e]));
      gmin = fmin(gmin, g);
      Hp[e] = g;  // grad held in the Htprev slot until the step below
    }
    amax = nes_block_reduce(amax, red, true);
    gmin = nes_block_reduce(gmin, red, false);
    if (amax <= 1e-2 && gmin >= -1e-2) break;
    if (it >= cap) {
      if (threadIdx.x == 0) atomicOr(status, 8);
      break;
    }
    if (threadIdx.x == 0) {
      // get_alpha (:40-54) and beta (:160-161).
      const double ap = alpha;
      const double b = ap * ap - q, c2 = -ap * ap;
      const double an = (-b + sqrt(b * b - 4 * c2)) / 2;
      coef[0] = an;
      coef[1] = (ap * (1 - ap)) / (an + ap * ap);
    }
    __syncthreads();
    alpha = coef[0];
    const double beta = coef[1];
    for (long e = threadIdx.x; e < tot; e += blockDim.x) {
      const double g = Hp[e];
      const double hold = Ht[e];
      double h = acc[e] - step * g;
      h = h > 0.0 ? h : 0.0;
      Ht[e] = h;
      acc[e] = h + beta * (h - hold);
    }
    __syncthreads();
  }
  for (long e = threadIdx.x; e < tot; e += blockDim.x) F[(e % I) + ldf * (e / I)] = (float)Ht[e];
}

// prox = Ht with unit columns (normalize_rows of the transposed factor),
// lambda_out = the norms; columns of zero norm stay zero.
__global__ void nes_prox_kernel(const double *Ht, long I, int k, float *prox, double *lambda_out) {
  __shared__ double red[4];
  const int c = blockIdx.x;
  double s = 0;
  for (long r = threadIdx.x; r < I; r += blockDim.x) s += Ht[r + I * c] * Ht[r + I * c];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  const double nrm = sqrt(red[0] + red[1] + red[2] + red[3]);
  if (threadIdx.x == 0) lambda_out[c] = nrm;
  for (long r = threadIdx.x; r < I; r += blockDim.x)
    prox[r + I * c] = (float)(nrm > 0 ? Ht[r + I * c] / nrm : Ht[r + I * c]);
}

// A = (sc F - sp P) step + sc F with sc, sp per column (1 unless scaled).
__global__ void nes_acc_kernel(const float *F, const float *Pv, long I, int k, const double *sc,
                               const double *sp, double step, float *A) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= I * k) return;
  const int c = (int)(e / I);
  const double a = sc ? sc[c] : 1.0, b = sp ? sp[c] : 1.0;
  const double cur = a * (double)F[e];
  A[e] = (float)((cur - b * (double)Pv[e]) * step + cur);
}

__global__ void scale_cols_by_kernel(const float *A, long I, int k, const double *s, float *out) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= I * k) return;
  out[e] = (float)((double)A[e] * s[e / I]);
}

// Householder reduction of a symmetric matrix to tridiagonal form (d, e).
void tridiagonalise(std::vector<double> a, int n, std::vector<double> &d, std::vector<double> &e) {
  d.assign(n, 0.0);
  e.assign(n, 0.0);
  for (int j = 0; j < n - 2; ++j) {
    double alpha = 0;
    for (int i = j + 1; i < n; ++i) alpha += a[i + n * j] * a[i + n * j];
    alpha = sqrt(alpha);
    if (alpha == 0) continue;
    if (a[j + 1 + n * j] > 0) alpha = -alpha;
    std::vector<double> v(n, 0.0);
    v[j + 1] = a[j + 1 + n * j] - alpha;
    for (int i = j + 2; i < n; ++i) v[i] = a[i + n * j];
    double vn = 0;
    for (int i = j + 1; i < n; ++i) vn += v[i] * v[i];
    if (vn == 0) continue;
    // A <- H A H with H = I - 2 v v^T / (v^T v).
    std::vector<double> p(n, 0.0);
    for (int i = 0; i < n; ++i) {
      double s = 0;
      for (int t = j + 1; t < n; ++t) s += a[i + n * t] * v[t];
      p[i] = 2 * s / vn;
    }
    double kf = 0;
    for (int i = j + 1; i < n; ++i) kf += v[i] * p[i];
    kf /= vn;
    for (int i = 0; i < n; ++i) p[i] -= kf * v[i];
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < n; ++c) a[r + n * c] -= v[r] * p[c] + p[r] * v[c];
  }
  for (int i = 0; i < n; ++i) d[i] = a[i + n * i];
  for (int i = 0; i + 1 < n; ++i) e[i] = a[i + 1 + n * i];
}

// Number of eigenvalues of the tridiagonal (d, e) below x (Sturm count).
int sturm_count(const std::vector<double> &d, const std::vector<double> &e, double x) {
  int cnt = 0;
  double q = 1;
  const int n = (int)d.size();
  for (int i = 0; i < n; ++i) {
    const double off = i > 0 ? e[i - 1] * e[i - 1] : 0.0;
    q = d[i] - x - (i > 0 ? off / q : 0.0);
    if (q == 0) q = -1e-300;
    if (q < 0) ++cnt;
  }
  return cnt;
}

}  // namespace

// Smallest and largest eigenvalue of the symmetric k x k matrix S (host).
void sym_eig_extremes(const std::vector<double> &S, int k, double *mn, double *mx) {
  std::vector<double> d, e;
  tridiagonalise(S, k, d, e);
  double lo = d[0], hi = d[0];
  for (int i = 0; i < k; ++i) {
    const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < k ? fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  auto kth = [&](int idx) {  // eigenvalue idx (0-based, ascending)
    double a = lo, b = hi;
    for (int it = 0; it < 200 && b - a > 1e-16 * std::max(fabs(a), fabs(b)); ++it) {
      const double m = 0.5 * (a + b);
      if (m == a || m == b) break;
      if (sturm_count(d, e, m) > idx)
        b = m;
      else
        a = m;
    }
    return 0.5 * (a + b);
  };
  *mn = kth(0);
  *mx = kth(k - 1);
}

// One NTFNES mode update on the device: S (k x k FP64 device), M (k x I,
// element (c, r) at M[r*mrs + c*mcs]), centre prox (I x k FP32, updated to
// the new prox), new factor into F; prox_l (device, k) gets the prox norms.
int nes_mode_update(planc_ctx *ctx, const double *dS, int k, const float *M, long mrs, long mcs, float *prox,
                    long I, float *F, long ldf, double *prox_l, int *status) {
  if (k <= 0 || k > 64) return set_error(ctx, PLANC_EUNSUPPORTED, "NES: k must be 1..64");
  std::vector<double> S((size_t)k * k);
  PLANC_HIP_CHECK(ctx, hipMemcpyAsync(S.data(), dS, sizeof(double) * k * k, hipMemcpyDeviceToHost, ctx->stream));
  PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  double mu, L;
  sym_eig_extremes(S, k, &mu, &L);
  const double q0 = L / mu;
  const double lambda = q0 > 1e6 ? 10 * mu : (q0 > 1e3 ? mu : mu / 10);  // get_lambda (:27-38)
  const double q = (mu + lambda) / (L + lambda);
  double *work = (double *)ws_get(ctx, 35, sizeof(double) * 4 * (size_t)I * k);
  if (!work) return set_error(ctx, PLANC_ENOMEM, "NES: workspace");
  nes_loop_kernel<<<1, NES_THR, 0, ctx->stream>>>(dS, k, lambda, L, q, M, mrs, mcs, prox, I, work, F, ldf, 100000,
                                                  status);
  PLANC_LAUNCH_CHECK(ctx);
  nes_prox_kernel<<<k, 256, 0, ctx->stream>>>(work, I, k, prox, prox_l);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

// A (I x k) = (sc F - sp P) step + sc F.
int nes_acc_build(planc_ctx *ctx, const float *F, const float *Pv, long I, int k, const double *sc,
                  const double *sp, double step, float *A) {
  const long tot = I * k;
  if (tot == 0) return PLANC_OK;
  nes_acc_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, ctx->stream>>>(F, Pv, I, k, sc, sp, step, A);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int scale_cols_by(planc_ctx *ctx, const float *A, long I, int k, const double *s, float *out) {
  const long tot = I * k;
  if (tot == 0) return PLANC_OK;
  scale_cols_by_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, ctx->stream>>>(A, I, k, s, out);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

}  // namespace planc
