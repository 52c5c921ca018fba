// AO-ADMM solves (AOADMMNMF, nmf/aoadmm.hpp; NTFAOADMM, ntf/ntfaoadmm.hpp)
// and the Gauss-Newton / conjugate-gradient step of symmetric NMF
// (GNSYMNMF, nmf/gnsym.hpp), as device-resident launch sequences: every
// data-dependent stop of the reference (the ADMM residual test, the CG
// residual test) is a device flag that the later launches of the sequence
// read, so no launch waits on the host.
//
// ADMM solve, for the n rows of X (n x k) given G (k x k FP64, the Gram with
// its regularisation), R (k x n) and the dual D (n x k):
//   rho = trace(G)/k (0.01 if not positive), L = chol(G + rho I)   [one block]
//   up to 5 sweeps:  T = L^-T L^-1 (R + rho (X + D)^T)  per row (a wavefront,
//     lane c holding component c, the triangular solves column by column),
//     X = max(T^T - D, 0), D += X - T^T, with per-block FP64 partial sums of
//     ||X - T^T||^2, ||X - X_prev||^2, ||X||^2 and ||D||^2; then one block
//     sums them in block order and sets the stop flag when ||X - T^T|| <
//     0.01 ||X|| and ||X - X_prev|| < 0.01 ||D||.
#include "common.h"

namespace planc {

namespace {

constexpr int RB = 4;         // rows (wavefronts) per 256-thread block
constexpr int ADMM_NB = 512;  // row blocks of the ADMM sweep (grid-stride)

struct AdmmCtl {
  double rho;
  int stop;
  int bad;  // G + rho I not positive definite
};

// L = chol(G + rho I) (lower, k x k, column-major), rho and the flags.
__global__ void admm_prepare_kernel(const double *G, int k, double *L, AdmmCtl *ctl) {
  __shared__ double sL[64 * 64];
  __shared__ double rho_s;
  __shared__ int bad_s;
  const int tid = threadIdx.x;
  if (tid == 0) {
    double tr = 0.0;
    for (int c = 0; c < k; ++c) tr += G[c + (long)k * c];
    double rho = tr / k;
    rho_s = rho > 0 ? rho : 0.01;
    bad_s = 0;
  }
  __syncthreads();
  for (int e = tid; e < k * k; e += blockDim.x) {
    const int a = e % k, b = e / k;
    sL[e] = G[e] + (a == b ? rho_s : 0.0);
  }
  __syncthreads();
  // Right-looking Cholesky, one column per step.
  for (int j = 0; j < k; ++j) {
    if (tid == 0) {
      double d = sL[j + k * j];
      if (!(d > 0.0)) {
        bad_s = 1;
        d = 1.0;
      }
      sL[j + k * j] = sqrt(d);
    }
    __syncthreads();
    const double ljj = sL[j + k * j];
    for (int i = j + 1 + tid; i < k; i += blockDim.x) sL[i + k * j] /= ljj;
    __syncthreads();
    for (int e = tid; e < (k - j - 1) * (k - j - 1); e += blockDim.x) {
      const int a = j + 1 + e % (k - j - 1), b = j + 1 + e / (k - j - 1);
      if (b <= a) sL[a + k * b] -= sL[a + k * j] * sL[b + k * j];
    }
    __syncthreads();
  }
  for (int e = tid; e < k * k; e += blockDim.x) {
    const int a = e % k, b = e / k;
    L[e] = a >= b ? sL[e] : 0.0;
  }
  if (tid == 0) {
    ctl->rho = rho_s;
    ctl->stop = 0;
    ctl->bad = bad_s;
  }
}

// One ADMM sweep over all rows (skipped once the stop flag is set).
__global__ __launch_bounds__(256) void admm_sweep_kernel(const double *L, int k, const float *R, long r_js,
                                                         long r_cs, float *X, long ldx, float *D, long ldd,
                                                         long n, AdmmCtl *ctl, double *part) {
  __shared__ double sL[64 * 64];
  __shared__ double red[4][RB];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool skip = ctl->stop || ctl->bad;
  double r2 = 0, s2 = 0, x2 = 0, d2 = 0;
  if (!skip) {
    for (int e = threadIdx.x; e < k * k; e += blockDim.x) sL[e] = L[e];
    __syncthreads();
    const double rho = ctl->rho;
    const bool on = lane < k;
    for (long j = (long)blockIdx.x * RB + wave; j < n; j += (long)gridDim.x * RB) {
      double xo = 0, dv = 0, b = 0;
      if (on) {
        xo = X[j + ldx * lane];
        dv = D[j + ldd * lane];
        b = (double)R[j * r_js + (long)lane * r_cs] + rho * (xo + dv);
      }
      // Forward substitution L y = b (column by column).
      for (int i = 0; i < k; ++i) {
        double yi = __shfl(b, i, 64);
        yi /= sL[i + k * i];
        if (lane == i) b = yi;
        if (on && lane > i) b -= sL[lane + k * i] * yi;
      }
      // Back substitution L^T t = y.
      for (int i = k - 1; i >= 0; --i) {
        double ti = __shfl(b, i, 64);
        ti /= sL[i + k * i];
        if (lane == i) b = ti;
        if (on && lane < i) b -= sL[i + k * lane] * ti;
      }
      if (on) {
        double x = b - dv;
        x = x > 0.0 ? x : 0.0;
        const double dn = dv + x - b;
        const float xf = (float)x, df = (float)dn;
        X[j + ldx * lane] = xf;
        D[j + ldd * lane] = df;
        r2 += (x - b) * (x - b);
        s2 += ((double)xf - xo) * ((double)xf - xo);
        x2 += (double)xf * (double)xf;
        d2 += (double)df * (double)df;
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    r2 += __shfl_xor(r2, off, 64);
    s2 += __shfl_xor(s2, off, 64);
    x2 += __shfl_xor(x2, off, 64);
    d2 += __shfl_xor(d2, off, 64);
  }
  if (lane == 0) {
    red[0][wave] = r2;
    red[1][wave] = s2;
    red[2][wave] = x2;
    red[3][wave] = d2;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double t = 0;
    for (int w = 0; w < RB; ++w) t += red[threadIdx.x][w];
    part[4 * blockIdx.x + threadIdx.x] = t;
  }
}

__global__ void admm_check_kernel(const double *part, int nb, AdmmCtl *ctl) {
  if (threadIdx.x != 0 || ctl->stop || ctl->bad) return;
  double s[4] = {0, 0, 0, 0};
  for (int b = 0; b < nb; ++b)
    for (int q = 0; q < 4; ++q) s[q] += part[4 * b + q];
  const double tol = 0.01;
  if (sqrt(s[0]) < tol * sqrt(s[2]) && sqrt(s[1]) < tol * sqrt(s[3])) ctl->stop = 1;
}

__global__ void admm_flag_kernel(const AdmmCtl *ctl, int *err) {
  if (threadIdx.x == 0 && ctl->bad) *err |= 4;
}

}  // namespace

// X (n x k, ld ldx) and its dual D (ld ldd) updated in place by one AO-ADMM
// solve with Gram G (k x k, regularised) and right-hand side R(c, j) at
// R[j * r_js + c * r_cs].  *err |= 4 when G + rho I is not positive definite.
int admm_solve(planc_ctx *ctx, const double *G, int k, const float *R, long r_js, long r_cs, float *X,
               long ldx, float *D, long ldd, long n, int *err) {
  if (k <= 0 || k > 64) return set_error(ctx, PLANC_EUNSUPPORTED, "admm: k must be 1..64");
  double *L = (double *)ws_get(ctx, 22, sizeof(double) * (64 * 64 + 4 * ADMM_NB) + sizeof(AdmmCtl));
  if (!L) return set_error(ctx, PLANC_ENOMEM, "admm: workspace");
  double *part = L + 64 * 64;
  AdmmCtl *ctl = (AdmmCtl *)(part + 4 * ADMM_NB);
  admm_prepare_kernel<<<1, 256, 0, ctx->stream>>>(G, k, L, ctl);
  PLANC_LAUNCH_CHECK(ctx);
  long nb = (n + RB - 1) / RB;
  if (nb > ADMM_NB) nb = ADMM_NB;
  if (nb < 1) nb = 1;
  for (int sweep = 0; sweep < 5; ++sweep) {
    admm_sweep_kernel<<<(unsigned)nb, 256, 0, ctx->stream>>>(L, k, R, r_js, r_cs, X, ldx, D, ldd, n, ctl, part);
    PLANC_LAUNCH_CHECK(ctx);
    admm_check_kernel<<<1, 64, 0, ctx->stream>>>(part, (int)nb, ctl);
    PLANC_LAUNCH_CHECK(ctx);
  }
  if (err) {
    admm_flag_kernel<<<1, 64, 0, ctx->stream>>>(ctl, err);
    PLANC_LAUNCH_CHECK(ctx);
  }
  return PLANC_OK;
}

// ---------------------------------------------------------------------------
// Symmetric NMF, one Gauss-Newton iteration (nmf/gnsym.hpp:178-247) on the
// device.  State: H (n x k FP32, ld n), HtH (k x k FP64) and AHt = H^T A
// (k x n FP32, ld k) of the current H.  CG vectors (k x n) in FP64.

namespace {

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

struct GnCtl {
  double rsold, alpha, beta;
  int active;
};

// grad = -2 (AHt - HtH H^T); P = grad; Dt = 0; partial sums of grad^2.
__global__ __launch_bounds__(256) void gn_grad_kernel(const float *AHt, const double *HtH, const float *H,
                                                      long n, int k, double *grad, double *P, double *Dt,
                                                      double *part) {
  __shared__ double sG[64 * 64];
  __shared__ double red[4];
  for (int e = threadIdx.x; e < k * k; e += blockDim.x) sG[e] = HtH[e];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double acc = 0;
  for (long j = (long)blockIdx.x * 4 + wave; j < n; j += (long)gridDim.x * 4) {
    if (lane < k) {
      double hh = 0;
      for (int t = 0; t < k; ++t) hh += sG[lane + k * t] * (double)H[j + n * t];
      const double g = -2.0 * ((double)AHt[lane + (long)k * j] - hh);
      const long e = lane + (long)k * j;
      grad[e] = g;
      P[e] = g;
      Dt[e] = 0.0;
      acc += g * g;
    }
  }
  acc = wave_sum(acc);
  if (lane == 0) red[wave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void gn_init_kernel(const double *part, int nb, double tol, GnCtl *ctl) {
  if (threadIdx.x != 0) return;
  double s = 0;
  for (int b = 0; b < nb; ++b) s += part[b];
  ctl->rsold = s;
  ctl->active = s > tol ? 1 : 0;
}

// XtX = P H (k x k), FP64 partials per block: part[b][a + k c] = sum_j P(a, j) H(j, c).
__global__ __launch_bounds__(256) void gn_pth_kernel(const double *P, const float *H, long n, int k,
                                                     const GnCtl *ctl, double *part) {
  if (!ctl->active) return;
  const int tid = threadIdx.x;
  double *out = part + (long)blockIdx.x * k * k;
  for (int e = tid; e < k * k; e += blockDim.x) {
    const int a = e % k, c = e / k;
    double s = 0;
    for (long j = blockIdx.x; j < n; j += gridDim.x) s += P[a + (long)k * j] * (double)H[j + n * c];
    out[e] = s;
  }
}

__global__ void gn_pth_reduce_kernel(const double *part, int nb, int k, const GnCtl *ctl, double *XtX) {
  if (!ctl->active) return;
  for (int e = threadIdx.x; e < k * k; e += blockDim.x) {
    double s = 0;
    for (int b = 0; b < nb; ++b) s += part[(long)b * k * k + e];
    XtX[e] = s;
  }
}

// Y = 2 (HtH P + XtX H^T); partial sums of P . Y.
__global__ __launch_bounds__(256) void gn_hess_kernel(const double *HtH, const double *XtX, const double *P,
                                                      const float *H, long n, int k, const GnCtl *ctl,
                                                      double *Y, double *part) {
  __shared__ double sG[64 * 64];
  __shared__ double red[4];
  if (!ctl->active) return;
  for (int e = threadIdx.x; e < k * k; e += blockDim.x) sG[e] = HtH[e];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double acc = 0;
  for (long j = (long)blockIdx.x * 4 + wave; j < n; j += (long)gridDim.x * 4) {
    const double pj = lane < k ? P[lane + (long)k * j] : 0.0;
    const double hj = lane < k ? (double)H[j + n * lane] : 0.0;
    double a = 0, b = 0;
    for (int t = 0; t < k; ++t) {
      const double pt = __shfl(pj, t, 64), ht = __shfl(hj, t, 64);
      if (lane < k) {
        a += sG[lane + k * t] * pt;
        b += XtX[lane + k * t] * ht;
      }
    }
    if (lane < k) {
      const double y = 2.0 * (a + b);
      Y[lane + (long)k * j] = y;
      acc += pj * y;
    }
  }
  acc = wave_sum(acc);
  if (lane == 0) red[wave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void gn_alpha_kernel(const double *part, int nb, GnCtl *ctl) {
  if (threadIdx.x != 0 || !ctl->active) return;
  double s = 0;
  for (int b = 0; b < nb; ++b) s += part[b];
  ctl->alpha = ctl->rsold / s;
}

// Dt += alpha P; grad -= alpha Y; partial sums of grad^2.
__global__ __launch_bounds__(256) void gn_step_kernel(const double *P, const double *Y, long tot,
                                                      const GnCtl *ctl, double *Dt, double *grad,
                                                      double *part) {
  __shared__ double red[4];
  if (!ctl->active) return;
  const double al = ctl->alpha;
  double acc = 0;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    Dt[e] += al * P[e];
    const double g = grad[e] - al * Y[e];
    grad[e] = g;
    acc += g * g;
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void gn_beta_kernel(const double *part, int nb, double tol, GnCtl *ctl) {
  if (threadIdx.x != 0 || !ctl->active) return;
  double rsnew = 0;
  for (int b = 0; b < nb; ++b) rsnew += part[b];
  if (rsnew < tol) {
    ctl->active = 0;
    return;
  }
  ctl->beta = rsnew / ctl->rsold;
  ctl->rsold = rsnew;
}

__global__ void gn_dir_kernel(const double *grad, long tot, const GnCtl *ctl, double *P) {
  if (!ctl->active) return;
  const double be = ctl->beta;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x)
    P[e] = grad[e] + be * P[e];
}

// H = max(H - Dt^T, 0).
__global__ void gn_project_kernel(float *H, long n, int k, const double *Dt) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * k) return;
  const long j = e % n;
  const int c = (int)(e / n);
  const double v = (double)H[e] - Dt[c + (long)k * j];
  H[e] = (float)(v > 0.0 ? v : 0.0);
}

}  // namespace

// One GNSYMNMF iteration up to (not including) the error: the gradient from
// AHt / HtH of the current H, at most k CG iterations, the projected step.
int gnsym_step(planc_ctx *ctx, float *H, long n, int k, const float *AHt, const double *HtH) {
  if (k <= 0 || k > 64) return set_error(ctx, PLANC_EUNSUPPORTED, "gnsym: k must be 1..64");
  const long tot = (long)k * n;
  const int nb = 256;
  const int npart = 1024;  // partial sums of the element-wise kernels (at most 1024 blocks)
  double *buf = (double *)ws_get(ctx, 23, sizeof(double) * (4 * (size_t)tot + (size_t)nb * k * k + k * k + npart) +
                                              sizeof(GnCtl));
  if (!buf) return set_error(ctx, PLANC_ENOMEM, "gnsym: workspace");
  double *grad = buf, *P = buf + tot, *Dt = buf + 2 * tot, *Y = buf + 3 * tot;
  double *ppart = buf + 4 * tot, *XtX = ppart + (size_t)nb * k * k, *part = XtX + k * k;
  GnCtl *ctl = (GnCtl *)(part + npart);
  const double tol = 0.0001;  // cg_tol, nmf/gnsym.hpp:85
  gn_grad_kernel<<<nb, 256, 0, ctx->stream>>>(AHt, HtH, H, n, k, grad, P, Dt, part);
  PLANC_LAUNCH_CHECK(ctx);
  gn_init_kernel<<<1, 64, 0, ctx->stream>>>(part, nb, tol, ctl);
  PLANC_LAUNCH_CHECK(ctx);
  const unsigned eb = (unsigned)std::min<long>(1024, (tot + 255) / 256);
  for (int cg = 0; cg < k; ++cg) {  // cg_max_iters = k
    gn_pth_kernel<<<nb, 256, 0, ctx->stream>>>(P, H, n, k, ctl, ppart);
    PLANC_LAUNCH_CHECK(ctx);
    gn_pth_reduce_kernel<<<1, 256, 0, ctx->stream>>>(ppart, nb, k, ctl, XtX);
    PLANC_LAUNCH_CHECK(ctx);
    gn_hess_kernel<<<nb, 256, 0, ctx->stream>>>(HtH, XtX, P, H, n, k, ctl, Y, part);
    PLANC_LAUNCH_CHECK(ctx);
    gn_alpha_kernel<<<1, 64, 0, ctx->stream>>>(part, nb, ctl);
    PLANC_LAUNCH_CHECK(ctx);
    gn_step_kernel<<<eb, 256, 0, ctx->stream>>>(P, Y, tot, ctl, Dt, grad, part);
    PLANC_LAUNCH_CHECK(ctx);
    gn_beta_kernel<<<1, 64, 0, ctx->stream>>>(part, (int)eb, tol, ctl);
    PLANC_LAUNCH_CHECK(ctx);
    gn_dir_kernel<<<eb, 256, 0, ctx->stream>>>(grad, tot, ctl, P);
    PLANC_LAUNCH_CHECK(ctx);
  }
  gn_project_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, ctx->stream>>>(H, n, k, Dt);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

}  // namespace planc
