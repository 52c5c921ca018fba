// Batched block-principal-pivoting NNLS in FP64: min ||A x - b||, x >= 0,
// given the normal equations G = A^T A (k x k, k <= 64) and right-hand sides
// b = A^T a (one per column).  Restates BPPNNLS::solveNNLSMultipleRHS
// (nnls/bppnnls.hpp:67-191) column by column: each column's pivoting
// (P, Ninf, exchange rules Cols1 / Cols2 / backup rule, 1e-12 clamps) is
// independent of the other columns, so the reference's chunk-wide loop and
// its grouping of equal passive sets (:204-259, a factorisation-sharing
// device) give the same per-column iterates.  MAX_ITERATIONS = 5k per
// column; a column that exceeds it reports -1 (the reference exits).
//
// One wavefront per right-hand side: lane i owns variable i.  The passive
// set is a 64-bit wave-uniform mask (ballots).  Each pivot iteration solves
// G(P,P) x_P = b_P by Gauss-Jordan elimination of the masked k x k system
// (rows/columns outside P replaced by the identity), which for SPD G(P,P)
// needs no pivoting; lane i keeps row i in registers and the pivot row is
// broadcast with v_readlane.  G itself is staged once per block in LDS.
#include "common.h"

namespace planc {

namespace {

__device__ inline double bcast(double v, int lane) {
  int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

__device__ inline unsigned long long ballot64(bool p) { return __ballot(p); }

template <int KM>
__global__ __launch_bounds__(256) void bpp_kernel(const double *G, int k, const float *B, long b_cs,
                                                  long b_rs, long nrhs, float *X, long x_cs, long x_rs,
                                                  int maxit, int *iters_out) {
  __shared__ double Gs[KM][KM + 1];
  for (int e = threadIdx.x; e < KM * KM; e += blockDim.x) {
    int r = e % KM, c = e / KM;
    Gs[r][c] = (r < k && c < k) ? G[r + (size_t)k * c] : 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long col = (long)blockIdx.x * (blockDim.x >> 6) + wave;
  if (col >= nrhs) return;  // whole wavefront leaves together
  const bool live = lane < k;
  const double b = live ? (double)B[col * b_cs + lane * b_rs] : 0.0;
  double x = 0.0, y = -b;
  unsigned long long P = 0ull;
  int pcnt = 3, ninf = k + 1;
  unsigned long long nonopt = ballot64(live && y < 0.0);
  unsigned long long infea = 0ull;
  int notgood = __popcll(nonopt) + __popcll(infea);
  int iter = 0;
  bool failed = false;
  while (notgood > 0) {
    ++iter;
    if (iter > maxit) {
      failed = true;
      break;
    }
    if (notgood < ninf) {
      pcnt = 3;
      ninf = notgood;
      P = (P | nonopt) & ~infea;
    } else if (pcnt >= 1) {
      pcnt -= 1;
      P = (P | nonopt) & ~infea;
    } else {
      int r = 63 - __clzll(nonopt | infea);
      P ^= (1ull << r);
    }
    // Masked system [G(P,P) 0; 0 I] x = [b_P; 0], lane = row.
    const bool pi = (P >> lane) & 1ull;
    double a[KM];
#pragma unroll
    for (int t = 0; t < KM; ++t) {
      bool pt = (P >> t) & 1ull;
      a[t] = (pi && pt) ? Gs[lane < KM ? lane : 0][t] : (lane == t ? 1.0 : 0.0);
    }
    double rhs = pi ? b : 0.0;
#pragma unroll
    for (int j = 0; j < KM; ++j) {
      if ((P >> j) & 1ull) {
        const double piv = bcast(a[j], j);
        const double rj = bcast(rhs, j);
        const double f = (lane != j) ? a[j] / piv : 0.0;
#pragma unroll
        for (int t = 0; t < KM; ++t) {
          if (t != j) {
            double arow = bcast(a[t], j);
            a[t] = fma(-f, arow, a[t]);
          }
        }
        a[j] = (lane != j) ? 0.0 : a[j];
        rhs = fma(-f, rj, rhs);
      }
    }
    double diag = 1.0;
#pragma unroll
    for (int t = 0; t < KM; ++t)
      if (t == lane) diag = a[t];
    x = pi ? rhs / diag : 0.0;
    // Y = G x - b with the full Gram.
    double gx = 0.0;
#pragma unroll
    for (int t = 0; t < KM; ++t) {
      double xt = bcast(x, t);
      if (live) gx = fma(Gs[lane][t], xt, gx);
    }
    y = live ? gx - b : 0.0;
    if (fabs(x) < 1e-12) x = 0.0;
    if (fabs(y) < 1e-12) y = 0.0;
    nonopt = ballot64(live && y < 0.0 && !pi);
    infea = ballot64(live && x < 0.0 && pi);
    notgood = __popcll(nonopt) + __popcll(infea);
  }
  if (live) X[col * x_cs + lane * x_rs] = (float)x;
  if (lane == 0 && iters_out) iters_out[col] = failed ? -1 : iter;
}

__global__ void any_negative_kernel(const int *v, long n, int *flag) {
  __shared__ int s;
  if (threadIdx.x == 0) s = 0;
  __syncthreads();
  int f = 0;
  for (long i = threadIdx.x; i < n; i += blockDim.x) f |= (v[i] < 0);
  if (f) s = 1;
  __syncthreads();
  if (threadIdx.x == 0 && s) *flag = 1;
}

}  // namespace

// Solves every column; *d_failed (device int) is set to 1 when a column
// needed more than 5k pivot iterations.
int bpp_solve(planc_ctx *ctx, const double *G, int k, const float *B, long b_cs, long b_rs, long nrhs,
              float *X, long x_cs, long x_rs, int *d_failed) {
  if (k <= 0 || k > 64) return set_error(ctx, PLANC_EUNSUPPORTED, "bpp: k must be 1..64");
  if (nrhs <= 0) return PLANC_OK;
  int *iters = (int *)ws_get(ctx, 10, sizeof(int) * (size_t)nrhs);
  if (!iters) return set_error(ctx, PLANC_ENOMEM, "bpp: workspace");
  unsigned nb = (unsigned)((nrhs + 3) / 4);
  if (k <= 16)
    bpp_kernel<16><<<nb, 256, 0, ctx->stream>>>(G, k, B, b_cs, b_rs, nrhs, X, x_cs, x_rs, 5 * k, iters);
  else if (k <= 32)
    bpp_kernel<32><<<nb, 256, 0, ctx->stream>>>(G, k, B, b_cs, b_rs, nrhs, X, x_cs, x_rs, 5 * k, iters);
  else
    bpp_kernel<64><<<nb, 256, 0, ctx->stream>>>(G, k, B, b_cs, b_rs, nrhs, X, x_cs, x_rs, 5 * k, iters);
  PLANC_LAUNCH_CHECK(ctx);
  if (d_failed) {
    any_negative_kernel<<<1, 256, 0, ctx->stream>>>(iters, nrhs, d_failed);
    PLANC_LAUNCH_CHECK(ctx);
  }
  return PLANC_OK;
}

}  // namespace planc
