// Synthetic inputs generated on the device from the Philox4x64-10 stream of
// philox.h (bit-identical uniforms to planc_amd/datagen.py on the host), so
// multi-gigabyte inputs never cross PCIe.  No reference counterpart: the
// reference draws its synthetic inputs from Armadillo's RNG.
#include "common.h"
#include "philox.h"

namespace planc {

namespace {

__global__ void gen_uniform_kernel(float *out, long count, uint64_t seed, uint64_t stream) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (long)gridDim.x * blockDim.x)
    out[e] = (float)planc_rng::u01(planc_rng::word(seed, stream, (uint64_t)e));
}

// A(i, j) = sum_c W0(i, c) H0(c, j) (FP64) + sigma * |N(0,1)|_(i + m j).
__global__ void gen_lowrank_kernel(long m, long n, int r, const float *W0, const float *H0,
                                   double sigma, uint64_t seed, uint64_t stream, float *A, long lda) {
  long tot = m * n;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    long i = e % m, j = e / m;
    double s = 0.0;
    for (int c = 0; c < r; ++c) s += (double)W0[i + m * c] * (double)H0[c + (long)r * j];
    if (sigma != 0.0) s += sigma * fabs(planc_rng::normal(seed, stream, (uint64_t)e));
    A[i + lda * j] = (float)s;
  }
}

struct FacSet {
  const float *f[8];
  long dim[8];
  int n;
};

// X(idx) = sum_c prod_i F_i(idx_i, c) + sigma * |N(0,1)|_idx, mode 0 fastest.
__global__ void gen_cp_kernel(FacSet fs, int r, long numel, double sigma, uint64_t seed,
                              uint64_t stream, float *X) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < numel; e += (long)gridDim.x * blockDim.x) {
    long idx[8];
    long rem = e;
    for (int t = 0; t < fs.n; ++t) {
      idx[t] = rem % fs.dim[t];
      rem /= fs.dim[t];
    }
    double s = 0.0;
    for (int c = 0; c < r; ++c) {
      double p = 1.0;
      for (int t = 0; t < fs.n; ++t) p *= (double)fs.f[t][idx[t] + fs.dim[t] * c];
      s += p;
    }
    if (sigma != 0.0) s += sigma * fabs(planc_rng::normal(seed, stream, (uint64_t)e));
    X[e] = (float)s;
  }
}

unsigned grid_for(long n) {
  long b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace

}  // namespace planc

using namespace planc;

extern "C" {

int planc_gen_uniform(planc_ctx *ctx, float *d_out, long count, unsigned long long seed,
                      unsigned long long stream) {
  if (!ctx) return PLANC_EINVAL;
  if (count < 0 || (count > 0 && !d_out)) return set_error(ctx, PLANC_EINVAL, "planc_gen_uniform: bad arguments");
  if (count == 0) return PLANC_OK;
  gen_uniform_kernel<<<grid_for(count), 256, 0, ctx->stream>>>(d_out, count, seed, stream);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int planc_gen_lowrank(planc_ctx *ctx, long m, long n, int r, const float *d_W0, const float *d_H0,
                      double sigma, unsigned long long seed, unsigned long long stream, float *d_A,
                      long lda) {
  if (!ctx) return PLANC_EINVAL;
  if (m <= 0 || n <= 0 || r <= 0 || lda < m || !d_W0 || !d_H0 || !d_A)
    return set_error(ctx, PLANC_EINVAL, "planc_gen_lowrank: bad arguments");
  gen_lowrank_kernel<<<grid_for(m * n), 256, 0, ctx->stream>>>(m, n, r, d_W0, d_H0, sigma, seed, stream,
                                                               d_A, lda);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int planc_gen_cp_tensor(planc_ctx *ctx, int nmodes, const long *dims, const float *const *d_factors,
                        int r, double sigma, unsigned long long seed, unsigned long long stream,
                        float *d_X) {
  if (!ctx) return PLANC_EINVAL;
  if (nmodes < 1 || nmodes > 8 || !dims || !d_factors || r <= 0 || !d_X)
    return set_error(ctx, PLANC_EINVAL, "planc_gen_cp_tensor: bad arguments");
  FacSet fs{};
  fs.n = nmodes;
  long numel = 1;
  for (int i = 0; i < nmodes; ++i) {
    if (dims[i] <= 0) return set_error(ctx, PLANC_EINVAL, "planc_gen_cp_tensor: empty mode");
    fs.f[i] = d_factors[i];
    fs.dim[i] = dims[i];
    numel *= dims[i];
  }
  gen_cp_kernel<<<grid_for(numel), 256, 0, ctx->stream>>>(fs, r, numel, sigma, seed, stream, d_X);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

}  // extern "C"
