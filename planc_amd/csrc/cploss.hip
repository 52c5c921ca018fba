// CP reconstruction loss 0.5 ||X - [[A_0 .. A_{N-1}]]||^2 and its gradient.
//
// Forward: with the modes split at s (left block L = prod_{i<=s} I_i, right
// R), the reconstruction is Xhat_(L x R) = KRP_L * KRP_R^T.  The fused kernel
// builds Khatri-Rao tiles of both sides in LDS from the factors (never
// stored in HBM), forms a 128 (r) x 64 (l) tile of Xhat with
// v_mfma_f32_16x16x4f32, reads the matching X tile once (16 consecutive l
// per lane group, coalesced) and accumulates (x - xhat)^2 in FP64.  Blocks
// keep their KRP_R tile and walk several l tiles.  Partials per block are
// summed in block order.
//
// Backward: dL/dA_n = A_n * (Hadamard of the other Grams) - MTTKRP_n^T, the
// transposed NTFNES gradient S Y - M (ntf/ntfnes.hpp:138-147) without the
// lambda shift.  All N MTTKRPs come from the two partial contractions of
// the dimension tree, each one pass over X with the KRP fused.
#include <mutex>
#include <string>

#include "common.h"

namespace planc {

namespace {

constexpr int RT = 128;  // rows of the Xhat^T tile (right index r)
constexpr int LT = 64;   // cols (left index l)

struct CPParams {
  const float *X;
  long L, R;
  int k;
  int nl, nr;  // modes in the left / right block
  const float *fl[6];
  long dl[6];
  const float *fr[6];
  long dr[6];
  int ltiles_per_block;
  double *part;
};

template <int KP>
__device__ inline void krp_tile(float (*dst)[KP + 1], int rows, long base, long limit, int nf,
                                const float *const *f, const long *dims, int k) {
  for (int e = threadIdx.x; e < rows * KP; e += blockDim.x) {
    int r = e / KP, c = e % KP;
    long g = base + r;
    float v = 0.0f;
    if (g < limit && c < k) {
      v = 1.0f;
      long rem = g;
      for (int q = 0; q < nf; ++q) {
        long d = rem % dims[q];
        rem /= dims[q];
        v *= f[q][d + dims[q] * c];
      }
    }
    dst[r][c] = v;
  }
}

template <int KP>
__global__ __launch_bounds__(256) void cp_loss_fwd_kernel(CPParams p) {
  __shared__ float KR[RT][KP + 1];
  __shared__ float KL[LT][KP + 1];
  __shared__ double red[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long r0 = (long)blockIdx.x * RT;
  krp_tile<KP>(KR, RT, r0, p.R, p.nr, p.fr, p.dr, p.k);
  double acc2 = 0.0;
  const long ltiles = (p.L + LT - 1) / LT;
  const long lt_begin = (long)blockIdx.y * p.ltiles_per_block;
  const long lt_end = min(ltiles, lt_begin + p.ltiles_per_block);
  for (long lt = lt_begin; lt < lt_end; ++lt) {
    const long l0 = lt * LT;
    __syncthreads();
    krp_tile<KP>(KL, LT, l0, p.L, p.nl, p.fl, p.dl, p.k);
    __syncthreads();
    floatx4 acc[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < KP; kk += 4) {
      const int kq = kk + (lane >> 4);
      float a0 = KR[wave * 32 + (lane & 15)][kq];
      float a1 = KR[wave * 32 + 16 + (lane & 15)][kq];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        float bv = KL[16 * b + (lane & 15)][kq];
        acc[0][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, bv, acc[0][b], 0, 0, 0);
        acc[1][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, bv, acc[1][b], 0, 0, 0);
      }
    }
    // acc[a][b][i] = Xhat(l = l0 + 16b + lane%16, r = r0 + wave*32 + 16a + 4*(lane/16) + i)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        long r = r0 + wave * 32 + 16 * a + 4 * (lane >> 4) + i;
        if (r >= p.R) continue;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          long l = l0 + 16 * b + (lane & 15);
          if (l < p.L) {
            float d = p.X[l + p.L * r] - acc[a][b][i];
            acc2 += (double)d * (double)d;
          }
        }
      }
  }
  // Block reduction (fixed order).
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc2 += __shfl_xor(acc2, off, 64);
  __syncthreads();
  if (lane == 0) red[wave] = acc2;
  __syncthreads();
  if (threadIdx.x == 0) p.part[blockIdx.y * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void sum_half_kernel(const double *part, long n, double *out) {
  if (threadIdx.x >= 64) return;
  double t = 0.0;
  for (long b = threadIdx.x; b < n; b += 64) t += part[b];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
  if (threadIdx.x == 0) *out = 0.5 * t;
}

// grad(i, c) = sum_t A(i, t) S(t, c) - M(c, i); A I x k, M k x I, S k x k.
__global__ void cp_grad_kernel(const float *A, long I, int k, const double *S, const float *M,
                               float *grad) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= I * k) return;
  long i = e % I;
  int c = (int)(e / I);
  double s = 0.0;
  for (int t = 0; t < k; ++t) s += (double)A[i + I * t] * S[t + (long)k * c];
  grad[i + I * c] = (float)(s - (double)M[c + (long)k * i]);
}

int balanced(int N, const long *dims) {
  int best = N / 2 < N - 1 ? N / 2 : N - 2;
  long bestv = -1;
  for (int s = 0; s <= N - 2; ++s) {
    long L = 1, R = 1;
    for (int i = 0; i <= s; ++i) L *= dims[i];
    for (int i = s + 1; i < N; ++i) R *= dims[i];
    long v = L > R ? L : R;
    if (bestv < 0 || v < bestv || (v == bestv && s == N / 2)) {
      bestv = v;
      best = s;
    }
  }
  return best;
}

}  // namespace

namespace {

// Block partition of one forward evaluation: the grid and how many FP64
// block partials it writes.  Both entry points use the same partition, so
// their sums are formed in the same order.
struct CPLaunch {
  CPParams p;
  dim3 grid;
  long nparts;
};

int cp_loss_plan(planc_ctx *ctx, const float *X, int N, const long *dims, const float *const *F,
                 int k, CPLaunch *out) {
  if (N < 2 || N > 7 || k <= 0 || k > 64) return set_error(ctx, PLANC_EINVAL, "cp_loss: 2..7 modes, k 1..64");
  int s = balanced(N, dims);
  CPParams p{};
  p.X = X;
  p.k = k;
  p.L = 1;
  p.R = 1;
  p.nl = s + 1;
  p.nr = N - 1 - s;
  if (p.nl > 6 || p.nr > 6) return set_error(ctx, PLANC_EINVAL, "cp_loss: too many modes per side");
  for (int i = 0; i <= s; ++i) {
    p.fl[i] = F[i];
    p.dl[i] = dims[i];
    p.L *= dims[i];
  }
  for (int i = s + 1; i < N; ++i) {
    p.fr[i - s - 1] = F[i];
    p.dr[i - s - 1] = dims[i];
    p.R *= dims[i];
  }
  long rblocks = (p.R + RT - 1) / RT;
  long ltiles = (p.L + LT - 1) / LT;
  long want = (2L * ctx->num_cu * 4 + rblocks - 1) / rblocks;  // ~8 blocks per CU
  if (want < 1) want = 1;
  if (want > ltiles) want = ltiles;
  p.ltiles_per_block = (int)((ltiles + want - 1) / want);
  long lgroups = (ltiles + p.ltiles_per_block - 1) / p.ltiles_per_block;
  out->p = p;
  out->grid = dim3((unsigned)rblocks, (unsigned)lgroups);
  out->nparts = rblocks * lgroups;
  return PLANC_OK;
}

// Every block writes its own partial (nothing is accumulated into part), and
// the final kernel reads exactly nparts of them: part needs no clearing.
int cp_loss_launch(planc_ctx *ctx, CPLaunch &c, double *part, double *d_loss) {
  c.p.part = part;
  const int k = c.p.k;
  if (k <= 16)
    cp_loss_fwd_kernel<16><<<c.grid, 256, 0, ctx->stream>>>(c.p);
  else if (k <= 32)
    cp_loss_fwd_kernel<32><<<c.grid, 256, 0, ctx->stream>>>(c.p);
  else if (k <= 48)
    cp_loss_fwd_kernel<48><<<c.grid, 256, 0, ctx->stream>>>(c.p);
  else
    cp_loss_fwd_kernel<64><<<c.grid, 256, 0, ctx->stream>>>(c.p);
  PLANC_LAUNCH_CHECK(ctx);
  sum_half_kernel<<<1, 64, 0, ctx->stream>>>(part, c.nparts, d_loss);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

// Partials of the allocation-free forward: a file-scope device array, one
// copy per GPU, held until the process ends.  Its name is a host-side
// stand-in; kernels get the per-device address that hipGetSymbolAddress
// returns at context creation (cp_noalloc_setup), never the name itself.
constexpr long kNoallocParts = 16384;  // 128 KiB
static __device__ __align__(256) double scratch[kNoallocParts];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices];
int g_scratch_state[kMaxDevices];  // 0 not set up, 1 ready, -1 lookup failed

}  // namespace

int cp_loss_forward(planc_ctx *ctx, const float *X, int N, const long *dims, const float *const *F,
                    int k, double *d_loss) {
  CPLaunch c;
  int rc = cp_loss_plan(ctx, X, N, dims, F, k, &c);
  if (rc) return rc;
  double *part = (double *)ws_get(ctx, 3, sizeof(double) * (size_t)c.nparts);
  if (!part) return set_error(ctx, PLANC_ENOMEM, "cp_loss: workspace");
  return cp_loss_launch(ctx, c, part, d_loss);
}

// Called by planc_create with `device` current: looks up this device's copy
// of the partials array once per process.
int cp_noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return PLANC_EUNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_state[device] != 0) return g_scratch_state[device] > 0 ? PLANC_OK : PLANC_EHIP;
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    g_scratch_state[device] = -1;
    return PLANC_EHIP;
  }
  g_scratch_ptr[device] = (double *)ptr;
  g_scratch_state[device] = 1;
  return PLANC_OK;
}

int cp_loss_forward_noalloc(planc_ctx *ctx, const float *X, int N, const long *dims,
                            const float *const *F, int k, double *d_loss) {
  const int dev = ctx->device;
  double *part = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dev >= 0 && dev < kMaxDevices && g_scratch_state[dev] > 0) part = g_scratch_ptr[dev];
  }
  if (!part)
    return set_error(ctx, PLANC_EHIP,
                     "cp_loss_fwd_noalloc: the device's partials array could not be located "
                     "(hipGetSymbolAddress failed at context creation)");
  CPLaunch c;
  int rc = cp_loss_plan(ctx, X, N, dims, F, k, &c);
  if (rc) return rc;
  if (c.nparts > kNoallocParts)
    return set_error(ctx, PLANC_EINVAL,
                     "cp_loss_fwd_noalloc: " + std::to_string(c.nparts) +
                         " block partials exceed the fixed array of " + std::to_string(kNoallocParts) +
                         " (the right block of modes has too many rows); use planc_cp_loss_fwd");
  return cp_loss_launch(ctx, c, part, d_loss);
}

int cp_grad_from_mttkrp(planc_ctx *ctx, const float *A, long I, int k, const double *S, const float *M,
                        float *grad) {
  long tot = I * k;
  if (tot == 0) return PLANC_OK;
  cp_grad_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, ctx->stream>>>(A, I, k, S, M, grad);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int cp_split(int N, const long *dims) { return balanced(N, dims); }

}  // namespace planc
