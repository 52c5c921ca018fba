// C ABI of libplanc_hip.so: contexts, memory helpers, kernel-level entry
// points and the device-resident computeNMF / computeNTF loops.
#include <algorithm>
#include <functional>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"

namespace planc {

int set_error(planc_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->last_error = msg;
  return code;
}

void *ws_get(planc_ctx *ctx, int slot, size_t bytes) {
  Workspace &w = ctx->ws[slot];
  if (bytes == 0) bytes = 256;
  if (w.bytes >= bytes) return w.ptr;
  if (w.ptr) {
    // Queued kernels may still read the old buffer.
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return nullptr;
    (void)hipFree(w.ptr);
    w.ptr = nullptr;
    w.bytes = 0;
  }
  size_t b = bytes + bytes / 8;
  if (hipMalloc(&w.ptr, b) != hipSuccess) {
    w.ptr = nullptr;
    return nullptr;
  }
  w.bytes = b;
  return w.ptr;
}

// Dimension-tree split: the s in [0, N-2] minimising the larger of the two
// projections (left product L = prod_{i<=s} I_i, right R = prod_{i>s} I_i);
// the reference's s = N/2 (ntf/auntf.hpp:123) when it ties.
int balanced_split(int N, const long *dims) {
  int best = N / 2 < N - 1 ? N / 2 : N - 2;
  long bestv = -1;
  for (int s = 0; s <= N - 2; ++s) {
    long L = 1, R = 1;
    for (int i = 0; i <= s; ++i) L *= dims[i];
    for (int i = s + 1; i < N; ++i) R *= dims[i];
    long v = std::max(L, R);
    if (bestv < 0 || v < bestv || (v == bestv && s == N / 2)) {
      bestv = v;
      best = s;
    }
  }
  return best;
}

}  // namespace planc

using namespace planc;

namespace {

int check_ctx(planc_ctx *ctx) { return ctx ? PLANC_OK : PLANC_EINVAL; }

bool valid_dtype(int dt) { return dt == PLANC_F32 || dt == PLANC_BF16 || dt == PLANC_F16; }

// MTTKRP of mode n (k x I_n, ld ldo) with the KRP fused.  P is an L' x k
// scratch for the middle modes.  Direct form: mode 0 and N-1 are one
// product each; a middle mode n is the right contraction over modes > n
// followed by the contraction of modes < n.
int mttkrp_direct(planc_ctx *ctx, int dtype, const void *X, int N, const long *dims,
                  const float *const *F, int k, int n, float *out, long ldo, float *P) {
  long numel = 1;
  for (int i = 0; i < N; ++i) numel *= dims[i];
  if (N == 1) return set_error(ctx, PLANC_EINVAL, "mttkrp: need >= 2 modes");
  if (n == 0) {
    KrpDesc kd{};
    kd.nf = N - 1;
    for (int q = 1; q < N; ++q) {
      kd.f[q - 1] = F[q];
      kd.dim[q - 1] = dims[q];
      kd.ld[q - 1] = dims[q];
    }
    return thin_gemm(ctx, dtype, 0, X, dims[0], dims[0], numel / dims[0], nullptr, 0, &kd, nullptr,
                     nullptr, k, out, ldo, 1, 0);
  }
  if (n == N - 1) {
    KrpDesc kd{};
    kd.nf = N - 1;
    for (int q = 0; q < N - 1; ++q) {
      kd.f[q] = F[q];
      kd.dim[q] = dims[q];
      kd.ld[q] = dims[q];
    }
    long rest = numel / dims[N - 1];
    return thin_gemm(ctx, dtype, 1, X, rest, dims[N - 1], rest, nullptr, 0, &kd, nullptr, nullptr, k,
                     out, ldo, 1, 0);
  }
  long Lp = 1;
  for (int i = 0; i <= n; ++i) Lp *= dims[i];
  KrpDesc kd{};
  kd.nf = N - 1 - n;
  for (int q = n + 1; q < N; ++q) {
    kd.f[q - n - 1] = F[q];
    kd.dim[q - n - 1] = dims[q];
    kd.ld[q - n - 1] = dims[q];
  }
  int rc = thin_gemm(ctx, dtype, 0, X, Lp, Lp, numel / Lp, nullptr, 0, &kd, nullptr, nullptr, k, P, Lp,
                     0, 0);
  if (rc) return rc;
  long ldf[8];
  for (int i = 0; i <= n; ++i) ldf[i] = dims[i];
  return contract_projection(ctx, P, Lp, n + 1, dims, F, ldf, n, k, out, ldo);
}

__global__ void counter_set_kernel(int *c, int v) { *c = v; }
__global__ void counter_append_kernel(const double *src, double *dst, int *c) {
  dst[*c] = *src;
  *c += 1;
}

int counter_set(planc_ctx *ctx, int *c, int v) {
  counter_set_kernel<<<1, 1, 0, ctx->stream>>>(c, v);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int counter_append(planc_ctx *ctx, const double *src, double *dst, int *c) {
  counter_append_kernel<<<1, 1, 0, ctx->stream>>>(src, dst, c);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

bool graph_lookup(planc_ctx *ctx, const std::vector<uintptr_t> &key, hipGraphExec_t *exec) {
  for (auto &g : ctx->graphs)
    if (g.key == key) {
      *exec = g.exec;
      return true;
    }
  return false;
}

// Caches exec under key; at 16 entries the oldest one other than keep is
// destroyed (no cached graph is in flight: callers store before launching).
void graph_store(planc_ctx *ctx, const std::vector<uintptr_t> &key, hipGraphExec_t exec,
                 hipGraphExec_t keep = nullptr) {
  if (ctx->graphs.size() >= 16) {
    auto victim = ctx->graphs.begin();
    if (victim->exec == keep) ++victim;
    (void)hipGraphExecDestroy(victim->exec);
    ctx->graphs.erase(victim);
  }
  ctx->graphs.push_back({key, exec});
}

// The computeNMF loop of MUNMF / HALSNMF / BPPNMF (nmf/mu.hpp:48-91,
// nmf/hals.hpp:50-116, nmf/bppnmf.hpp:187-241) shared by the dense and the
// sparse entry points; only the two products with A and ||A||^2 differ.
struct NMFProducts {
  std::function<int(const float *, float *)> wta;  // F^T A (k x n, ld k), F m x k (ld m)
  std::function<int(const float *, float *)> ah;   // A F (m x k, ld m), F n x k (ld n)
  std::function<int(double *)> norma2; // ||A||_F^2 into a device double
  bool bpp_prenormalize = false;       // sparse builds' BPPNMF start (see below)
};

int nmf_loop(planc_ctx *ctx, const planc_nmf_desc *desc, NMFProducts &P, float *d_W, float *d_H,
             double *h_relerr) {
  const long m = desc->m, n = desc->n;
  const int k = desc->k, iters = desc->iters, algo = desc->algo;
  // Loop buffers.
  float *fbuf = (float *)ws_get(ctx, 4, sizeof(float) * ((size_t)k * n + (size_t)m * k));
  const size_t kk = (size_t)k * k;
  double *dbuf = (double *)ws_get(ctx, 5, sizeof(double) * (4 * kk + 8 + (size_t)iters));
  if (!fbuf || !dbuf) return set_error(ctx, PLANC_ENOMEM, "planc_nmf_run: workspace");
  float *WtA = fbuf;                // k x n
  float *AH = fbuf + (size_t)k * n;  // m x k
  double *WtW = dbuf, *HtH = dbuf + kk, *Greg = dbuf + 2 * kk, *WtW2 = dbuf + 3 * kk;
  double *normA2 = dbuf + 4 * kk, *dotv = normA2 + 1, *nrm = normA2 + 2;  // nrm unused slot
  double *rel = normA2 + 8;
  (void)nrm;
  int rc;
#define RC(x)          \
  do {                 \
    rc = (x);          \
    if (rc) return rc; \
  } while (0)
  if (desc->compute_error) {
    if (desc->normA2 > 0)
      PLANC_HIP_CHECK(ctx, hipMemcpyAsync(normA2, &desc->normA2, sizeof(double), hipMemcpyHostToDevice,
                                          ctx->stream));
    else
      RC(P.norma2(normA2));
  }
  const double sr = desc->symm_reg;
  if (sr > 0 && (algo != PLANC_ANLSBPP || m != n))
    return set_error(ctx, PLANC_EINVAL,
                     "planc_nmf_run: symmetric regularisation needs ANLS-BPP and a square A");
  if (algo == PLANC_AOADMM) {
    // AOADMMNMF::computeNMF (nmf/aoadmm.hpp:88-192): constructor
    // normalize_by_W; per iteration an ADMM solve for H (Gram W^T W + regH,
    // right-hand side W^T A, dual V) then for W (H^T H + regW, (A H)^T, dual
    // U); duals start at zero; no final normalisation.
    int *dflag = (int *)ws_get(ctx, 11, sizeof(int));
    float *duals = (float *)ws_get(ctx, 33, sizeof(float) * (size_t)(m + n) * k);
    if (!dflag || !duals) return set_error(ctx, PLANC_ENOMEM, "planc_nmf_run: workspace");
    float *U = duals, *V = duals + (size_t)m * k;
    PLANC_HIP_CHECK(ctx, hipMemsetAsync(dflag, 0, sizeof(int), ctx->stream));
    PLANC_HIP_CHECK(ctx, hipMemsetAsync(duals, 0, sizeof(float) * (size_t)(m + n) * k, ctx->stream));
    RC(planc_normalize_by_w(ctx, d_W, m, m, d_H, n, n, k));
    for (int it = 0; it < iters; ++it) {
      RC(P.wta(d_W, WtA));
      RC(gram(ctx, d_W, m, k, m, WtW));
      PLANC_HIP_CHECK(ctx, hipMemcpyAsync(Greg, WtW, kk * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
      RC(apply_reg(ctx, Greg, k, desc->regH));
      RC(admm_solve(ctx, Greg, k, WtA, k, 1, d_H, n, V, n, n, dflag));
      RC(P.ah(d_H, AH));
      RC(gram(ctx, d_H, n, k, n, HtH));
      PLANC_HIP_CHECK(ctx, hipMemcpyAsync(Greg, HtH, kk * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
      RC(apply_reg(ctx, Greg, k, desc->regW));
      RC(admm_solve(ctx, Greg, k, AH, 1, m, d_W, m, U, m, m, dflag));
      if (desc->compute_error) {
        RC(gram(ctx, d_W, m, k, m, WtW2));
        RC(dot(ctx, d_W, 1, m, AH, 1, m, m, k, dotv, 0));
        RC(nmf_relerr(ctx, normA2, dotv, WtW2, HtH, k, rel + it, nullptr));
      }
    }
    int hflag = 0;
    PLANC_HIP_CHECK(ctx, hipMemcpyAsync(&hflag, dflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    if (h_relerr && desc->compute_error && iters > 0)
      PLANC_HIP_CHECK(ctx, hipMemcpyAsync(h_relerr, rel, sizeof(double) * iters, hipMemcpyDeviceToHost, ctx->stream));
    PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (hflag) return set_error(ctx, PLANC_ENOCONV, "AO-ADMM: shifted Gram not positive definite");
    return PLANC_OK;
  }
  if (algo == PLANC_GNSYM) {
    // GNSYMNMF::computeNMF (nmf/gnsym.hpp:178-260): A ~ H H^T.  H^T A and
    // H^T H of the current H feed the gradient; after the step they are
    // recomputed for the error and reused by the next gradient.
    if (m != n) return set_error(ctx, PLANC_EINVAL, "planc_nmf_run: GNSYM needs a square A");
    RC(gram(ctx, d_H, n, k, n, HtH));
    RC(P.wta(d_H, WtA));
    for (int it = 0; it < iters; ++it) {
      RC(gnsym_step(ctx, d_H, n, k, WtA, HtH));
      RC(gram(ctx, d_H, n, k, n, HtH));
      RC(P.wta(d_H, WtA));
      if (desc->compute_error) {
        RC(dot(ctx, d_H, 1, n, WtA, k, 1, n, k, dotv, 0));
        RC(nmf_relerr(ctx, normA2, dotv, HtH, HtH, k, rel + it, nullptr));
      }
    }
    if (h_relerr && desc->compute_error && iters > 0)
      PLANC_HIP_CHECK(ctx, hipMemcpyAsync(h_relerr, rel, sizeof(double) * iters, hipMemcpyDeviceToHost, ctx->stream));
    PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return PLANC_OK;
  }
  if (algo == PLANC_ANLSBPP) {
    // BPPNMF::computeNMF (nmf/bppnmf.hpp:187-241): W given H first, then H
    // given W, each a batched NNLS on the FP64 normal equations.
    int *dflag = (int *)ws_get(ctx, 11, sizeof(int));
    if (!dflag) return set_error(ctx, PLANC_ENOMEM, "planc_nmf_run: workspace");
    PLANC_HIP_CHECK(ctx, hipMemsetAsync(dflag, 0, sizeof(int), ctx->stream));
    // Sparse builds construct a HALSNMF from W, H "to get proper
    // initializations" and take its factors without running it
    // (nmf/bppnmf.hpp:195-201): only the HALSNMF constructor's
    // normalize_by_W (nmf/hals.hpp:46) takes effect.
    if (P.bpp_prenormalize) RC(planc_normalize_by_w(ctx, d_W, m, m, d_H, n, n, k));
    for (int it = 0; it < iters; ++it) {
      RC(gram(ctx, d_H, n, k, n, HtH));
      PLANC_HIP_CHECK(ctx, hipMemcpyAsync(Greg, HtH, kk * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
      RC(apply_reg(ctx, Greg, k, desc->regW));
      RC(P.ah(d_H, AH));
      RC(symm_reg_apply(ctx, Greg, k, AH, 1, m, d_H, n, sr));
      RC(bpp_solve(ctx, Greg, k, AH, 1, m, m, d_W, 1, m, dflag));
      RC(gram(ctx, d_W, m, k, m, WtW));
      PLANC_HIP_CHECK(ctx, hipMemcpyAsync(Greg, WtW, kk * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
      RC(apply_reg(ctx, Greg, k, desc->regH));
      RC(P.wta(d_W, WtA));
      if (sr > 0) {
        // The error below needs the plain W^T A: keep the regularised
        // right-hand side in the AH buffer (m x k >= k x n when m == n).
        PLANC_HIP_CHECK(ctx, hipMemcpyAsync(AH, WtA, sizeof(float) * (size_t)k * n,
                                            hipMemcpyDeviceToDevice, ctx->stream));
        RC(symm_reg_apply(ctx, Greg, k, AH, k, 1, d_W, m, sr));
        RC(bpp_solve(ctx, Greg, k, AH, k, 1, n, d_H, 1, n, dflag));
      } else {
        RC(bpp_solve(ctx, Greg, k, WtA, k, 1, n, d_H, 1, n, dflag));
      }
      if (desc->compute_error) {
        // tr(H^T A^T W) = <H, (W^T A)^T> with the W^T A of the H half-step.
        RC(dot(ctx, d_H, 1, n, WtA, k, 1, n, k, dotv, 0));
        RC(gram(ctx, d_H, n, k, n, HtH));
        RC(nmf_relerr(ctx, normA2, dotv, WtW, HtH, k, rel + it, nullptr));
      }
    }
    RC(planc_normalize_by_w(ctx, d_W, m, m, d_H, n, n, k));
    int hflag = 0;
    PLANC_HIP_CHECK(ctx, hipMemcpyAsync(&hflag, dflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    if (h_relerr && desc->compute_error && iters > 0)
      PLANC_HIP_CHECK(ctx, hipMemcpyAsync(h_relerr, rel, sizeof(double) * iters, hipMemcpyDeviceToHost, ctx->stream));
    PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (hflag) return set_error(ctx, PLANC_ENOCONV, "BPP failed: more than 5k pivot iterations");
    return PLANC_OK;
  }
  double *wnorms = (double *)ws_get(ctx, 9, sizeof(double) * k);
  if (!wnorms) return set_error(ctx, PLANC_ENOMEM, "planc_nmf_run: workspace");
  if (algo == PLANC_HALS) {
    RC(hals_sync_reset(ctx));
    RC(planc_normalize_by_w(ctx, d_W, m, m, d_H, n, n, k));
  }
  const bool regW_on = desc->regW[0] > 0 || desc->regW[1] > 0;
  const bool regH_on = desc->regH[0] > 0 || desc->regH[1] > 0;
  // W^T W of the current W: computed once per iteration (after the W
  // half-step) and carried into the next H half-step -- rescaled by the
  // normalisation for HALS -- instead of a second pass over W.
  bool wtw_valid = false;
  // The H rows' update and H^T H in one pass over H (and, for MU, the W
  // rows' update, W^T W and <W, A H> in one pass over W), where the layouts
  // allow; the Grams are the same sums, accumulated in another order.
  const bool fuseH = rows_gram_ok(d_H, n, WtA, k, 1, k);
  const bool fuseW = algo == PLANC_MU && rows_gram_ok(d_W, m, AH, 1, m, k);
  for (int it = 0; it < iters; ++it) {
    // H half-step: W^T A, W^T W (+ regH).
    RC(P.wta(d_W, WtA));
    if (!wtw_valid) RC(gram(ctx, d_W, m, k, m, WtW));
    const double *Gh = WtW;
    if (regH_on) {
      PLANC_HIP_CHECK(ctx, hipMemcpyAsync(Greg, WtW, kk * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
      RC(apply_reg(ctx, Greg, k, desc->regH));
      Gh = Greg;
    }
    if (fuseH)
      RC(row_update_gram(ctx, algo == PLANC_MU ? 0 : 1, d_H, n, k, n, WtA, k, 1, Gh, 1e-17f, HtH, nullptr,
                         nullptr));
    else
      RC(row_update(ctx, algo == PLANC_MU ? 0 : 1, d_H, n, k, n, WtA, k, 1, Gh, 1e-17f));
    // W half-step: A H, H^T H (+ regW).
    RC(P.ah(d_H, AH));
    if (!fuseH) RC(gram(ctx, d_H, n, k, n, HtH));
    const double *Gw = HtH;
    if (regW_on) {
      PLANC_HIP_CHECK(ctx, hipMemcpyAsync(Greg, HtH, kk * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
      RC(apply_reg(ctx, Greg, k, desc->regW));
      Gw = Greg;
    }
    // <W, A H> with the H of this iteration equals tr(H^T A^T W) of
    // computeObjectiveError (common/nmf.hpp:391-396) without the third
    // product; both traces are invariant under normalize_by_W.
    double *dparts = nullptr;
    long ndp = 0;
    if (fuseW) {
      RC(row_update_gram(ctx, 0, d_W, m, k, m, AH, 1, m, Gw, 1e-17f, WtW2,
                         desc->compute_error ? &dparts : nullptr, &ndp));
    } else {
      if (algo == PLANC_MU)
        RC(row_update(ctx, 0, d_W, m, k, m, AH, 1, m, Gw, 1e-17f));
      else
        RC(hals_cols(ctx, d_W, m, k, m, AH, m, Gw));
      RC(gram(ctx, d_W, m, k, m, WtW2));
      if (desc->compute_error) RC(dot_parts(ctx, d_W, 1, m, AH, 1, m, m, k, &dparts, &ndp));
    }
    if (desc->compute_error) RC(nmf_relerr_parts(ctx, normA2, dparts, ndp, WtW2, HtH, k, rel + it));
    if (algo == PLANC_HALS) {
      // normalize_by_W: the column norms are the square roots of diag(W^T W);
      // the next W^T W is D^-1 (W^T W) D^-1.
      RC(gram_norms(ctx, WtW2, k, wnorms));
      RC(scale_cols(ctx, d_W, m, m, d_H, n, n, k, wnorms));
      RC(gram_rescale(ctx, WtW2, k, wnorms, WtW));
    } else {
      std::swap(WtW, WtW2);  // this W^T W serves the next H half-step
    }
    wtw_valid = true;
  }
  if (algo == PLANC_MU) RC(planc_normalize_by_w(ctx, d_W, m, m, d_H, n, n, k));
#undef RC
  if (h_relerr && desc->compute_error && iters > 0)
    PLANC_HIP_CHECK(ctx, hipMemcpyAsync(h_relerr, rel, sizeof(double) * iters, hipMemcpyDeviceToHost, ctx->stream));
  PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (algo == PLANC_HALS) return hals_sync_check(ctx);
  return PLANC_OK;
}

}  // namespace

extern "C" {

const char *planc_version(void) { return "planc_amd 0.1 (gfx950)"; }

int planc_create(int device, planc_ctx **out) {
  if (!out) return PLANC_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return PLANC_EHIP;
  if (hipSetDevice(device) != hipSuccess) return PLANC_EHIP;
  planc_ctx *c = new planc_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return PLANC_EHIP;
  }
  c->own_stream = true;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
    c->num_cu = prop.multiProcessorCount;
  // PLANC_THIN_PLANES=2: every 16-bit product carries its thin operand as two
  // 16-bit planes (NTF MTTKRPs and ANLS-BPP always do); unset or 1: one plane
  // for NMF-MU / HALS and the GEMM entry points.
  const char *tp = getenv("PLANC_THIN_PLANES");
  if (tp && tp[0] == '2') c->thin_planes = 2;
  // Locates this device's copy of the allocation-free CP-loss partials; a
  // failure is reported by planc_cp_loss_fwd_noalloc, not here.
  (void)cp_noalloc_setup(device);
  *out = c;
  return PLANC_OK;
}

int planc_destroy(planc_ctx *ctx) {
  if (!ctx) return PLANC_EINVAL;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &g : ctx->graphs) (void)hipGraphExecDestroy(g.exec);
  if (ctx->capture_stream) (void)hipStreamDestroy(ctx->capture_stream);
  for (auto &w : ctx->ws)
    if (w.ptr) (void)hipFree(w.ptr);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return PLANC_OK;
}

int planc_sync(planc_ctx *ctx) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return PLANC_OK;
}

const char *planc_last_error(planc_ctx *ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int planc_set_stream(planc_ctx *ctx, void *stream) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
  ctx->stream = (hipStream_t)stream;
  ctx->own_stream = false;
  return PLANC_OK;
}

void *planc_get_stream(planc_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int planc_malloc(planc_ctx *ctx, void **d_ptr, size_t bytes) {
  if (check_ctx(ctx) || !d_ptr) return PLANC_EINVAL;
  if (hipMalloc(d_ptr, bytes ? bytes : 1) != hipSuccess) return set_error(ctx, PLANC_ENOMEM, "planc_malloc");
  return PLANC_OK;
}

int planc_free(planc_ctx *ctx, void *d_ptr) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  PLANC_HIP_CHECK(ctx, hipFree(d_ptr));
  return PLANC_OK;
}

int planc_memcpy_h2d(planc_ctx *ctx, void *d_dst, const void *h_src, size_t bytes) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  PLANC_HIP_CHECK(ctx, hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
  PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return PLANC_OK;
}

int planc_memcpy_d2h(planc_ctx *ctx, void *h_dst, const void *d_src, size_t bytes) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  PLANC_HIP_CHECK(ctx, hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return PLANC_OK;
}

int planc_convert(planc_ctx *ctx, int dtype, const float *d_src, long rows, long cols, long ld_src,
                  void *d_dst, long ld_dst, float *d_colscale) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if ((dtype != PLANC_BF16 && dtype != PLANC_F16) || rows < 0 || cols < 0 || ld_src < rows ||
      ld_dst < rows || (rows * cols > 0 && (!d_src || !d_dst)))
    return set_error(ctx, PLANC_EINVAL, "planc_convert: bad arguments");
  return convert16(ctx, dtype, d_src, rows, cols, ld_src, d_dst, ld_dst, d_colscale);
}

int planc_transpose16(planc_ctx *ctx, const void *d_in, long rows, long cols, long ldi, void *d_out,
                      long ldo) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (rows < 0 || cols < 0 || ldi < rows || ldo < cols || (rows * cols > 0 && (!d_in || !d_out)))
    return set_error(ctx, PLANC_EINVAL, "planc_transpose16: bad arguments");
  return transpose16(ctx, d_in, rows, cols, ldi, d_out, ldo);
}

int planc_gemm_nn(planc_ctx *ctx, int dtype, const void *d_A, long lda, long L, long K,
                  const float *d_colscale, const float *d_B, long ldb, int k, float *d_C, long ldc) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (!valid_dtype(dtype) || lda < L || ldb < K || ldc < L || k <= 0)
    return set_error(ctx, PLANC_EINVAL, "planc_gemm_nn: bad arguments");
  return thin_gemm(ctx, dtype, 0, d_A, lda, L, K, d_B, ldb, nullptr, d_colscale, nullptr, k, d_C, ldc, 0,
                   0);
}

int planc_gemm_tn(planc_ctx *ctx, int dtype, const void *d_A, long lda, long K, long L,
                  const float *d_colscale, const float *d_B, long ldb, int k, float *d_C, long ldc) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (!valid_dtype(dtype) || lda < K || ldb < K || ldc < k || k <= 0)
    return set_error(ctx, PLANC_EINVAL, "planc_gemm_tn: bad arguments");
  return thin_gemm(ctx, dtype, 1, d_A, lda, L, K, d_B, ldb, nullptr, nullptr, d_colscale, k, d_C, ldc, 1,
                   0);
}

int planc_read_bandwidth(planc_ctx *ctx, const void *d_buf, size_t bytes, int reps, double *gbps) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (!d_buf || !gbps || bytes < (16u << 20) || reps < 1 || ((uintptr_t)d_buf % 16) != 0)
    return set_error(ctx, PLANC_EINVAL, "planc_read_bandwidth: bad arguments");
  return read_bandwidth(ctx, d_buf, bytes, reps, gbps);
}

int planc_gram(planc_ctx *ctx, const float *d_X, long rows, int k, long ldx, double *d_G) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (ldx < rows || !d_G) return set_error(ctx, PLANC_EINVAL, "planc_gram: bad arguments");
  return gram(ctx, d_X, rows, k, ldx, d_G);
}

int planc_mu_update(planc_ctx *ctx, float *d_H, long n, int k, long ldh, const float *d_N, long n_rs,
                    long n_cs, const double *d_G, double eps) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (ldh < n || k <= 0) return set_error(ctx, PLANC_EINVAL, "planc_mu_update: bad arguments");
  return row_update(ctx, 0, d_H, n, k, ldh, d_N, n_rs, n_cs, d_G, (float)eps);
}

int planc_hals_update_rows(planc_ctx *ctx, float *d_H, long n, int k, long ldh, const float *d_N,
                           long n_rs, long n_cs, const double *d_G) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (ldh < n || k <= 0) return set_error(ctx, PLANC_EINVAL, "planc_hals_update_rows: bad arguments");
  return row_update(ctx, 1, d_H, n, k, ldh, d_N, n_rs, n_cs, d_G, 0.0f);
}

int planc_hals_update_cols(planc_ctx *ctx, float *d_W, long m, int k, long ldw, const float *d_AH,
                           long ldah, const double *d_G) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (ldw < m || ldah < m || k <= 0) return set_error(ctx, PLANC_EINVAL, "planc_hals_update_cols: bad arguments");
  int rc = hals_sync_reset(ctx);
  if (rc) return rc;
  rc = hals_cols(ctx, d_W, m, k, ldw, d_AH, ldah, d_G);
  if (rc) return rc;
  return hals_sync_check(ctx);
}

int planc_normalize_by_w(planc_ctx *ctx, float *d_W, long m, long ldw, float *d_H, long n, long ldh,
                         int k) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (ldw < m || ldh < n || k <= 0) return set_error(ctx, PLANC_EINVAL, "planc_normalize_by_w: bad arguments");
  double *norms = (double *)ws_get(ctx, 9, sizeof(double) * k);
  if (!norms) return set_error(ctx, PLANC_ENOMEM, "normalize_by_w: workspace");
  int rc = col_norms(ctx, d_W, m, k, ldw, norms);
  if (rc) return rc;
  return scale_cols(ctx, d_W, m, ldw, d_H, n, ldh, k, norms);
}

int planc_normalize_cols(planc_ctx *ctx, float *d_A, long rows, int k, long lda, double *d_lambda) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (lda < rows || k <= 0 || !d_lambda) return set_error(ctx, PLANC_EINVAL, "planc_normalize_cols: bad arguments");
  int rc = col_norms(ctx, d_A, rows, k, lda, d_lambda);
  if (rc) return rc;
  return scale_cols(ctx, d_A, rows, lda, nullptr, 0, 0, k, d_lambda);
}

int planc_hadamard_grams(planc_ctx *ctx, const double *const *d_grams, int nmodes, int skip, int k,
                         double *d_out) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (!d_grams || k <= 0) return set_error(ctx, PLANC_EINVAL, "planc_hadamard_grams: bad arguments");
  return hadamard_grams(ctx, d_grams, nmodes, skip, k, d_out);
}

int planc_rel_error_gram(planc_ctx *ctx, const double *d_normA2, const double *d_trace, const double *d_WtW,
                         const double *d_HtH, int k, double *d_rel) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (!d_normA2 || !d_trace || !d_WtW || !d_HtH || !d_rel || k <= 0)
    return set_error(ctx, PLANC_EINVAL, "planc_rel_error_gram: bad arguments");
  PLANC_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  return nmf_relerr(ctx, d_normA2, d_trace, d_WtW, d_HtH, k, d_rel, nullptr);
}

int planc_mttkrp(planc_ctx *ctx, int dtype, const void *d_X, int nmodes, const long *dims,
                 const float *const *d_factors, int k, int mode, float *d_out, long ldo) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (!valid_dtype(dtype) || nmodes < 2 || nmodes > 7 || !dims || !d_factors || mode < 0 ||
      mode >= nmodes || k <= 0 || ldo < k)
    return set_error(ctx, PLANC_EINVAL, "planc_mttkrp: bad arguments");
  for (int i = 0; i < nmodes; ++i)
    if (dims[i] <= 0) return set_error(ctx, PLANC_EINVAL, "planc_mttkrp: empty mode");
  ThinPlanes planes(ctx, dtype != PLANC_F32 ? 2 : 1);  // as in planc_ntf_run
  float *P = nullptr;
  if (mode > 0 && mode < nmodes - 1) {
    long Lp = 1;
    for (int i = 0; i <= mode; ++i) Lp *= dims[i];
    P = (float *)ws_get(ctx, 4, sizeof(float) * (size_t)Lp * k);
    if (!P) return set_error(ctx, PLANC_ENOMEM, "planc_mttkrp: workspace");
  }
  return mttkrp_direct(ctx, dtype, d_X, nmodes, dims, d_factors, k, mode, d_out, ldo, P);
}

int planc_mttkrp_partial(planc_ctx *ctx, int dtype, const void *d_X, int nmodes, const long *dims,
                         const float *const *d_factors, int k, int split, int side, float *d_out,
                         long ldo) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (!valid_dtype(dtype) || nmodes < 2 || nmodes > 7 || !dims || !d_factors || split < 0 ||
      split > nmodes - 2 || (side != 0 && side != 1) || k <= 0 || !d_X || !d_out)
    return set_error(ctx, PLANC_EINVAL, "planc_mttkrp_partial: bad arguments");
  ThinPlanes planes(ctx, dtype != PLANC_F32 ? 2 : 1);  // as in planc_ntf_run
  long L = 1, R = 1;
  for (int i = 0; i < nmodes; ++i) {
    if (dims[i] <= 0) return set_error(ctx, PLANC_EINVAL, "planc_mttkrp_partial: empty mode");
    if (i <= split)
      L *= dims[i];
    else
      R *= dims[i];
  }
  KrpDesc kd{};
  if (side == 0) {
    if (ldo < L) return set_error(ctx, PLANC_EINVAL, "planc_mttkrp_partial: ldo < L");
    kd.nf = nmodes - 1 - split;
    for (int q = split + 1; q < nmodes; ++q) {
      kd.f[q - split - 1] = d_factors[q];
      kd.dim[q - split - 1] = dims[q];
      kd.ld[q - split - 1] = dims[q];
    }
    return thin_gemm(ctx, dtype, 0, d_X, L, L, R, nullptr, 0, &kd, nullptr, nullptr, k, d_out, ldo, 0, 0);
  }
  if (ldo < R) return set_error(ctx, PLANC_EINVAL, "planc_mttkrp_partial: ldo < R");
  kd.nf = split + 1;
  for (int q = 0; q <= split; ++q) {
    kd.f[q] = d_factors[q];
    kd.dim[q] = dims[q];
    kd.ld[q] = dims[q];
  }
  return thin_gemm(ctx, dtype, 1, d_X, L, R, L, nullptr, 0, &kd, nullptr, nullptr, k, d_out, ldo, 0, 0);
}

int planc_multi_ttv(planc_ctx *ctx, const float *d_P, long ldp, int nb, const long *dims,
                    const float *const *d_factors, int target, int k, float *d_out, long ldo) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (!d_P || nb < 1 || nb > 7 || !dims || !d_factors || target < 0 || target >= nb || k <= 0 ||
      ldo < k)
    return set_error(ctx, PLANC_EINVAL, "planc_multi_ttv: bad arguments");
  long E = 1, ldf[8];
  for (int i = 0; i < nb; ++i) {
    if (dims[i] <= 0) return set_error(ctx, PLANC_EINVAL, "planc_multi_ttv: empty mode");
    E *= dims[i];
    ldf[i] = dims[i];
  }
  if (ldp < E) return set_error(ctx, PLANC_EINVAL, "planc_multi_ttv: ldp < rows of P");
  return contract_projection(ctx, d_P, ldp, nb, dims, d_factors, ldf, target, k, d_out, ldo);
}

// ---------------------------------------------------------------------------
int planc_nmf_run(planc_ctx *ctx, const planc_nmf_desc *desc, const float *d_A, long lda,
                  const void *d_A16, const float *d_colscale, float *d_W, float *d_H,
                  double *h_relerr) {
  if (check_ctx(ctx) || !desc) return PLANC_EINVAL;
  const long m = desc->m, n = desc->n;
  const int k = desc->k, iters = desc->iters, algo = desc->algo, dt = desc->dtype;
  if (m <= 0 || n <= 0 || k <= 0 || k > 64 || iters < 0 || lda < m || !d_A || !d_W || !d_H ||
      !valid_dtype(dt))
    return set_error(ctx, PLANC_EINVAL, "planc_nmf_run: bad arguments (k must be 1..64)");
  if (algo != PLANC_MU && algo != PLANC_HALS && algo != PLANC_ANLSBPP && algo != PLANC_AOADMM &&
      algo != PLANC_GNSYM)
    return set_error(ctx, PLANC_EUNSUPPORTED, "planc_nmf_run: algorithm not supported");
  PLANC_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const void *Aop = d_A;
  const float *cs = nullptr;
  long ldop = lda;
  if (dt != PLANC_F32) {
    if (d_A16) {
      Aop = d_A16;
      cs = d_colscale;
    } else {
      void *a16 = ws_get(ctx, 7, sizeof(uint16_t) * (size_t)m * n);
      float *sc = (float *)ws_get(ctx, 8, sizeof(float) * (size_t)n);
      if (!a16 || !sc) return set_error(ctx, PLANC_ENOMEM, "planc_nmf_run: 16-bit copy");
      int rc = convert16(ctx, dt, d_A, m, n, lda, a16, m, sc);
      if (rc) return rc;
      Aop = a16;
      cs = sc;
      ldop = m;
    }
  }
  // 16-bit: a transposed copy makes A*H a TN product too (its contraction
  // index n contiguous), so both products stream with 16-byte MFMA operands.
  const void *AopT = nullptr;
  if (dt != PLANC_F32 && desc->A16T) {
    AopT = desc->A16T;
  } else if (dt != PLANC_F32) {
    void *at = ws_get(ctx, 16, sizeof(uint16_t) * (size_t)m * n);
    if (!at) return set_error(ctx, PLANC_ENOMEM, "planc_nmf_run: transposed 16-bit copy");
    int rc = transpose16(ctx, Aop, m, n, ldop, at, n);
    if (rc) return rc;
    AopT = at;
  }
  // A*F (m x k, ld m) for F n x k.
  auto a_times_h = [&](const float *F, float *out) -> int {
    if (AopT) return thin_gemm(ctx, dt, 1, AopT, n, m, n, F, n, nullptr, cs, nullptr, k, out, m, 0, 0);
    return thin_gemm(ctx, dt, 0, Aop, ldop, m, n, F, n, nullptr, cs, nullptr, k, out, m, 0, 0);
  };
  NMFProducts prod;
  prod.wta = [&](const float *F, float *out) -> int {
    return thin_gemm(ctx, dt, 1, Aop, ldop, n, m, F, m, nullptr, nullptr, cs, k, out, k, 1, 0);
  };
  prod.ah = a_times_h;
  // ANLS-BPP solves normal equations whose Gram comes from the unrounded
  // factor; a right-hand side from the 16-bit-rounded factor would make them
  // inconsistent, an error that the solve amplifies by cond(Gram).  The
  // factor therefore enters the 16-bit products as two 16-bit planes (its
  // rounding and the rounded remainder): only the rounding of A remains.
  ThinPlanes planes(ctx, dt != PLANC_F32 && algo == PLANC_ANLSBPP ? 2 : ctx->thin_planes);
  prod.norma2 = [&](double *out) -> int { return dot(ctx, d_A, 1, lda, d_A, 1, lda, m, n, out, 0); };
  return nmf_loop(ctx, desc, prod, d_W, d_H, h_relerr);
}

// ---- sparse input -----------------------------------------------------------
namespace {
int read_nnz(planc_ctx *ctx, const long *d_indptr, long m, long *nnz) {
  PLANC_HIP_CHECK(ctx, hipMemcpyAsync(nnz, d_indptr + m, sizeof(long), hipMemcpyDeviceToHost, ctx->stream));
  PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (*nnz < 0) return set_error(ctx, PLANC_EINVAL, "sparse: negative entry count");
  return PLANC_OK;
}
}  // namespace

int planc_csr_transpose(planc_ctx *ctx, long m, long n, const long *d_indptr, const int *d_indices,
                        const float *d_values, long *d_out_indptr, int *d_out_indices, float *d_out_values) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (m < 0 || n < 0 || !d_indptr || !d_out_indptr)
    return set_error(ctx, PLANC_EINVAL, "planc_csr_transpose: bad arguments");
  long nnz = 0;
  int rc = read_nnz(ctx, d_indptr, m, &nnz);
  if (rc) return rc;
  if (nnz > 0 && (!d_indices || !d_values || !d_out_indices || !d_out_values))
    return set_error(ctx, PLANC_EINVAL, "planc_csr_transpose: null entry arrays");
  if ((rc = csr_transpose(ctx, m, n, nnz, d_indptr, d_indices, d_values, d_out_indptr, d_out_indices,
                          d_out_values)))
    return rc;
  PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return PLANC_OK;
}

int planc_spmm(planc_ctx *ctx, long rows, long cols, const long *d_indptr, const int *d_indices,
               const float *d_values, const float *d_F, long ldf, int k, float *d_out, long ldo) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (rows < 0 || cols < 0 || k <= 0 || k > 64 || ldf < cols || ldo < rows || !d_indptr || !d_F || !d_out)
    return set_error(ctx, PLANC_EINVAL, "planc_spmm: bad arguments (k must be 1..64)");
  return spmm_colmajor(ctx, rows, cols, d_indptr, d_indices, d_values, d_F, ldf, k, d_out, ldo);
}

int planc_sparse_fit_error(planc_ctx *ctx, long m, long n, const long *d_indptr, const int *d_indices,
                           const float *d_values, const float *d_W, long ldw, const float *d_H, long ldh,
                           int k, double *h_err) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (m <= 0 || n <= 0 || k <= 0 || k > 64 || ldw < m || ldh < n || !d_indptr || !d_W || !d_H || !h_err)
    return set_error(ctx, PLANC_EINVAL, "planc_sparse_fit_error: bad arguments");
  double *d = (double *)ws_get(ctx, 12, sizeof(double));
  if (!d) return set_error(ctx, PLANC_ENOMEM, "planc_sparse_fit_error: workspace");
  int rc = sparse_fit_error(ctx, m, n, d_indptr, d_indices, d_values, d_W, ldw, d_H, ldh, k, d);
  if (rc) return rc;
  PLANC_HIP_CHECK(ctx, hipMemcpyAsync(h_err, d, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return PLANC_OK;
}

int planc_nmf_sparse_run(planc_ctx *ctx, const planc_nmf_desc *desc, const long *d_indptr,
                         const int *d_indices, const float *d_values, float *d_W, float *d_H,
                         double *h_relerr) {
  if (check_ctx(ctx) || !desc) return PLANC_EINVAL;
  const long m = desc->m, n = desc->n;
  const int k = desc->k, iters = desc->iters, algo = desc->algo;
  if (m <= 0 || n <= 0 || k <= 0 || k > 64 || iters < 0 || !d_indptr || !d_W || !d_H)
    return set_error(ctx, PLANC_EINVAL, "planc_nmf_sparse_run: bad arguments (k must be 1..64)");
  if (algo != PLANC_MU && algo != PLANC_HALS && algo != PLANC_ANLSBPP && algo != PLANC_AOADMM &&
      algo != PLANC_GNSYM)
    return set_error(ctx, PLANC_EUNSUPPORTED, "planc_nmf_sparse_run: algorithm not supported");
  if (desc->dtype != PLANC_F32)
    return set_error(ctx, PLANC_EUNSUPPORTED, "planc_nmf_sparse_run: sparse products are FP32 only");
  PLANC_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  long nnz = 0;
  int rc = read_nnz(ctx, d_indptr, m, &nnz);
  if (rc) return rc;
  if (nnz > 0 && (!d_indices || !d_values))
    return set_error(ctx, PLANC_EINVAL, "planc_nmf_sparse_run: null entry arrays");
  SparseOps ops;
  if ((rc = sparse_products(ctx, m, n, nnz, d_indptr, d_indices, d_values, &ops))) return rc;
  NMFProducts prod;
  prod.wta = [&](const float *F, float *out) -> int { return sparse_wta(ctx, ops, m, n, F, k, out); };
  prod.ah = [&](const float *F, float *out) -> int {
    return sparse_ah(ctx, ops, m, n, d_indptr, d_indices, d_values, F, k, out);
  };
  prod.norma2 = [&](double *out) -> int {
    if (nnz == 0) {
      PLANC_HIP_CHECK(ctx, hipMemsetAsync(out, 0, sizeof(double), ctx->stream));
      return PLANC_OK;
    }
    return dot(ctx, d_values, 1, nnz, d_values, 1, nnz, nnz, 1, out, 0);
  };
  prod.bpp_prenormalize = true;
  return nmf_loop(ctx, desc, prod, d_W, d_H, h_relerr);
}

int planc_ntf_tree_split(int nmodes, const long *dims) {
  if (nmodes < 2 || nmodes > 7 || !dims) return -1;
  for (int i = 0; i < nmodes; ++i)
    if (dims[i] <= 0) return -1;
  return balanced_split(nmodes, dims);
}

int planc_ntf_run(planc_ctx *ctx, const planc_ntf_desc *desc, const float *d_X, const void *d_X16,
                  float *const *d_factors, double *d_lambda, double *h_relerr) {
  if (check_ctx(ctx) || !desc) return PLANC_EINVAL;
  const int N = desc->nmodes, k = desc->k, iters = desc->iters, algo = desc->algo, dt = desc->dtype;
  if (N < 2 || N > 7 || k <= 0 || k > 64 || iters < 0 || !d_X || !d_factors || !d_lambda ||
      !valid_dtype(dt))
    return set_error(ctx, PLANC_EINVAL, "planc_ntf_run: bad arguments (2..7 modes, k 1..64)");
  if (algo != PLANC_MU && algo != PLANC_HALS && algo != PLANC_ANLSBPP && algo != PLANC_AOADMM &&
      algo != PLANC_NESTEROV)
    return set_error(ctx, PLANC_EUNSUPPORTED, "planc_ntf_run: algorithm not supported");
  PLANC_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  // 16-bit MTTKRPs: the factors / Khatri-Rao rows enter as two 16-bit planes,
  // so only X's rounding remains (a one-plane KRP error is amplified by the
  // mode updates several times over).
  ThinPlanes planes(ctx, dt != PLANC_F32 ? 2 : 1);
  const long *dims = desc->dims;
  long numel = 1, Imax = 0;
  for (int i = 0; i < N; ++i) {
    if (dims[i] <= 0) return set_error(ctx, PLANC_EINVAL, "planc_ntf_run: empty mode");
    numel *= dims[i];
    Imax = std::max(Imax, dims[i]);
  }
  const bool dimtree = desc->dimtree && N >= 3;
  const int s = dimtree ? balanced_split(N, dims) : 0;
  long L = 1, R = 1;
  for (int i = 0; i <= s; ++i) L *= dims[i];
  for (int i = s + 1; i < N; ++i) R *= dims[i];
  const void *Xop = d_X;
  // Per-column power-of-two scales s_j of the 16-bit copy, columns of the
  // dimension tree's L x R unfolding (the copy made here; a caller's copy is
  // unscaled).  The left contraction takes s_j as scales of its output
  // columns (FP32).  The right one contracts over j: its Khatri-Rao rows take
  // r_j = s_j / S (S = max_j s_j, so the 16-bit operand stays at most 1 and
  // within FP16's range) and its FP32 output is multiplied by S.
  const float *xscale = nullptr, *xrscale = nullptr;
  float xS = 1.0f;
  if (dt != PLANC_F32) {
    if (d_X16) {
      Xop = d_X16;
    } else {
      void *x16 = ws_get(ctx, 7, sizeof(uint16_t) * (size_t)numel);
      if (!x16) return set_error(ctx, PLANC_ENOMEM, "planc_ntf_run: 16-bit copy");
      int rc;
      if (dimtree) {
        float *sc = (float *)ws_get(ctx, 8, sizeof(float) * 2 * (size_t)R);
        if (!sc) return set_error(ctx, PLANC_ENOMEM, "planc_ntf_run: 16-bit scales");
        rc = convert16(ctx, dt, d_X, L, R, L, x16, L, sc);
        if (rc) return rc;
        std::vector<float> hs((size_t)R);
        PLANC_HIP_CHECK(ctx, hipMemcpyAsync(hs.data(), sc, sizeof(float) * (size_t)R, hipMemcpyDeviceToHost,
                                            ctx->stream));
        PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        for (float v : hs) xS = std::max(xS, v);
        for (float &v : hs) v /= xS;  // powers of two: exact
        PLANC_HIP_CHECK(ctx, hipMemcpyAsync(sc + R, hs.data(), sizeof(float) * (size_t)R, hipMemcpyHostToDevice,
                                            ctx->stream));
        PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        xscale = sc;
        xrscale = sc + R;
      } else {
        rc = convert16(ctx, dt, d_X, numel, 1, numel, x16, numel, nullptr);
      }
      if (rc) return rc;
      Xop = x16;
    }
  }
  // 16-bit dimension tree: both unfoldings, X_(LxR) and its transpose
  // (R x L, the same values and scales), so that each contraction runs on the
  // streaming 16-bit NN kernel where its rows come in 16-byte groups: the
  // right one X_(LxR) KRP_R on X when L % 8 == 0, else as a TN product on the
  // transpose; the left one X_(LxR)^T KRP_L on the transpose when R % 8 == 0,
  // else as a TN product on X.  The run makes the transpose with its own
  // copy; with a caller's copy it is desc->X16T, or made here when the right
  // contraction needs it.
  const bool nnR = L % 8 == 0 && (uintptr_t)Xop % 16 == 0;
  const void *XopT = nullptr;
  if (dt != PLANC_F32 && dimtree) {
    if (d_X16 && desc->X16T) {
      XopT = desc->X16T;
    } else if (!d_X16 || !nnR) {
      void *xt = ws_get(ctx, 16, sizeof(uint16_t) * (size_t)numel);
      if (!xt) return set_error(ctx, PLANC_ENOMEM, "planc_ntf_run: transposed 16-bit copy");
      int rc = transpose16(ctx, Xop, L, R, L, xt, R);
      if (rc) return rc;
      XopT = xt;
    }
  }
  const bool nnQ = XopT && R % 8 == 0 && (uintptr_t)XopT % 16 == 0;
  // Float scratch: P (L x k) and Q (R x k) for the tree (or the direct
  // middle-mode projection), M (k x Imax).
  long Pmax = 0;
  if (dimtree) {
    Pmax = L * k + R * k;
  } else {
    long Lp = 1;
    for (int i = 0; i < N - 1; ++i) {
      Lp *= dims[i];
      if (i > 0) Pmax = std::max(Pmax, Lp * k);
    }
  }
  float *fbuf = (float *)ws_get(ctx, 4, sizeof(float) * ((size_t)Pmax + (size_t)k * Imax));
  const size_t kk = (size_t)k * k;
  double *dbuf = (double *)ws_get(ctx, 5, sizeof(double) * ((N + 2) * kk + 8 + k + (size_t)iters + 1));
  if (!fbuf || !dbuf) return set_error(ctx, PLANC_ENOMEM, "planc_ntf_run: workspace");
  float *P = fbuf, *Q = dimtree ? fbuf + (size_t)L * k : nullptr;
  float *M = fbuf + Pmax;
  std::vector<double *> G(N);
  for (int i = 0; i < N; ++i) G[i] = dbuf + i * kk;
  double *S = dbuf + N * kk;
  double *normX2 = dbuf + (N + 1) * kk;
  double *norms = normX2 + 8;
  double *rel = norms + k;
  int rc;
#define RC(x)          \
  do {                 \
    rc = (x);          \
    if (rc) return rc; \
  } while (0)
  int *dflag = (int *)ws_get(ctx, 11, sizeof(int));
  if (!dflag) return set_error(ctx, PLANC_ENOMEM, "planc_ntf_run: workspace");
  PLANC_HIP_CHECK(ctx, hipMemsetAsync(dflag, 0, sizeof(int), ctx->stream));
  // NTFAOADMM's auxiliary (dual) factors, zero at construction
  // (ntf/ntfaoadmm.hpp:84-86), one I_n x k block per mode, and a norm scratch.
  std::vector<float *> duals(N, nullptr);
  double *dnorm = nullptr;
  if (algo == PLANC_AOADMM) {
    long tot = 0;
    for (int i = 0; i < N; ++i) tot += dims[i] * k;
    float *db = (float *)ws_get(ctx, 32, sizeof(float) * (size_t)tot);
    dnorm = (double *)ws_get(ctx, 34, sizeof(double) * k);
    if (!db || !dnorm) return set_error(ctx, PLANC_ENOMEM, "planc_ntf_run: AO-ADMM duals");
    PLANC_HIP_CHECK(ctx, hipMemsetAsync(db, 0, sizeof(float) * (size_t)tot, ctx->stream));
    for (int i = 0; i < N; ++i) {
      duals[i] = db;
      db += dims[i] * k;
    }
  }
  // NTFNES state (ntf/ntfnes.hpp:9-188): the proximal centres and previous
  // iterates (zero at construction, :175-176, held here as I_n x k), the
  // accelerated factors, a lambda-scaled copy of mode 0 for the trial error,
  // and the lambdas (ones at construction).
  const bool nes = algo == PLANC_NESTEROV;
  std::vector<float *> nprox(N), nprev(N), nacc(N);
  float *ntmp0 = nullptr;
  double *nl = nullptr;  // prox_l, prev_l, acc_l (k each), trial loss
  if (nes) {
    long tot = 0;
    for (int i = 0; i < N; ++i) tot += dims[i] * k;
    float *b = (float *)ws_get(ctx, 36, sizeof(float) * (size_t)(3 * tot + dims[0] * k));
    nl = (double *)ws_get(ctx, 37, sizeof(double) * (3 * (size_t)k + 1));
    if (!b || !nl) return set_error(ctx, PLANC_ENOMEM, "planc_ntf_run: NES state");
    PLANC_HIP_CHECK(ctx, hipMemsetAsync(b, 0, sizeof(float) * (size_t)(3 * tot), ctx->stream));
    for (int i = 0; i < N; ++i) {
      nprox[i] = b;
      nprev[i] = b + tot;
      nacc[i] = b + 2 * tot;
      b += dims[i] * k;
    }
    ntmp0 = nacc[N - 1] + dims[N - 1] * k;
    std::vector<double> ones(2 * (size_t)k, 1.0);
    PLANC_HIP_CHECK(ctx, hipMemcpyAsync(nl, ones.data(), sizeof(double) * 2 * k, hipMemcpyHostToDevice, ctx->stream));
  }
  if (desc->compute_error || nes) {
    if (desc->normX2 > 0)
      PLANC_HIP_CHECK(ctx, hipMemcpyAsync(normX2, &desc->normX2, sizeof(double), hipMemcpyHostToDevice,
                                          ctx->stream));
    else
      RC(dot(ctx, d_X, 1, numel, d_X, 1, numel, numel, 1, normX2, 0));
  }
  // NCPFactors::normalize() of the initial factors (ntf/auntf.hpp:88) and
  // their Grams: small factors in one launch for all modes.
  if (init_normalize_fits(dims, N, k)) {
    RC(init_normalize(ctx, d_factors, dims, N, k, G.data(), d_lambda));
  } else {
    for (int i = 0; i < N; ++i) {
      RC(col_norms(ctx, d_factors[i], dims[i], k, dims[i], norms));
      RC(scale_cols(ctx, d_factors[i], dims[i], dims[i], nullptr, 0, 0, k, norms));
      RC(lambda_mul(ctx, d_lambda, norms, k, i == 0));
      RC(gram(ctx, d_factors[i], dims[i], k, dims[i], G[i]));
    }
  }
  std::vector<const double *> Gc(G.begin(), G.end());
  long ldfL[8], ldfR[8];
  for (int i = 0; i < N; ++i) {
    ldfL[i] = dims[i];
    ldfR[i] = dims[i];
  }
  // The dimension tree's two contractions: P = X_(LxR) KRP_R (right) and
  // Q = X_(LxR)^T KRP_L (left), each in one of its two forms.
  auto right_contraction = [&](bool tn_on_t) -> int {
    KrpDesc kd{};
    kd.nf = N - 1 - s;
    for (int q = s + 1; q < N; ++q) {
      kd.f[q - s - 1] = d_factors[q];
      kd.dim[q - s - 1] = dims[q];
      kd.ld[q - s - 1] = dims[q];
    }
    if (tn_on_t) return thin_gemm(ctx, dt, 1, XopT, R, L, R, nullptr, 0, &kd, xrscale, nullptr, k, P, L, 0, 0, xS);
    return thin_gemm(ctx, dt, 0, Xop, L, L, R, nullptr, 0, &kd, xrscale, nullptr, k, P, L, 0, 0, xS);
  };
  auto left_contraction = [&](bool nn_on_t) -> int {
    KrpDesc kd{};
    kd.nf = s + 1;
    for (int q = 0; q <= s; ++q) {
      kd.f[q] = d_factors[q];
      kd.dim[q] = dims[q];
      kd.ld[q] = dims[q];
    }
    if (nn_on_t) return thin_gemm(ctx, dt, 0, XopT, R, R, L, nullptr, 0, &kd, nullptr, xscale, k, Q, R, 0, 0);
    return thin_gemm(ctx, dt, 1, Xop, L, R, L, nullptr, 0, &kd, nullptr, xscale, k, Q, R, 0, 0);
  };
  // With both unfoldings each contraction takes the faster of its feasible
  // forms on this shape (NN on X or TN on the transpose; TN on X or NN on
  // the transpose), timed once per context and shape.
  bool tnR = XopT && !nnR, nnL = nnQ;
  if (dimtree && dt != PLANC_F32 && XopT && (nnR || nnQ)) {
    std::vector<long> key = {dt, N, k, s, L, R, nnR ? 1 : 0, nnQ ? 1 : 0};
    for (int i = 0; i < N; ++i) key.push_back(dims[i]);
    int form = -1;
    for (const auto &e : ctx->tree_forms)
      if (e.first == key) form = e.second;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    PLANC_HIP_CHECK(ctx, hipStreamIsCapturing(ctx->stream, &cs));
    if (form < 0 && cs == hipStreamCaptureStatusNone) {
      hipEvent_t e0, e1;
      PLANC_HIP_CHECK(ctx, hipEventCreate(&e0));
      PLANC_HIP_CHECK(ctx, hipEventCreate(&e1));
      auto timed = [&](const std::function<int()> &f, float *ms) -> int {
        for (int r = 0; r < 2; ++r)  // geometry tuning and warm-up
          if (int e = f()) return e;
        PLANC_HIP_CHECK(ctx, hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < 3; ++r)
          if (int e = f()) return e;
        PLANC_HIP_CHECK(ctx, hipEventRecord(e1, ctx->stream));
        PLANC_HIP_CHECK(ctx, hipEventSynchronize(e1));
        PLANC_HIP_CHECK(ctx, hipEventElapsedTime(ms, e0, e1));
        return PLANC_OK;
      };
      form = 0;
      float a = 0, b = 0;
      if (nnR) {
        RC(timed([&] { return right_contraction(false); }, &a));
        RC(timed([&] { return right_contraction(true); }, &b));
        if (b < a) form |= 1;
      } else {
        form |= 1;
      }
      if (nnQ) {
        RC(timed([&] { return left_contraction(true); }, &a));
        RC(timed([&] { return left_contraction(false); }, &b));
        if (a <= b) form |= 2;
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      ctx->tree_forms.push_back({key, form});
    }
    if (form >= 0) {
      tnR = (form & 1) != 0;
      nnL = (form & 2) != 0;
    }
  }
  // One outer iteration (all mode updates + the error), written so the same
  // launch sequence can be captured into a hipGraph: no allocation, no host
  // synchronisation, the error written to *rel_out.
  // rel_counter null: the error goes to *rel_out; otherwise to
  // rel_out[*rel_counter] (device counter, incremented).
  double *relslot = rel + iters;  // one spare slot after rel[]
  auto one_iteration = [&](double *rel_out, int *rel_counter) -> int {
    int rc;
    bool err_done = false;
    for (int n = 0; n < N; ++n) {
      const bool fused = mode_update_fits(dims[n], k) && algo != PLANC_AOADMM && !nes;
      if (!fused || algo == PLANC_ANLSBPP) RC(hadamard_grams(ctx, Gc.data(), N, n, k, S));
      // This mode's MTTKRP: element (x, i) at Mp[i * mrs + x * mcs].
      const float *Mp = M;
      long mrs = k, mcs = 1;
      if (dimtree) {
        if (n == 0) RC(right_contraction(tnR));
        if (n == s + 1) RC(left_contraction(nnL));
        if (n <= s) {
          RC(contract_projection(ctx, P, L, s + 1, dims, d_factors, ldfL, n, k, M, k));
        } else if (fused && N - 1 - s == 1) {
          // A one-mode right block: the projection is the MTTKRP itself (R x k).
          Mp = Q;
          mrs = 1;
          mcs = R;
        } else {
          RC(contract_projection(ctx, Q, R, N - 1 - s, dims + s + 1, d_factors + s + 1, ldfR + s + 1,
                                 n - s - 1, k, M, k));
        }
      } else {
        RC(mttkrp_direct(ctx, dt, Xop, N, dims, d_factors, k, n, M, k, P));
      }
      float *Fn = d_factors[n];
      if (fused) {
        int kind = algo == PLANC_HALS ? 1 : 0;
        if (algo == PLANC_ANLSBPP) {
          RC(bpp_solve(ctx, S, k, Mp, mrs, mcs, dims[n], Fn, 1, dims[n], dflag));
          kind = 2;
        }
        const bool last_err = desc->compute_error && n == N - 1;
        RC(mode_update(ctx, kind, 1e-6f, Fn, dims[n], dims[n], k, Mp, mrs, mcs, Gc.data(), N, n, G[n],
                       d_lambda, normX2, last_err ? rel_out : nullptr, rel_counter));
        err_done = err_done || last_err;
        continue;
      }
      if (algo == PLANC_HALS) {
        RC(row_update(ctx, 1, Fn, dims[n], k, dims[n], M, k, 1, S, 0.0f));
      } else if (algo == PLANC_MU) {
        RC(row_update(ctx, 0, Fn, dims[n], k, dims[n], M, k, 1, S, 1e-6f));
      } else if (nes) {
        // NTFNES::update: the previous iterate (and at mode 0 its lambda)
        // saved, the proximal Nesterov loop from the centre, the centre
        // replaced by the result with unit columns.
        if (n == 0)
          PLANC_HIP_CHECK(ctx, hipMemcpyAsync(nl + k, nl, sizeof(double) * k, hipMemcpyDeviceToDevice, ctx->stream));
        PLANC_HIP_CHECK(ctx, hipMemcpyAsync(nprev[n], nprox[n], sizeof(float) * dims[n] * k,
                                            hipMemcpyDeviceToDevice, ctx->stream));
        RC(nes_mode_update(ctx, S, k, M, k, 1, nprox[n], dims[n], Fn, dims[n], nl, dflag));
      } else if (algo == PLANC_AOADMM) {
        // NTFAOADMM::update (ntf/ntfaoadmm.hpp:23-77): ADMM solve on S and
        // the MTTKRP with this mode's dual, then the dual's columns to unit
        // norm (m_ncp_aux.normalize(mode)).
        RC(admm_solve(ctx, S, k, M, k, 1, Fn, dims[n], duals[n], dims[n], dims[n], dflag));
        RC(col_norms(ctx, duals[n], dims[n], k, dims[n], dnorm));
        RC(scale_cols(ctx, duals[n], dims[n], dims[n], nullptr, 0, 0, k, dnorm));
      } else {
        RC(bpp_solve(ctx, S, k, M, k, 1, dims[n], Fn, 1, dims[n], dflag));
      }
      // update_factor_mode: normalize(mode) replaces lambda (ntf/auntf.hpp:66-79).
      RC(col_norms(ctx, Fn, dims[n], k, dims[n], d_lambda));
      RC(scale_cols(ctx, Fn, dims[n], dims[n], nullptr, 0, 0, k, d_lambda));
      RC(gram(ctx, Fn, dims[n], k, dims[n], G[n]));
    }
    if ((desc->compute_error || nes) && !err_done) {
      RC(ntf_relerr(ctx, normX2, d_factors[N - 1], dims[N - 1], dims[N - 1], M, k, k, d_lambda,
                    Gc.data(), N, rel_counter ? relslot : rel_out));
      if (rel_counter) RC(counter_append(ctx, relslot, rel_out, rel_counter));
    }
    return PLANC_OK;
  };
  if (nes) {
    // Eager iterations: accelerate() (ntf/ntfnes.hpp:68-117) decides on the
    // host between the accelerated factors and the proximal ones.
    double normX2_h = 0;
    PLANC_HIP_CHECK(ctx, hipMemcpyAsync(&normX2_h, normX2, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    int acc_exp = 3, acc_fails = 0;
    const int acc_start = 5, fail_limit = 5;
    double *prox_l = nl, *prev_l = nl + k, *acc_l = nl + 2 * k, *loss = nl + 3 * k;
    auto reset_to = [&](const std::vector<float *> &src, const double *lam) -> int {
      // reset(new_factors, true) (ntf/auntf.hpp:186-198)
      int rc2;
      for (int i = 0; i < N; ++i) {
        PLANC_HIP_CHECK(ctx, hipMemcpyAsync(d_factors[i], src[i], sizeof(float) * dims[i] * k,
                                            hipMemcpyDeviceToDevice, ctx->stream));
        if ((rc2 = col_norms(ctx, d_factors[i], dims[i], k, dims[i], d_lambda))) return rc2;
        if ((rc2 = scale_cols(ctx, d_factors[i], dims[i], dims[i], nullptr, 0, 0, k, d_lambda))) return rc2;
      }
      PLANC_HIP_CHECK(ctx, hipMemcpyAsync(d_lambda, lam, sizeof(double) * k, hipMemcpyDeviceToDevice, ctx->stream));
      return PLANC_OK;
    };
    for (int it = 0; it < iters; ++it) {
      RC(one_iteration(rel + it, nullptr));
      if (it <= acc_start) continue;
      double cur_err = 0;
      PLANC_HIP_CHECK(ctx, hipMemcpyAsync(&cur_err, rel + it, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
      const double step = std::pow((double)(it + 1), 1.0 / acc_exp);
      for (int i = 0; i < N; ++i) {
        const bool last = i == N - 1;  // only the last mode carries lambda
        RC(nes_acc_build(ctx, d_factors[i], nprev[i], dims[i], k, last ? d_lambda : nullptr,
                         last ? prev_l : nullptr, step, nacc[i]));
        RC(col_norms(ctx, nacc[i], dims[i], k, dims[i], acc_l));
        RC(scale_cols(ctx, nacc[i], dims[i], dims[i], nullptr, 0, 0, k, acc_l));
      }
      // computeObjectiveError(m_acc_t): the factors reset to the accelerated
      // ones, the error from the fused CP loss with lambda folded into mode 0.
      RC(reset_to(nacc, acc_l));
      RC(scale_cols_by(ctx, d_factors[0], dims[0], k, d_lambda, ntmp0));
      std::vector<const float *> fl(N);
      fl[0] = ntmp0;
      for (int i = 1; i < N; ++i) fl[i] = d_factors[i];
      RC(cp_loss_forward(ctx, d_X, N, dims, fl.data(), k, loss));
      double loss_h = 0;
      PLANC_HIP_CHECK(ctx, hipMemcpyAsync(&loss_h, loss, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
      const double acc_err = std::sqrt(2.0 * loss_h / normX2_h);
      if (acc_err < cur_err) {
        for (int i = 0; i < N; ++i) {
          PLANC_HIP_CHECK(ctx, hipMemcpyAsync(nprox[i], nacc[i], sizeof(float) * dims[i] * k,
                                              hipMemcpyDeviceToDevice, ctx->stream));
          RC(col_norms(ctx, nprox[i], dims[i], k, dims[i], prox_l));
          RC(scale_cols(ctx, nprox[i], dims[i], dims[i], nullptr, 0, 0, k, prox_l));
        }
      } else {
        RC(reset_to(nprox, prox_l));
        if (++acc_fails > fail_limit) {
          acc_fails = 0;
          acc_exp++;
        }
      }
      for (int i = 0; i < N; ++i) RC(gram(ctx, d_factors[i], dims[i], k, dims[i], G[i]));
    }
  } else if (iters > 0) {
    RC(one_iteration(rel, nullptr));
  }
  if (!nes && iters > 1) {
    // Iterations 1.. replay a captured graph of GU iterations (at small
    // shapes an iteration is a few tens of microseconds and every graph
    // launch costs several); the remainder runs eagerly.  The graph is
    // captured (not run) on the first call with iters > 1 whatever its
    // length, so later calls of any length find it cached.  The error goes
    // through a device counter into rel[1..].
    int *counter = (int *)ws_get(ctx, 15, sizeof(int));
    if (!counter) return set_error(ctx, PLANC_ENOMEM, "planc_ntf_run: workspace");
    RC(counter_set(ctx, counter, 1));
    const int GU = 8;
    const int rest = iters - 1, full = rest / GU, rem = rest % GU;
    auto graph_of = [&](int n_it, hipGraphExec_t *exec, hipGraphExec_t keep) -> int {
      std::vector<uintptr_t> key = {1, (uintptr_t)N, (uintptr_t)k, (uintptr_t)algo, (uintptr_t)dt,
                                    (uintptr_t)n_it, (uintptr_t)dimtree, (uintptr_t)desc->compute_error,
                                    (uintptr_t)Xop, (uintptr_t)XopT, (uintptr_t)d_lambda, (uintptr_t)rel};
      for (int i = 0; i < N; ++i) {
        key.push_back((uintptr_t)d_factors[i]);
        key.push_back((uintptr_t)dims[i]);
      }
      for (auto &w : ctx->ws) key.push_back((uintptr_t)w.ptr);
      const char *oe = getenv("PLANC_MODE_ONE_BLOCK");  // the mode-update form captured
      key.push_back(oe && oe[0] == '0' ? 0 : 1);
      uint32_t xs_bits;  // the 16-bit copy's output scale is a captured kernel argument
      std::memcpy(&xs_bits, &xS, sizeof xs_bits);
      key.push_back(xs_bits);
      if (graph_lookup(ctx, key, exec)) return PLANC_OK;
      int rc2 = capture_graph(ctx, [&]() -> int {
        for (int u = 0; u < n_it; ++u) {
          const int r = one_iteration(rel, counter);
          if (r) return r;
        }
        return PLANC_OK;
      }, exec);
      if (rc2) return rc2;
      graph_store(ctx, key, *exec, keep);
      return PLANC_OK;
    };
    hipGraphExec_t exec_full = nullptr;
    RC(graph_of(GU, &exec_full, nullptr));
    for (int it = 0; it < full; ++it) PLANC_HIP_CHECK(ctx, hipGraphLaunch(exec_full, ctx->stream));
    for (int it = 0; it < rem; ++it) RC(one_iteration(rel, counter));
  }
#undef RC
  int hflag = 0;
  PLANC_HIP_CHECK(ctx, hipMemcpyAsync(&hflag, dflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  if (h_relerr && desc->compute_error && iters > 0)
    PLANC_HIP_CHECK(ctx, hipMemcpyAsync(h_relerr, rel, sizeof(double) * iters, hipMemcpyDeviceToHost, ctx->stream));
  PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (hflag & 4) return set_error(ctx, PLANC_ENOCONV, "AO-ADMM: shifted Gram not positive definite");
  if (hflag & 8) return set_error(ctx, PLANC_ENOCONV, "NES: the Nesterov loop ran past 100000 steps");
  if (hflag) return set_error(ctx, PLANC_ENOCONV, "BPP failed: more than 5k pivot iterations");
  return PLANC_OK;
}

int planc_cp_loss_fwd(planc_ctx *ctx, const float *d_X, int nmodes, const long *dims,
                      const float *const *d_factors, int k, double *h_loss) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (!d_X || !dims || !d_factors || !h_loss || nmodes < 2 || nmodes > 7 || k <= 0 || k > 64)
    return set_error(ctx, PLANC_EINVAL, "planc_cp_loss_fwd: bad arguments");
  for (int i = 0; i < nmodes; ++i)
    if (dims[i] <= 0) return set_error(ctx, PLANC_EINVAL, "planc_cp_loss_fwd: empty mode");
  double *dl = (double *)ws_get(ctx, 12, sizeof(double));
  if (!dl) return set_error(ctx, PLANC_ENOMEM, "planc_cp_loss_fwd: workspace");
  int rc = cp_loss_forward(ctx, d_X, nmodes, dims, d_factors, k, dl);
  if (rc) return rc;
  PLANC_HIP_CHECK(ctx, hipMemcpyAsync(h_loss, dl, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return PLANC_OK;
}

int planc_cp_loss_fwd_noalloc(planc_ctx *ctx, const float *d_X, int nmodes, const long *dims,
                              const float *const *d_factors, int k, double *d_loss) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (!d_X || !dims || !d_factors || !d_loss || nmodes < 2 || nmodes > 7 || k <= 0 || k > 64)
    return set_error(ctx, PLANC_EINVAL, "planc_cp_loss_fwd_noalloc: bad arguments");
  for (int i = 0; i < nmodes; ++i)
    if (dims[i] <= 0) return set_error(ctx, PLANC_EINVAL, "planc_cp_loss_fwd_noalloc: empty mode");
  return cp_loss_forward_noalloc(ctx, d_X, nmodes, dims, d_factors, k, d_loss);
}

int planc_cp_loss_bwd(planc_ctx *ctx, const float *d_X, int nmodes, const long *dims,
                      const float *const *d_factors, int k, float *const *d_grads) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  const int N = nmodes;
  if (!d_X || !dims || !d_factors || !d_grads || N < 3 || N > 7 || k <= 0 || k > 64)
    return set_error(ctx, PLANC_EINVAL, "planc_cp_loss_bwd: bad arguments (3..7 modes, k 1..64)");
  long Imax = 0;
  for (int i = 0; i < N; ++i) {
    if (dims[i] <= 0) return set_error(ctx, PLANC_EINVAL, "planc_cp_loss_bwd: empty mode");
    Imax = std::max(Imax, dims[i]);
  }
  const int s = cp_split(N, dims);
  long L = 1, R = 1;
  for (int i = 0; i <= s; ++i) L *= dims[i];
  for (int i = s + 1; i < N; ++i) R *= dims[i];
  float *fb = (float *)ws_get(ctx, 13, sizeof(float) * ((size_t)(L + R) * k + (size_t)k * Imax));
  const size_t kk = (size_t)k * k;
  double *db = (double *)ws_get(ctx, 14, sizeof(double) * (N + 1) * kk);
  if (!fb || !db) return set_error(ctx, PLANC_ENOMEM, "planc_cp_loss_bwd: workspace");
  float *P = fb, *Q = fb + (size_t)L * k, *M = fb + (size_t)(L + R) * k;
  std::vector<const double *> G(N);
  double *S = db + N * kk;
  int rc;
  for (int i = 0; i < N; ++i) {
    G[i] = db + i * kk;
    rc = gram(ctx, d_factors[i], dims[i], k, dims[i], db + i * kk);
    if (rc) return rc;
  }
  KrpDesc kr{}, kl{};
  kr.nf = N - 1 - s;
  for (int q = s + 1; q < N; ++q) {
    kr.f[q - s - 1] = d_factors[q];
    kr.dim[q - s - 1] = dims[q];
    kr.ld[q - s - 1] = dims[q];
  }
  kl.nf = s + 1;
  for (int q = 0; q <= s; ++q) {
    kl.f[q] = d_factors[q];
    kl.dim[q] = dims[q];
    kl.ld[q] = dims[q];
  }
  rc = thin_gemm(ctx, PLANC_F32, 0, d_X, L, L, R, nullptr, 0, &kr, nullptr, nullptr, k, P, L, 0, 0);
  if (rc) return rc;
  rc = thin_gemm(ctx, PLANC_F32, 1, d_X, L, R, L, nullptr, 0, &kl, nullptr, nullptr, k, Q, R, 0, 0);
  if (rc) return rc;
  long ldf[8];
  for (int i = 0; i < N; ++i) ldf[i] = dims[i];
  for (int n = 0; n < N; ++n) {
    if (n <= s)
      rc = contract_projection(ctx, P, L, s + 1, dims, d_factors, ldf, n, k, M, k);
    else
      rc = contract_projection(ctx, Q, R, N - 1 - s, dims + s + 1, d_factors + s + 1, ldf + s + 1,
                               n - s - 1, k, M, k);
    if (rc) return rc;
    rc = hadamard_grams(ctx, G.data(), N, n, k, S);
    if (rc) return rc;
    rc = cp_grad_from_mttkrp(ctx, d_factors[n], dims[n], k, S, M, d_grads[n]);
    if (rc) return rc;
  }
  PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return PLANC_OK;
}

int planc_admm_update(planc_ctx *ctx, const double *d_G, int k, const float *d_R, long r_js, long r_cs,
                      float *d_X, long ldx, float *d_D, long ldd, long n) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (!d_G || !d_R || !d_X || !d_D || k <= 0 || k > 64 || n < 0 || ldx < n || ldd < n)
    return set_error(ctx, PLANC_EINVAL, "planc_admm_update: bad arguments (k must be 1..64)");
  int *dflag = (int *)ws_get(ctx, 11, sizeof(int));
  if (!dflag) return set_error(ctx, PLANC_ENOMEM, "planc_admm_update: workspace");
  PLANC_HIP_CHECK(ctx, hipMemsetAsync(dflag, 0, sizeof(int), ctx->stream));
  int rc = admm_solve(ctx, d_G, k, d_R, r_js, r_cs, d_X, ldx, d_D, ldd, n, dflag);
  if (rc) return rc;
  int hflag = 0;
  PLANC_HIP_CHECK(ctx, hipMemcpyAsync(&hflag, dflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (hflag) return set_error(ctx, PLANC_ENOCONV, "AO-ADMM: shifted Gram not positive definite");
  return PLANC_OK;
}

int planc_bpp(planc_ctx *ctx, const double *d_G, int k, const float *d_B, long b_cs, long b_rs,
              long nrhs, float *d_X, long x_cs, long x_rs) {
  if (check_ctx(ctx)) return PLANC_EINVAL;
  if (!d_G || !d_B || !d_X || k <= 0 || nrhs < 0)
    return set_error(ctx, PLANC_EINVAL, "planc_bpp: bad arguments");
  int *dflag = (int *)ws_get(ctx, 11, sizeof(int));
  if (!dflag) return set_error(ctx, PLANC_ENOMEM, "planc_bpp: workspace");
  PLANC_HIP_CHECK(ctx, hipMemsetAsync(dflag, 0, sizeof(int), ctx->stream));
  int rc = bpp_solve(ctx, d_G, k, d_B, b_cs, b_rs, nrhs, d_X, x_cs, x_rs, dflag);
  if (rc) return rc;
  int hflag = 0;
  PLANC_HIP_CHECK(ctx, hipMemcpyAsync(&hflag, dflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (hflag) return set_error(ctx, PLANC_ENOCONV, "BPP failed: more than 5k pivot iterations");
  return PLANC_OK;
}

}  // extern "C"
