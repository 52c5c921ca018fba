// Shared definitions of the planc_amd HIP library (gfx950 / MI355X only).
#ifndef PLANC_AMD_COMMON_H_
#define PLANC_AMD_COMMON_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/planc_hip.h"

namespace planc {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));

// 16-bit storage types of the mixed-precision copies of A / X.
struct bf16_t {
  uint16_t v;
};
struct f16_t {
  uint16_t v;
};

__host__ __device__ inline float bf16_to_float(uint16_t h) {
  union {
    uint32_t u;
    float f;
  } c;
  c.u = (uint32_t)h << 16;
  return c.f;
}

// Round to nearest even; NaN stays NaN.
__host__ __device__ inline uint16_t float_to_bf16(float f) {
  union {
    uint32_t u;
    float f;
  } c;
  c.f = f;
  if ((c.u & 0x7f800000u) == 0x7f800000u && (c.u & 0x007fffffu)) return (uint16_t)((c.u >> 16) | 0x40);
  uint32_t lsb = (c.u >> 16) & 1u;
  c.u += 0x7fffu + lsb;
  return (uint16_t)(c.u >> 16);
}

__device__ inline float f16_to_float(uint16_t h) {
  _Float16 x;
  __builtin_memcpy(&x, &h, 2);
  return (float)x;
}
__device__ inline uint16_t float_to_f16(float f) {
  _Float16 x = (_Float16)f;
  uint16_t h;
  __builtin_memcpy(&h, &x, 2);
  return h;
}

// Value of element e of a storage array of the given dtype as float.
template <typename T>
__device__ inline float load_as_float(const T *p, long e);
template <>
__device__ inline float load_as_float<float>(const float *p, long e) {
  return p[e];
}
template <>
__device__ inline float load_as_float<bf16_t>(const bf16_t *p, long e) {
  return bf16_to_float(p[e].v);
}
template <>
__device__ inline float load_as_float<f16_t>(const f16_t *p, long e) {
  return f16_to_float(p[e].v);
}

// Workspace cache owned by a context: grows, never shrinks until destroy.
struct Workspace {
  void *ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace planc

struct planc_graph_entry {
  std::vector<uintptr_t> key;
  hipGraphExec_t exec;
};

struct planc_ctx {
  std::vector<planc_graph_entry> graphs;  // instantiated loop graphs, keyed by shapes and pointers
  hipStream_t capture_stream = nullptr;   // private stream graphs are captured on
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string last_error;
  planc::Workspace ws[40];
  int num_cu = 256;
  int max_coop_blocks_hals = 0;
  // 2: 16-bit products carry the thin operand as two 16-bit planes (hi and
  // the rounded remainder lo), so only the big operand is rounded; 1: one
  // plane.  Set per run by ThinPlanes.
  int thin_planes = 1;
  // Tuned streaming-product geometries: shape key -> {cw, wk, nsplit, dual, shared}.
  std::vector<std::pair<std::vector<long>, std::vector<long>>> gemm_tuned;
  // 16-bit dimension-tree contraction forms chosen per shape: key -> bit 0
  // (right contraction as TN on the transposed copy), bit 1 (left
  // contraction as NN on the transposed copy).
  std::vector<std::pair<std::vector<long>, int>> tree_forms;
};

namespace planc {

// Returns a device buffer of at least `bytes` for workspace slot `slot`.
// Slots: 0 split-K partials, 1 Gram partials, 2 column-norm / HALS-sweep
// partials, 3 dot partials, 4 loop float buffers, 5 loop double buffers,
// 6 packed 16-bit thin operand, 7 16-bit copy of A / X, 8 column scales of that copy, 9 entry-point norms,
// 10 BPP per-column iteration counts, 11 BPP failure flag, 12 CP-loss scalar,
// 13 CP-gradient float buffers, 14 CP-gradient Grams, 15 loop-graph counter,
// 16 transposed 16-bit copy of A / X, 17 HALS grid-barrier counter and flag,
// 18 NTF mode-update partials (two-launch form), 19/21 (unused), 20 (unused), 22 AO-ADMM factor / partials,
// 23 GNSym CG vectors, 24..31 sparse path (sparse.hip: k-major factor
// copies, error partials, transposed CSR, sort keys / entries / temporary
// storage), 32 NTF AO-ADMM duals, 33 NMF AO-ADMM duals, 34 dual norms,
// 35 NES iterates, 36 NES factors (centres, previous, accelerated), 37 NES
// lambdas.
void *ws_get(planc_ctx *ctx, int slot, size_t bytes);

// Scoped choice of ctx->thin_planes (restored on exit).
struct ThinPlanes {
  planc_ctx *ctx;
  int saved;
  ThinPlanes(planc_ctx *c, int np) : ctx(c), saved(c->thin_planes) { c->thin_planes = np; }
  ~ThinPlanes() { ctx->thin_planes = saved; }
  ThinPlanes(const ThinPlanes &) = delete;
  ThinPlanes &operator=(const ThinPlanes &) = delete;
};

int set_error(planc_ctx *ctx, int code, const std::string &msg);

#define PLANC_HIP_CHECK(ctx, call)                                                  \
  do {                                                                              \
    hipError_t _e = (call);                                                         \
    if (_e != hipSuccess)                                                           \
      return ::planc::set_error((ctx), PLANC_EHIP,                                  \
                                std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define PLANC_LAUNCH_CHECK(ctx) PLANC_HIP_CHECK(ctx, hipGetLastError())

// Captures everything `record` launches on ctx->stream into a graph: the
// work is recorded on the context's private capture stream (the caller's
// stream may be the legacy NULL stream, which cannot be captured) and the
// instantiated graph is returned for launching on ctx->stream.
template <class F>
int capture_graph(planc_ctx *ctx, F record, hipGraphExec_t *exec) {
  if (!ctx->capture_stream) {
    if (hipStreamCreateWithFlags(&ctx->capture_stream, hipStreamNonBlocking) != hipSuccess)
      return set_error(ctx, PLANC_EHIP, "capture stream");
  }
  hipStream_t saved = ctx->stream;
  ctx->stream = ctx->capture_stream;
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    ctx->stream = saved;
    return set_error(ctx, PLANC_EHIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
  }
  int rc = record();
  e = hipStreamEndCapture(ctx->stream, &graph);
  ctx->stream = saved;
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return set_error(ctx, PLANC_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return set_error(ctx, PLANC_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  return PLANC_OK;
}

// Kernel-family entry points shared between translation units.
struct KrpDesc {
  int nf = 0;               // number of factors in the Khatri-Rao product (0: plain B)
  const float *f[6];        // factor matrices, column-major I_q x k
  long dim[6];              // I_q
  long ld[6];               // leading dimension of each factor
};

// C = A * B (trans == 0; A is L x K col-major, C is L x k col-major) or
// C = B^T * A (trans == 1; A is K x L col-major, C is k x L col-major), with
// B the K x k operand: a column-major matrix (krp.nf == 0) or the Khatri-Rao
// product of krp.f (lowest factor fastest).  rowscale (K, optional) scales
// the rows of B; colscale (L, optional; trans == 1 or a 16-bit A) scales
// output element l (TN: C's column l; NN: C's row l).
// cmajor selects the output layout: 1 -> C[c + ldc*l] (k x L), 0 -> C[l + ldc*c].
// splits_hint > 0 forces the split-K factor (0: automatic).  oscale (16-bit
// A only) multiplies every output element after the FP32 accumulation.
int thin_gemm(planc_ctx *ctx, int dtype, int trans, const void *A, long lda, long L, long K,
              const float *B, long ldb, const KrpDesc *krp, const float *rowscale,
              const float *colscale, int k, float *C, long ldc, int cmajor, int splits_hint,
              float oscale = 1.0f);

// nmf_kernels.hip
int gram(planc_ctx *ctx, const float *X, long rows, int k, long ldx, double *G);
int read_bandwidth(planc_ctx *ctx, const void *buf, size_t bytes, int reps, double *gbps);
// kind 0: MU (eps), kind 1: HALS rows.  N(j, c) = N[j*nrs + c*ncs].
int row_update(planc_ctx *ctx, int kind, float *H, long n, int k, long ldh, const float *N, long nrs,
               long ncs, const double *G, float eps);
int hals_cols(planc_ctx *ctx, float *W, long m, int k, long ldw, const float *AH, long ldah,
              const double *G);
// row_update fused with the FP64 Gram of the updated rows (into Gout, k x k)
// and, with dparts, FP64 block partials of <updated rows, N> (*ndp of them,
// workspace slot 3, for nmf_relerr_parts).  rows_gram_ok: the layouts it
// takes (N column-major, nrs == 1, or row-major, ncs == 1; leading
// dimensions multiples of 4; 16-byte aligned H and N; k <= 64).
bool rows_gram_ok(const float *H, long ldh, const float *N, long nrs, long ncs, int k);
int row_update_gram(planc_ctx *ctx, int kind, float *H, long n, int k, long ldh, const float *N, long nrs,
                    long ncs, const double *G, float eps, double *Gout, double **dparts, long *ndp);
int col_norms(planc_ctx *ctx, const float *X, long rows, int k, long ldx, double *norms);
// NCPFactors::normalize() of n tight factors (dims[i] x k, ld dims[i]) with
// lambda = the product of their column norms and G[i] = the FP64 Gram of
// normalised factor i, in two launches; only for factors of at most 1024
// rows and 16384 values (init_normalize_fits), k <= 64.
bool init_normalize_fits(const long *dims, int n, int k);
int init_normalize(planc_ctx *ctx, float *const *F, const long *dims, int n, int k, double *const *G,
                   double *lambda);
int scale_cols(planc_ctx *ctx, float *X, long rows, long ldx, float *Y, long yrows, long ldy, int k,
               const double *norms);
int dot(planc_ctx *ctx, const float *X, long xrs, long xcs, const float *Y, long yrs, long ycs,
        long rows, long cols, double *out, int accumulate);
int nmf_relerr(planc_ctx *ctx, const double *normA2, const double *dotv, const double *WtW,
               const double *HtH, int k, double *relerr_out, double *fit_out);
// dot() without its final sum: the block partials (workspace slot 3) and
// their count; nmf_relerr_parts sums them inside the error kernel.
int dot_parts(planc_ctx *ctx, const float *X, long xrs, long xcs, const float *Y, long yrs, long ycs,
              long rows, long cols, double **part, long *nparts);
int nmf_relerr_parts(planc_ctx *ctx, const double *normA2, const double *part, long nparts, const double *WtW,
                     const double *HtH, int k, double *relerr_out);
int apply_reg(planc_ctx *ctx, double *G, int k, const double *reg);
// hi = X rounded to dtype's 16-bit format (as float), lo = X - hi; Y += X.
// G += sr I, R(j, c) += sr F(j, c) (R element (j, c) at R[j*rs + c*cs]); no-op for sr <= 0.
int symm_reg_apply(planc_ctx *ctx, double *G, int k, float *R, long rs, long cs, const float *F,
                   long n, double sr);
// Time-out flag of the HALS column sweep's grid barrier: reset before a
// run, checked (synchronising) after it.
int hals_sync_reset(planc_ctx *ctx);
int hals_sync_check(planc_ctx *ctx);
// out = D^-1 G D^-1 with D = diag(norms) (entries 0 taken as 1).
int gram_rescale(planc_ctx *ctx, const double *G, int k, const double *norms, double *out);
// norms[c] = sqrt(G(c, c)).
int gram_norms(planc_ctx *ctx, const double *G, int k, double *norms);
int convert16(planc_ctx *ctx, int dtype, const float *A, long rows, long cols, long lda, void *out,
              long ldo, float *colscale);
// out (cols x rows) = in^T, 16-bit elements.
int transpose16(planc_ctx *ctx, const void *in, long rows, long cols, long ldi, void *out, long ldo);

// bpp.hip: column j of the right-hand side is B[j*b_cs + i*b_rs]; the
// solution goes to X[j*x_cs + i*x_rs]; *d_failed |= (some column > 5k pivots).
int bpp_solve(planc_ctx *ctx, const double *G, int k, const float *B, long b_cs, long b_rs, long nrhs,
              float *X, long x_cs, long x_rs, int *d_failed);

// cploss.hip
int cp_loss_forward(planc_ctx *ctx, const float *X, int N, const long *dims, const float *const *F,
                    int k, double *d_loss);
// Sparse (CSR) products of the sparse-input NMF (sparse.hip).  SparseOps
// holds the CSR of A^T built once per run.
struct SparseOps {
  const long *t_indptr = nullptr;
  const int *t_indices = nullptr;
  const float *t_values = nullptr;
};
int csr_transpose(planc_ctx *ctx, long m, long n, long nnz, const long *indptr, const int *indices,
                  const float *values, long *out_indptr, int *out_indices, float *out_values);
int spmm_csr(planc_ctx *ctx, long rows, const long *indptr, const int *indices, const float *values,
             const float *F, long ldf, int k, float *out, long ors, long ocs);
int spmm_colmajor(planc_ctx *ctx, long rows, long cols, const long *indptr, const int *indices,
                  const float *values, const float *F, long ldf, int k, float *out, long ldo);
int sparse_fit_error(planc_ctx *ctx, long m, long n, const long *indptr, const int *indices,
                     const float *values, const float *W, long ldw, const float *H, long ldh, int k,
                     double *d_err);
int sparse_products(planc_ctx *ctx, long m, long n, long nnz, const long *indptr, const int *indices,
                    const float *values, SparseOps *ops);
int sparse_wta(planc_ctx *ctx, const SparseOps &o, long m, long n, const float *W, int k, float *out);
int sparse_ah(planc_ctx *ctx, const SparseOps &o, long m, long n, const long *indptr, const int *indices,
              const float *values, const float *H, int k, float *out);

// AO-ADMM solve of X (n x k) and its dual D given the regularised Gram G
// and R(c, j) at R[j*r_js + c*r_cs]; *err |= 4 when G + rho I is not
// positive definite (admm_gn.hip).
int admm_solve(planc_ctx *ctx, const double *G, int k, const float *R, long r_js, long r_cs, float *X,
               long ldx, float *D, long ldd, long n, int *err);
// One symmetric-NMF Gauss-Newton step of H (n x k) from AHt = H^T A (k x n)
// and HtH of the current H.
int gnsym_step(planc_ctx *ctx, float *H, long n, int k, const float *AHt, const double *HtH);

// NTFNES pieces (nes.hip): the extreme eigenvalues of a symmetric k x k
// matrix (host), one mode update, the accelerated factor and a column scaling.
void sym_eig_extremes(const std::vector<double> &S, int k, double *mn, double *mx);
int nes_mode_update(planc_ctx *ctx, const double *dS, int k, const float *M, long mrs, long mcs, float *prox,
                    long I, float *F, long ldf, double *prox_l, int *status);
int nes_acc_build(planc_ctx *ctx, const float *F, const float *Pv, long I, int k, const double *sc,
                  const double *sp, double step, float *A);
int scale_cols_by(planc_ctx *ctx, const float *A, long I, int k, const double *s, float *out);

// Allocation-free forward: block partials in a per-device static array
// located once per device by cp_noalloc_setup (called from planc_create).
int cp_noalloc_setup(int device);
int cp_loss_forward_noalloc(planc_ctx *ctx, const float *X, int N, const long *dims,
                            const float *const *F, int k, double *d_loss);
int cp_grad_from_mttkrp(planc_ctx *ctx, const float *A, long I, int k, const double *S, const float *M,
                        float *grad);
int cp_split(int N, const long *dims);

// ntf_kernels.hip
int contract_projection(planc_ctx *ctx, const float *P, long ldp, int nb, const long *d,
                        const float *const *F, const long *ldf, int tn, int k, float *M, long ldo);
int hadamard_grams(planc_ctx *ctx, const double *const *grams, int n, int skip, int k, double *out);
int ntf_relerr(planc_ctx *ctx, const double *normX2, const float *F, long I, long ldf, const float *M,
               long ldm, int k, const double *lambda, const double *const *grams, int n,
               double *rel_out);
int lambda_mul(planc_ctx *ctx, double *lambda, const double *norms, int k, int init);
int transpose_to_k_major(planc_ctx *ctx, const float *in, long I, long ldi, int k, float *out, long ldo);
// Fused AUNTF mode update (Hadamard, row update, normalize(mode), Gram, error)
// in two launches; kind 0 MU, 1 HALS, 2 none (factor already solved).
bool mode_update_fits(long I, int k);
// rel_out: where the relative error goes (null: none); with rel_counter it
// is rel_out[*rel_counter] and the device counter is incremented.
int mode_update(planc_ctx *ctx, int kind, float eps, float *F, long I, long ldf, int k, const float *M,
                long mrs, long mcs, const double *const *grams, int ng, int skip, double *gout,
                double *lambda, const double *normX2, double *rel_out, int *rel_counter);

}  // namespace planc

#endif  // PLANC_AMD_COMMON_H_
