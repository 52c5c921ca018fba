/*
 * planc_hip.h -- C ABI of the MI355X (gfx950) implementation of PLANC's
 * shared-memory NMF/NTF hot path.
 *
 * The reference (PLANC) binds its arithmetic through CBLAS/LAPACK calls and
 * Armadillo expressions inside C++ class templates; each entry point below
 * replaces one of those call sites or one of the class-level loops, cited as
 * reference file:line (paths relative to the PLANC source tree).
 *
 * Conventions
 *  - Matrices are column-major (Armadillo's layout): W is m x k, H is n x k,
 *    NTF factors are I_n x k, an MTTKRP result is k x I_n, tensors are
 *    stored with mode 0 fastest.  Leading dimensions are in elements.
 *  - Every pointer named d_* is device memory (hipMalloc / planc_malloc).
 *    Host buffers are named h_*.
 *  - Calls are ordered on the context's stream and return once enqueued,
 *    except the *_run entry points, planc_sync and the copy helpers, which
 *    return after the stream has drained.
 *  - Every function returns 0 (PLANC_OK) or a negative status; no C++
 *    exception crosses this boundary.  planc_last_error() describes the last
 *    failure on a context.
 *  - One context per host thread.  The caller owns every buffer it passes;
 *    the context owns its scratch.
 */
#ifndef PLANC_HIP_H_
#define PLANC_HIP_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (replace the reference's exit()/assert failure paths). */
#define PLANC_OK 0
#define PLANC_EINVAL (-1)     /* invalid argument (shape, pointer, enum) */
#define PLANC_EHIP (-2)       /* HIP runtime error */
#define PLANC_ENOCONV (-3)    /* BPP exceeded 5k pivot iterations (nnls/bppnnls.hpp:185-188) */
#define PLANC_ENOMEM (-4)     /* device allocation failed */
#define PLANC_EUNSUPPORTED (-5)

/* Arithmetic of the bulk products (A*H, W^T*A, MTTKRP).  FP32 is the plain
 * form; BF16/FP16 read a 16-bit copy of A / X on MFMA with FP32
 * accumulation.  Grams, normal equations and error sums are FP64 always. */
#define PLANC_F32 0
#define PLANC_BF16 1
#define PLANC_F16 2

/* Update algorithms; same numbering as the reference's algotype enum
 * (common/utils.h:10-11). */
#define PLANC_MU 0
#define PLANC_HALS 1
#define PLANC_ANLSBPP 2
#define PLANC_AOADMM 4
#define PLANC_NESTEROV 5
#define PLANC_GNSYM 7
#define PLANC_R2 8

typedef struct planc_ctx planc_ctx;

/* ---- context, memory, streams ------------------------------------------- */
const char *planc_version(void);
int planc_create(int device, planc_ctx **out);
int planc_destroy(planc_ctx *ctx);
int planc_sync(planc_ctx *ctx);
const char *planc_last_error(planc_ctx *ctx);
/* Use an existing hipStream_t (passed as void*) instead of the context's own. */
int planc_set_stream(planc_ctx *ctx, void *stream);
void *planc_get_stream(planc_ctx *ctx);
int planc_malloc(planc_ctx *ctx, void **d_ptr, size_t bytes);
int planc_free(planc_ctx *ctx, void *d_ptr);
int planc_memcpy_h2d(planc_ctx *ctx, void *d_dst, const void *h_src, size_t bytes);
int planc_memcpy_d2h(planc_ctx *ctx, void *h_dst, const void *d_src, size_t bytes);

/* ---- data preparation ---------------------------------------------------- */
/* 16-bit copy of an FP32 matrix (rows x cols, column-major) for the BF16/FP16
 * forms.  When d_colscale is non-null each column j is divided by a power of
 * two s_j (written to d_colscale[j]) chosen so that its largest magnitude
 * lies in [0.5, 1); the products undo the scale exactly.  No reference
 * counterpart (the reference computes in FP64 only). */
int planc_convert(planc_ctx *ctx, int dtype, const float *d_src, long rows, long cols, long ld_src,
                  void *d_dst, long ld_dst, float *d_colscale);

/* d_out (cols x rows, ld ldo) = d_in^T for a 16-bit matrix d_in (rows x
 * cols, ld ldi): the transposed copy A16T of planc_nmf_desc, the analogue of
 * the reference's At = A.t() (nmf/mu.hpp:42,46). */
int planc_transpose16(planc_ctx *ctx, const void *d_in, long rows, long cols, long ldi, void *d_out,
                      long ldo);

/* Synthetic inputs on the device (Philox4x64-10; element e of stream
 * (seed, stream) equals numpy.random.Philox(key=[seed, 0],
 * counter=[0, stream, 0, 0]) word e).  No reference counterpart: the
 * reference draws synthetic inputs from Armadillo's RNG (nmf/nmf.cpp:
 * 132-167), which cannot be reproduced without Armadillo.
 *  planc_gen_uniform:   d_out[e] = U[0,1)_e (53-bit double rounded to float).
 *  planc_gen_lowrank:   A = W0 (m x r) * H0 (r x n) + sigma * |N(0,1)|, the
 *                       product in FP64, the normal by Box-Muller.
 *  planc_gen_cp_tensor: X = [[F_0..F_{N-1}]] + sigma * |N(0,1)|, mode 0 fastest. */
int planc_gen_uniform(planc_ctx *ctx, float *d_out, long count, unsigned long long seed,
                      unsigned long long stream);
int planc_gen_lowrank(planc_ctx *ctx, long m, long n, int r, const float *d_W0, const float *d_H0,
                      double sigma, unsigned long long seed, unsigned long long stream, float *d_A,
                      long lda);
int planc_gen_cp_tensor(planc_ctx *ctx, int nmodes, const long *dims, const float *const *d_factors,
                        int r, double sigma, unsigned long long seed, unsigned long long stream,
                        float *d_X);

/* ---- kernel-level entry points ------------------------------------------- */
/* C (L x k) = A (L x K) * B (K x k).  Replaces the dgemm behind `A * H`
 * (nmf/mu.hpp:71, nmf/hals.hpp:81) and, with B a formed KRP, the mode-0
 * MTTKRP dgemm (common/tensor.hpp:279-281).  dtype is that of A; B is FP32
 * and is rounded to dtype on the fly; d_colscale (optional, K) are the column
 * scales produced by planc_convert for A. */
int planc_gemm_nn(planc_ctx *ctx, int dtype, const void *d_A, long lda, long L, long K,
                  const float *d_colscale, const float *d_B, long ldb, int k, float *d_C, long ldc);

/* C (k x L) = B^T (k x K) * A (K x L).  Replaces `W.t() * A`
 * (nmf/hals.hpp:57, nmf/bppnmf.hpp:39) and `At * W` (nmf/mu.hpp:55, returned
 * transposed).  d_colscale (optional, L) scales the output columns. */
int planc_gemm_tn(planc_ctx *ctx, int dtype, const void *d_A, long lda, long K, long L,
                  const float *d_colscale, const float *d_B, long ldb, int k, float *d_C, long ldc);

/* Device read bandwidth (no counterpart in the reference; a measurement
 * utility for roofline fractions): reads the caller's buffer d_buf of
 * `bytes` (16-byte aligned, >= 16 MiB) `reps` times with plain 16-byte
 * grid-stride loads on the context stream and returns the mean rate in GB/s
 * in *gbps.  The buffer's contents are not changed. */
int planc_read_bandwidth(planc_ctx *ctx, const void *d_buf, size_t bytes, int reps, double *gbps);

/* G (k x k, FP64) = X^T X for X rows x k (FP32).  Replaces the Gram products
 * `W.t() * W` / `H.t() * H` (nmf/mu.hpp:56,72; common/ncpfactors.hpp:144). */
int planc_gram(planc_ctx *ctx, const float *d_X, long rows, int k, long ldx, double *d_G);

/* MU update H <- H .* N ./ (H * G + eps), in place, for H n x k, the
 * numerator N n x k with element (j, c) at d_N[j*n_rs + c*n_cs] (so a k x n
 * product such as W^T A is passed with n_rs = k, n_cs = 1) and G k x k FP64.
 * nmf/mu.hpp:64,79 (eps = 1e-17); ntf/ntfmu.hpp:13 (eps = 1e-6). */
int planc_mu_update(planc_ctx *ctx, float *d_H, long n, int k, long ldh, const float *d_N,
                    long n_rs, long n_cs, const double *d_G, double eps);

/* HALS sweep over the columns of H whose rows are independent:
 * for x = 0..k-1: H(:,x) <- max(H(:,x) + N(:,x) - H * G(:,x), 1e-17)
 * (values below 1e-17 become 1e-17); N as in planc_mu_update.
 * nmf/hals.hpp:67-75 and ntf/ntfhals.hpp:13-29. */
int planc_hals_update_rows(planc_ctx *ctx, float *d_H, long n, int k, long ldh, const float *d_N,
                           long n_rs, long n_cs, const double *d_G);

/* HALS sweep over the columns of W with per-column normalisation (Gauss-Seidel
 * across columns, one grid-wide reduction per column):
 * Wx = W(:,x)*G(x,x) + AH(:,x) - W*G(:,x); clamp 1e-17; W(:,x) = Wx/||Wx||.
 * nmf/hals.hpp:89-102. */
int planc_hals_update_cols(planc_ctx *ctx, float *d_W, long m, int k, long ldw, const float *d_AH,
                           long ldah, const double *d_G);

/* normalize_by_W (common/nmf.hpp:152-161): W columns to unit 2-norm, H
 * columns multiplied by the same norms; zero columns are left alone. */
int planc_normalize_by_w(planc_ctx *ctx, float *d_W, long m, long ldw, float *d_H, long n,
                         long ldh, int k);

/* NCPFactors::normalize(mode) (common/ncpfactors.hpp:346-351): columns of A
 * (rows x k) to unit norm, d_lambda (k, FP64) replaced by the norms. */
int planc_normalize_cols(planc_ctx *ctx, float *d_A, long rows, int k, long lda, double *d_lambda);

/* Hadamard product of the k x k FP64 Grams d_grams[i], i != skip
 * (gram_leave_out_one, common/ncpfactors.hpp:139-148).  d_grams is a host
 * array of device pointers. */
int planc_hadamard_grams(planc_ctx *ctx, const double *const *d_grams, int nmodes, int skip, int k,
                         double *d_out);

/* MTTKRP of mode `mode` for a dense tensor (mode 0 fastest, dims[nmodes]),
 * with the Khatri-Rao product formed on the fly, never stored.  Output is
 * k x I_mode (ld ldo).  Replaces Tensor::mttkrp (common/tensor.hpp:246-327)
 * plus krp_leave_out_one (common/ncpfactors.hpp:171-233).  d_factors is a
 * host array of device pointers (I_i x k, column-major).  nmodes 2..7. */
int planc_mttkrp(planc_ctx *ctx, int dtype, const void *d_X, int nmodes, const long *dims,
                 const float *const *d_factors, int k, int mode, float *d_out, long ldo);

/* Partial MTTKRP of the dimension tree (partial_MTTKRP_with_KRP,
 * dimtree/dimtrees.hpp:176-237): with the modes split at s (left = modes
 * 0..s, L = prod of their dims; right = the rest, R), side 0 computes the
 * right contraction P = X_(L x R) * KRP(right factors), an L x k projection;
 * side 1 the left contraction KRP(left factors)^T * X_(L x R), stored as an
 * R x k projection.  d_out is column-major with ld ldo (>= L or R). */
int planc_mttkrp_partial(planc_ctx *ctx, int dtype, const void *d_X, int nmodes, const long *dims,
                         const float *const *d_factors, int k, int split, int side, float *d_out,
                         long ldo);

/* Multi-TTV of the dimension tree (multi_TTV_with_KRP, dimtree/dimtrees.hpp:
 * 418-455): contracts a projection P (E x k, column-major, ld ldp) over all
 * modes of the block d_factors[0..nb) except `target`, giving the MTTKRP of
 * that mode as k x dims[target] (ld ldo).  dims/d_factors describe the
 * block's modes, lowest mode fastest in P's row index. */
int planc_multi_ttv(planc_ctx *ctx, const float *d_P, long ldp, int nb, const long *dims,
                    const float *const *d_factors, int target, int k, float *d_out, long ldo);

/* Batched block-principal-pivoting NNLS, FP64 arithmetic: for every column j
 * solves min ||A x - a_j|| s.t. x >= 0 from the normal equations
 * G = A^T A (k x k FP64, k <= 64) and b_j = A^T a_j, starting from x = 0 with
 * the reference's exchange rules, 1e-12 clamps and 5k iteration cap.
 * Replaces BPPNNLS::solveNNLS (nnls/bppnnls.hpp:23-191) as called per
 * 2000-column chunk by nmf/bppnmf.hpp:57-89 and ntf/ntfanlsbpp.hpp:22-55.
 * b_j(i) = d_B[j*b_cs + i*b_rs]; x_j(i) is written to d_X[j*x_cs + i*x_rs].
 * Synchronises; returns PLANC_ENOCONV when a column exceeded 5k pivots
 * (the reference calls exit(EXIT_FAILURE), nnls/bppnnls.hpp:185-188). */
int planc_bpp(planc_ctx *ctx, const double *d_G, int k, const float *d_B, long b_cs, long b_rs,
              long nrhs, float *d_X, long x_cs, long x_rs);

/* One AO-ADMM solve (the inner loop of AOADMMNMF, nmf/aoadmm.hpp:105-141 /
 * :157-185, and of NTFAOADMM::update, ntf/ntfaoadmm.hpp:34-72): with
 * rho = trace(G)/k (0.01 if not positive) and L = chol(G + rho I), at most 5
 * sweeps of  T = (G + rho I)^-1 (R + rho (X + D)^T),  X = max(T^T - D, 0),
 * D += X - T^T,  stopping when ||X - T^T|| < 0.01 ||X|| and ||X - X_prev|| <
 * 0.01 ||D||.  G k x k FP64 (the Gram with its regularisation), R(c, j) at
 * d_R[j * r_js + c * r_cs] (k x n), X and its dual D n x k (lds ldx, ldd),
 * updated in place; the dual carries over between calls as the reference's
 * does.  Synchronises; PLANC_ENOCONV when G + rho I is not positive definite
 * (the reference's arma::chol fails).  k <= 64. */
int planc_admm_update(planc_ctx *ctx, const double *d_G, int k, const float *d_R, long r_js, long r_cs,
                      float *d_X, long ldx, float *d_D, long ldd, long n);

/* CP reconstruction loss 0.5 ||X - [[A_0..A_{N-1}]]||^2 (no lambda), FP64
 * result in *h_loss: one fused kernel reconstructs tiles of X from
 * Khatri-Rao tiles formed in LDS and reduces the squared residual, reading X
 * once.  The reference evaluates the same residual by forming the full
 * reconstruction (AUNTF::computeObjectiveError, ntf/auntf.hpp:200-246).
 * Synchronises. */
int planc_cp_loss_fwd(planc_ctx *ctx, const float *d_X, int nmodes, const long *dims,
                      const float *const *d_factors, int k, double *h_loss);

/* The same loss, allocating nothing: no reference counterpart.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving or training loops, harnesses that forbid hipMalloc).  Same
 * arithmetic, block partition and summation order as planc_cp_loss_fwd, but
 * the block partials live in a fixed per-GPU device array of 16384 doubles
 * (128 KiB, held until the process ends) instead of context scratch, and the
 * loss is written to the caller's device double *d_loss.  Stream-ordered:
 * returns once enqueued, no synchronisation.  Returns PLANC_EINVAL when the
 * shape needs more than 16384 partials (the right block of modes -- the
 * modes after the balanced split -- longer than about 1.8 million rows).
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (two contexts or streams in flight at once) race and must be ordered by
 * the caller. */
int planc_cp_loss_fwd_noalloc(planc_ctx *ctx, const float *d_X, int nmodes, const long *dims,
                              const float *const *d_factors, int k, double *d_loss);

/* Gradient of that loss with respect to every factor:
 * d_grads[n] (I_n x k) = A_n * (Hadamard of the other Grams) - MTTKRP_n^T,
 * the transposed NTFNES gradient S*Y - M (ntf/ntfnes.hpp:138-147) without
 * its lambda shift; the MTTKRPs come from the two partial contractions of
 * the dimension tree with the KRP recomputed, never stored.  3..7 modes.
 * Synchronises. */
int planc_cp_loss_bwd(planc_ctx *ctx, const float *d_X, int nmodes, const long *dims,
                      const float *const *d_factors, int k, float *const *d_grads);

/* One rank-2 NNLS update (DistR2::update, distnmf/distr2.hpp:35-61): for
 * each of n columns, X(:, j) solves the 2 x 2 normal equations G x = R(:, j)
 * (LU with partial pivoting, as arma::solve's fast path); a column with a
 * negative entry is replaced by the better one-variable solution
 * (r_c / G_cc, compared as sqrt(G_cc) r_c / G_cc).  d_G 2 x 2 column-major
 * FP64; R(c, j) at d_R[j * r_js + c * r_cs], X(c, j) at d_X[j * x_js + c *
 * x_cs] (c = 0, 1).  Stream-ordered. */
int planc_r2_update(planc_ctx *ctx, const double *d_G, const float *d_R, long r_js, long r_cs, long n,
                    float *d_X, long x_js, long x_cs);

/* Relative error of NMF from the Gram trace identity (computeObjectiveError,
 * common/nmf.hpp:390-424, with printObjective's relative form, :462):
 * fit^2 = ||A||^2 - 2 tr(H^T A^T W) + sum(W^T W .* H^T H), *d_rel =
 * sqrt(fit^2) / ||A|| (1 when fit^2 <= 0).  All inputs are device FP64:
 * *d_normA2 = ||A||_F^2, *d_trace = tr(H^T A^T W) = <W, A H>, the k x k Grams.
 * Stream-ordered. */
int planc_rel_error_gram(planc_ctx *ctx, const double *d_normA2, const double *d_trace, const double *d_WtW,
                         const double *d_HtH, int k, double *d_rel);

/* Column subset of a CSR matrix (m x n): the node matrix A0(:, cols) of
 * hiernmf (Node::allocate, the sparse selection A0 * S, hiernmf/node.hpp:
 * 46-64).  h_cols: sorted, unique column ids (host).  Rows are kept, columns
 * renumbered 0..ncols-1, entry order and values unchanged (bit-exact).
 * d_out_indptr has m + 1 entries; the outputs must hold *h_out_nnz entries
 * (call once with null outputs to get the count).  Synchronises. */
int planc_csr_col_subset(planc_ctx *ctx, long m, long n, const long *d_indptr, const int *d_indices,
                         const float *d_data, const long *h_cols, long ncols, long *d_out_indptr,
                         int *d_out_indices, float *d_out_data, long *h_out_nnz);

/* Hierarchical rank-2 NMF (HierNMFDriver::buildTree, hiernmf/hiernmf.cpp:
 * 65-184).  Splitting a node runs `iters` iterations of the rank-2 NMF of
 * Node::split (hiernmf/node.hpp:104-188; DistAUNMF order H then W, DistR2
 * update); its columns split by H(:,0) >= H(:,1), the larger side becoming
 * child 2i+1; every child gets sigma by an `iters`-step power iteration
 * (Node::compute_sigma) and its parent the score of Node::compute_score.
 * Node seeds: initseed * index (rank 0); W from stream 0, H from stream 1,
 * the power-iteration start from normals of stream 2 (Philox).  Nodes are
 * processed in batches (all nodes of a batch as one block-diagonal matrix,
 * one hipGraph-captured iteration loop).
 *   desc->nodes <= 0: the complete tree of `depth` levels, level by level;
 *     h_nodes in level order (capacity 2^depth - 1).
 *   desc->nodes > 0: the reference's best-first frontier with that node
 *     budget (hiernmf/hiernmf.cpp:138-147): the root is split and accepted,
 *     then nodes - 2 times the frontier node of highest score (ties: smaller
 *     index) is accepted; h_nodes holds the reference's output queue (the
 *     root, its children, then the children of each accepted node in order;
 *     capacity max(3, 2 nodes - 1)); `depth` (2 .. 62) caps the levels (a
 *     node on level depth - 1 is never split); the complete tree takes
 *     depth 1 .. 24.
 * A is a host CSR (m x n, FP32 values, sorted column indices).  Outputs
 * (host): h_cols (capacity n * depth) each node's sorted columns at its
 * col_offset, optional h_H (2 doubles per entry of h_cols: the node's own
 * final H rows when it was split) and h_W (2*m doubles per reported node, in
 * h_nodes order: the node's own W; zero rows where the node's matrix is
 * empty).  The reference writes for every node but the root its parent's
 * factor column on its side (Node's W and H): column parent_side of the
 * parent's W and H. */
typedef struct planc_hier_desc {
  long m, n;
  int depth;     /* levels, root = level 0 (the cap in frontier mode) */
  int iters;     /* NMF and power iterations per node (-t) */
  int initseed;  /* --seed */
  int nodes;     /* --nodes: > 0 best-first frontier with this budget; <= 0 complete tree */
} planc_hier_desc;

typedef struct planc_hier_node {
  long index;      /* 0 root, children 2i+1 / 2i+2 */
  int level;
  long ncols;
  long col_offset; /* into h_cols / h_H */
  double sigma;    /* 0 for the root (RootNode) */
  double score;    /* Node::compute_score, hiernmf/node.hpp:71-81 */
  int lvalid, rvalid;
  int split;       /* 1: the node's rank-2 NMF ran (its W / H are in h_W / h_H) */
  int parent_side; /* column of the parent's W / H this node carries (-1: root) */
} planc_hier_node;

int planc_hiernmf_run(planc_ctx *ctx, const planc_hier_desc *desc, const long *h_indptr,
                      const int *h_indices, const float *h_data, planc_hier_node *h_nodes,
                      int *h_num_nodes, long *h_cols, double *h_H, double *h_W);

/* ---- whole loops (computeNMF / computeNTF on the device) ----------------- */
typedef struct planc_nmf_desc {
  int algo;          /* PLANC_MU, PLANC_HALS, PLANC_ANLSBPP, PLANC_AOADMM, PLANC_GNSYM */
  int dtype;         /* PLANC_F32 / PLANC_BF16 / PLANC_F16 */
  long m, n;         /* A is m x n */
  int k;             /* rank */
  int iters;         /* outer iterations (num_iterations, common/nmf.hpp:481) */
  double regW[2];    /* {L2, L1} (regW, common/nmf.hpp:487) */
  double regH[2];
  int compute_error; /* 1: relative error every iteration (always on in the
                        reference's shared-memory driver, nmf/nmf.cpp:205) */
  const void *A16T;  /* optional, 16-bit forms only: the transposed copy
                        (n x m, ld n) of the 16-bit A, made once by
                        planc_transpose16 -- the reference likewise forms
                        At = A.t() once in the MUNMF constructor
                        (nmf/mu.hpp:42,46).  Null: made in scratch on every
                        call. */
  double normA2;     /* > 0: ||A||_F^2 supplied by the caller, as the reference
                        computes normA once in the NMF constructor
                        (common/nmf.hpp:167); <= 0: computed from d_A on
                        every call. */
  double symm_reg;   /* > 0: symmetric regularisation of ANLS-BPP
                        (applySymmetricReg, common/nmf.hpp:123-129, applied in
                        nmf/bppnmf.hpp:40-44): Gram + symm_reg*I and right-hand
                        side + symm_reg*given^T; needs m == n.  <= 0: off */
} planc_nmf_desc;

/* Runs desc->iters iterations of MUNMF / HALSNMF / BPPNMF / AOADMMNMF /
 * GNSYMNMF computeNMF (nmf/mu.hpp:48-91, nmf/hals.hpp:50-116,
 * nmf/bppnmf.hpp:187-241, nmf/aoadmm.hpp:88-192, nmf/gnsym.hpp:178-260) from
 * the initial W (m x k) and H (n x k), overwriting them, including the
 * constructor/final normalize_by_W of each algorithm.  AO-ADMM keeps its
 * duals (zero at the start of each call) in context scratch and returns
 * PLANC_ENOCONV when a shifted Gram is not positive definite (the
 * reference's arma::chol fails there).  GNSYM factors a square A as H H^T:
 * W is not used (the reference clears it) and is left unchanged.  d_A is FP32 m x n
 * (lda); when dtype is 16-bit, d_A16 (optional) is a planc_convert copy with
 * d_colscale, else one is made in scratch.  h_relerr (host, iters, optional)
 * receives the relative error of printObjective (common/nmf.hpp:461-468). */
int planc_nmf_run(planc_ctx *ctx, const planc_nmf_desc *desc, const float *d_A, long lda,
                  const void *d_A16, const float *d_colscale, float *d_W, float *d_H,
                  double *h_relerr);

/* ---- sparse input (the reference's BUILD_SPARSE build, SP_MAT A) ---------- */
/* A sparse m x n matrix is passed as CSR: d_indptr (m + 1, int64), d_indices
 * (int32 column ids, sorted within a row) and d_values (FP32), all device
 * memory.  At most 2^31 - 1 rows, columns and entries. */

/* CSR of A^T (n x m) from the CSR of A: entries grouped by column of A, rows
 * increasing within a column, values unchanged (bit-exact; the reference
 * forms At = A.t() once in the MUNMF / HALSNMF / BPPNMF constructors,
 * nmf/mu.hpp:42,46).  Outputs hold n + 1 and nnz entries.  Synchronises. */
int planc_csr_transpose(planc_ctx *ctx, long m, long n, const long *d_indptr, const int *d_indices,
                        const float *d_values, long *d_out_indptr, int *d_out_indices, float *d_out_values);

/* d_out (rows x k, column-major, ld ldo) = A (rows x cols, CSR) * F (cols x k,
 * column-major, ld ldf): the sparse A * H of nmf/mu.hpp:71 / nmf/hals.hpp:81
 * (and, with the CSR of A^T, W^T A transposed).  k <= 64; entries summed in
 * CSR order in FP32. */
int planc_spmm(planc_ctx *ctx, long rows, long cols, const long *d_indptr, const int *d_indices,
               const float *d_values, const float *d_F, long ldf, int k, float *d_out, long ldo);

/* computeObjectiveError(A, W, H) for a sparse A (common/utils.hpp:190-241):
 * sqrt(sum over non-zeros of (a_ij - w_i.h_j)^2 + ||W H^T||_F^2 - sum over
 * non-zeros of (w_i.h_j)^2), FP64, into *h_err.  W m x k, H n x k.
 * Synchronises. */
int planc_sparse_fit_error(planc_ctx *ctx, long m, long n, const long *d_indptr, const int *d_indices,
                           const float *d_values, const float *d_W, long ldw, const float *d_H, long ldh,
                           int k, double *h_err);

/* planc_nmf_run for a sparse A (nmf/nmf.cpp:48-130 with BUILD_SPARSE): the
 * same MU / HALS / ANLS-BPP loop with the products W^T A and A H taken as
 * sparse products (FP32 only: desc->dtype must be PLANC_F32) and ||A||^2 over
 * the stored values.  ANLS-BPP starts, as the sparse build does, by
 * normalize_by_W: the reference builds a HALSNMF from W and H "to get proper
 * initializations" and takes its factors without running it
 * (nmf/bppnmf.hpp:195-201), so only HALSNMF's constructor normalisation
 * (nmf/hals.hpp:46) takes effect.  desc->A16T is ignored. */
int planc_nmf_sparse_run(planc_ctx *ctx, const planc_nmf_desc *desc, const long *d_indptr,
                         const int *d_indices, const float *d_values, float *d_W, float *d_H,
                         double *h_relerr);

typedef struct planc_ntf_desc {
  int algo;          /* PLANC_MU, PLANC_HALS, PLANC_ANLSBPP, PLANC_AOADMM, PLANC_NESTEROV */
  int dtype;         /* of the MTTKRP reads of X */
  int nmodes;        /* 2..7 */
  long dims[8];      /* dims[0..nmodes-1] used */
  int k;
  int iters;
  int dimtree;       /* 1: dimension-tree MTTKRP (2 full passes over X per sweep) */
  int compute_error; /* relative error after each iteration (-e 1) */
  const void *X16T;  /* optional, with d_X16 and the dimension tree: the R x L
                        transpose (ld R) of d_X16's L x R unfolding, L = dims[0]
                        * .. * dims[s] with s = planc_ntf_tree_split(); made once
                        by planc_transpose16.  The left contraction then runs as
                        an NN product on it.  Null: the run transposes when it
                        needs to. */
  double normX2;     /* > 0: ||X||_F^2 supplied by the caller, as AUNTF computes
                        the tensor's norm once in its constructor
                        (ntf/auntf.hpp:101); <= 0: computed from d_X on every
                        call that reports the error */
} planc_ntf_desc;

/* The dimension tree's split (AUNTF's dimension tree, ntf/auntf.hpp:123,
 * dimtree/dimtrees.hpp): the left block is modes 0..s, the right block modes
 * s+1..nmodes-1; s in [0, nmodes-2] minimises max(L, R) (L, R the products of
 * the two blocks' dimensions), the reference's nmodes/2 on ties.  Returns -1
 * for bad arguments. */
int planc_ntf_tree_split(int nmodes, const long *dims);

/* Runs AUNTF::computeNTF (ntf/auntf.hpp:128-177) with NTFHALS / NTFMU /
 * NTFANLSBPP / NTFAOADMM (ntf/ntfaoadmm.hpp:23-77, duals zero at the start
 * of each call) / NTFNES (ntf/ntfnes.hpp:9-188: the proximal Nesterov mode
 * update and, after iteration 5, accelerate(); the error is computed every
 * iteration as the reference's accelerated(true) turns it on; iterations run
 * eagerly with one host decision per acceleration; PLANC_ENOCONV if a mode's
 * Nesterov loop passes 100000 steps, which the reference does not bound)
 * updates.  d_factors (host array of device pointers, I_i x k)
 * hold the initial factors, which are first normalised as NCPFactors::
 * normalize() does in the AUNTF constructor (ntf/auntf.hpp:88); on return
 * they hold the unit-norm factors and d_lambda (k, FP64) the weights.
 * d_X16: optional unscaled 16-bit copy of X for dtype != F32 (and desc->X16T
 * its transpose); without it the run makes its own copy, which with the
 * dimension tree carries per-column power-of-two scales over the tree's L x R
 * unfolding (so FP16 holds data beyond its 65504 range; the scales are folded
 * into the Khatri-Rao rows of the right contraction and the output rows of the
 * left one), and the copy's transpose. */
int planc_ntf_run(planc_ctx *ctx, const planc_ntf_desc *desc, const float *d_X, const void *d_X16,
                  float *const *d_factors, double *d_lambda, double *h_relerr);

#ifdef __cplusplus
}
#endif

#endif /* PLANC_HIP_H_ */
