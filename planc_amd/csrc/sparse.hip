// Sparse-input NMF (the reference's BUILD_SPARSE build of nmf/nmf.cpp with
// SP_MAT A): the update rules are those of the dense path, only the two
// products with A change.  A is held as CSR (row pointers int64, column
// indices int32, FP32 values) and, for W^T A, as the CSR of A^T (built once
// per run on the device by a stable sort of the entries by column, the
// analogue of the reference's At = A.t() in the constructors).
//
// Products: out(r, :) = sum_e val(e) F(col(e), :) over the entries of row r
// (SpMM), one wavefront per row, lane c holding output column c; the dense
// factor is read from a row-major (k-major) copy so every entry gathers one
// contiguous row of F.  Entries are summed in CSR order in FP32.
//
// Also the sparse error helper of common/utils.hpp:190-241 (sum over the
// non-zeros of (a_ij - w_i h_j)^2 and of (w_i h_j)^2, plus ||W H^T||_F^2).
#include <algorithm>

#include "common.h"

// After common.h: hipCUB picks its backend from the platform macro that
// hip_runtime.h defines.
#include <hipcub/device/device_radix_sort.hpp>
#include <hipcub/device/device_scan.hpp>

namespace planc {

namespace {

constexpr int kSparseSlot0 = 24;  // context workspace slots 24..31 belong to this file

__global__ __launch_bounds__(256) void csr_entry_rows_kernel(long m, const long *indptr, int *rowid) {
  const long r = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= m) return;
  const int lane = threadIdx.x & 63;
  for (long e = indptr[r] + lane; e < indptr[r + 1]; e += 64) rowid[e] = (int)r;
}

__global__ void iota_kernel(long n, int *v) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) v[e] = (int)e;
}

__global__ void col_count_kernel(long nnz, const int *indices, int *cnt) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nnz) atomicAdd(&cnt[indices[e]], 1);
}

__global__ void set_long_kernel(long *p, long v) { *p = v; }

__global__ void csr_t_gather_kernel(long nnz, const int *ent, const int *rowid, const float *values,
                                    int *out_indices, float *out_values) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nnz) return;
  const int e = ent[p];
  out_indices[p] = rowid[e];
  out_values[p] = values[e];
}

// out(r, c) at out[r * ors + c * ocs] = sum over row r's entries of
// val(e) * F(col(e), c), F k-major (element (j, c) at F[j * ldf + c]).
__global__ __launch_bounds__(256) void spmm_rows_kernel(long rows, const long *indptr, const int *indices,
                                                         const float *values, const float *F, long ldf,
                                                         int k, float *out, long ors, long ocs) {
  const long r = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= rows) return;
  const int lane = threadIdx.x & 63;
  const long b = indptr[r], e = indptr[r + 1];
  const bool active = lane < k;
  float acc = 0.0f;
  for (long base = b; base < e; base += 64) {
    int j = 0;
    float v = 0.0f;
    if (base + lane < e) {
      j = indices[base + lane];
      v = values[base + lane];
    }
    const int cnt = (int)((e - base) < 64 ? (e - base) : 64);
    for (int t = 0; t < cnt; ++t) {
      const int jt = __shfl(j, t, 64);
      const float vt = __shfl(v, t, 64);
      if (active) acc += vt * F[(long)jt * ldf + lane];
    }
  }
  if (active) out[r * ors + (long)lane * ocs] = acc;
}

// Per block: sum over the non-zeros of its rows of (a - t)^2 and t^2 with
// t = w_i . h_j (FP64), W and H k-major.  part[2 * block + {0, 1}].
__global__ __launch_bounds__(256) void sparse_fit_kernel(long m, const long *indptr, const int *indices,
                                                          const float *values, const float *Wk,
                                                          const float *Hk, int k, double *part) {
  __shared__ double red[2][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double sse = 0.0, wh2 = 0.0;
  for (long r = (long)blockIdx.x * 4 + wave; r < m; r += (long)gridDim.x * 4) {
    const double w = lane < k ? (double)Wk[r * k + lane] : 0.0;
    for (long e = indptr[r]; e < indptr[r + 1]; ++e) {
      const int j = indices[e];
      double t = lane < k ? w * (double)Hk[(long)j * k + lane] : 0.0;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
      const double d = (double)values[e] - t;
      sse += d * d;  // identical in every lane
      wh2 += t * t;
    }
  }
  if (lane == 0) {
    red[0][wave] = sse;
    red[1][wave] = wh2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    part[2 * blockIdx.x + 1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
  }
}

// err = sqrt(nnzsse + (||W H^T||^2 - nnzwh)) with ||W H^T||^2 = tr(WtW HtH)
// (the reference takes it as ||R_W R_H^T||_F^2 from economy QRs, the same
// quantity).
__global__ void sparse_fit_finalize_kernel(const double *part, int nb, const double *WtW,
                                           const double *HtH, int k, double *err) {
  if (threadIdx.x != 0) return;
  double sse = 0.0, wh2 = 0.0;
  for (int b = 0; b < nb; ++b) {
    sse += part[2 * b];
    wh2 += part[2 * b + 1];
  }
  double tr = 0.0;
  for (int a = 0; a < k; ++a)
    for (int c = 0; c < k; ++c) tr += WtW[a + (long)k * c] * HtH[c + (long)k * a];
  *err = sqrt(sse + (tr - wh2));
}

}  // namespace

// Stable transpose of a CSR matrix (m x n) into the CSR of its transpose
// (n x m): entries grouped by column, rows increasing within a column.
int csr_transpose(planc_ctx *ctx, long m, long n, long nnz, const long *indptr, const int *indices,
                  const float *values, long *out_indptr, int *out_indices, float *out_values) {
  if (nnz >= (1L << 31) || n >= (1L << 31) || m >= (1L << 31))
    return set_error(ctx, PLANC_EUNSUPPORTED, "csr_transpose: at most 2^31-1 rows, columns and entries");
  if (nnz == 0) {
    PLANC_HIP_CHECK(ctx, hipMemsetAsync(out_indptr, 0, sizeof(long) * (n + 1), ctx->stream));
    return PLANC_OK;
  }
  int bits = 1;
  while (bits < 31 && (1L << bits) < n) ++bits;
  int *rowid = (int *)ws_get(ctx, kSparseSlot0 + 4, sizeof(int) * (size_t)nnz);
  unsigned *keys = (unsigned *)ws_get(ctx, kSparseSlot0 + 5, sizeof(unsigned) * 2 * (size_t)nnz);
  int *ent = (int *)ws_get(ctx, kSparseSlot0 + 6, sizeof(int) * 2 * (size_t)nnz + sizeof(int) * (size_t)(n + 1));
  if (!rowid || !keys || !ent) return set_error(ctx, PLANC_ENOMEM, "csr_transpose: workspace");
  unsigned *keys_s = keys + nnz;
  int *ent_s = ent + nnz;
  int *cnt = ent + 2 * nnz;
  csr_entry_rows_kernel<<<(unsigned)((m + 3) / 4), 256, 0, ctx->stream>>>(m, indptr, rowid);
  PLANC_LAUNCH_CHECK(ctx);
  iota_kernel<<<(unsigned)((nnz + 255) / 256), 256, 0, ctx->stream>>>(nnz, ent);
  PLANC_LAUNCH_CHECK(ctx);
  PLANC_HIP_CHECK(ctx, hipMemcpyAsync(keys, indices, sizeof(int) * (size_t)nnz, hipMemcpyDeviceToDevice,
                                      ctx->stream));
  size_t bytes = 0;
  PLANC_HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys, keys_s, ent, ent_s, (int)nnz, 0,
                                                          bits, ctx->stream));
  size_t scan_bytes = 0;
  PLANC_HIP_CHECK(ctx, hipcub::DeviceScan::ExclusiveScan(nullptr, scan_bytes, cnt, out_indptr, hipcub::Sum(),
                                                         0L, (int)n, ctx->stream));
  void *temp = ws_get(ctx, kSparseSlot0 + 7, std::max(bytes, scan_bytes));
  if (!temp) return set_error(ctx, PLANC_ENOMEM, "csr_transpose: sort workspace");
  PLANC_HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortPairs(temp, bytes, keys, keys_s, ent, ent_s, (int)nnz, 0,
                                                          bits, ctx->stream));
  csr_t_gather_kernel<<<(unsigned)((nnz + 255) / 256), 256, 0, ctx->stream>>>(nnz, ent_s, rowid, values,
                                                                              out_indices, out_values);
  PLANC_LAUNCH_CHECK(ctx);
  PLANC_HIP_CHECK(ctx, hipMemsetAsync(cnt, 0, sizeof(int) * (size_t)n, ctx->stream));
  col_count_kernel<<<(unsigned)((nnz + 255) / 256), 256, 0, ctx->stream>>>(nnz, indices, cnt);
  PLANC_LAUNCH_CHECK(ctx);
  PLANC_HIP_CHECK(ctx, hipcub::DeviceScan::ExclusiveScan(temp, scan_bytes, cnt, out_indptr, hipcub::Sum(), 0L,
                                                         (int)n, ctx->stream));
  set_long_kernel<<<1, 1, 0, ctx->stream>>>(out_indptr + n, nnz);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int spmm_csr(planc_ctx *ctx, long rows, const long *indptr, const int *indices, const float *values,
             const float *F, long ldf, int k, float *out, long ors, long ocs) {
  if (k <= 0 || k > 64) return set_error(ctx, PLANC_EUNSUPPORTED, "spmm: k must be 1..64");
  if (rows <= 0) return PLANC_OK;
  spmm_rows_kernel<<<(unsigned)((rows + 3) / 4), 256, 0, ctx->stream>>>(rows, indptr, indices, values, F, ldf,
                                                                        k, out, ors, ocs);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

// k-major copies of W (m x k) and H (n x k) in workspace slots; returns them.
static int kmajor(planc_ctx *ctx, int slot, const float *X, long rows, long ldx, int k, float **out) {
  float *p = (float *)ws_get(ctx, slot, sizeof(float) * (size_t)rows * k);
  if (!p) return set_error(ctx, PLANC_ENOMEM, "sparse: factor copy");
  *out = p;
  return transpose_to_k_major(ctx, X, rows, ldx, k, p, k);
}

int spmm_colmajor(planc_ctx *ctx, long rows, long cols, const long *indptr, const int *indices,
                  const float *values, const float *F, long ldf, int k, float *out, long ldo) {
  float *Fk = nullptr;
  int rc = kmajor(ctx, kSparseSlot0 + 1, F, cols, ldf, k, &Fk);
  if (rc) return rc;
  return spmm_csr(ctx, rows, indptr, indices, values, Fk, k, k, out, 1, ldo);
}

int sparse_fit_error(planc_ctx *ctx, long m, long n, const long *indptr, const int *indices,
                     const float *values, const float *W, long ldw, const float *H, long ldh, int k,
                     double *d_err) {
  if (k <= 0 || k > 64) return set_error(ctx, PLANC_EUNSUPPORTED, "sparse error: k must be 1..64");
  float *Wk = nullptr, *Hk = nullptr;
  int rc = kmajor(ctx, kSparseSlot0 + 0, W, m, ldw, k, &Wk);
  if (rc) return rc;
  if ((rc = kmajor(ctx, kSparseSlot0 + 1, H, n, ldh, k, &Hk))) return rc;
  const int nb = (int)std::min<long>(1024, std::max<long>(1, (m + 3) / 4));
  double *part = (double *)ws_get(ctx, kSparseSlot0 + 2, sizeof(double) * (2 * (size_t)nb + 2 * (size_t)k * k));
  if (!part) return set_error(ctx, PLANC_ENOMEM, "sparse error: workspace");
  double *WtW = part + 2 * nb, *HtH = WtW + (size_t)k * k;
  sparse_fit_kernel<<<nb, 256, 0, ctx->stream>>>(m, indptr, indices, values, Wk, Hk, k, part);
  PLANC_LAUNCH_CHECK(ctx);
  if ((rc = gram(ctx, W, m, k, ldw, WtW))) return rc;
  if ((rc = gram(ctx, H, n, k, ldh, HtH))) return rc;
  sparse_fit_finalize_kernel<<<1, 64, 0, ctx->stream>>>(part, nb, WtW, HtH, k, d_err);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int sparse_products(planc_ctx *ctx, long m, long n, long nnz, const long *indptr, const int *indices,
                    const float *values, SparseOps *ops) {
  long *tptr = (long *)ws_get(ctx, kSparseSlot0 + 3, sizeof(long) * (size_t)(n + 1) +
                                                         (sizeof(int) + sizeof(float)) * (size_t)std::max(nnz, 1L));
  if (!tptr) return set_error(ctx, PLANC_ENOMEM, "sparse: transposed copy");
  int *tidx = (int *)(tptr + n + 1);
  float *tval = (float *)(tidx + std::max(nnz, 1L));
  int rc = csr_transpose(ctx, m, n, nnz, indptr, indices, values, tptr, tidx, tval);
  if (rc) return rc;
  ops->t_indptr = tptr;
  ops->t_indices = tidx;
  ops->t_values = tval;
  return PLANC_OK;
}

int sparse_wta(planc_ctx *ctx, const SparseOps &o, long m, long n, const float *W, int k, float *out) {
  float *Wk = nullptr;
  int rc = kmajor(ctx, kSparseSlot0 + 0, W, m, m, k, &Wk);
  if (rc) return rc;
  return spmm_csr(ctx, n, o.t_indptr, o.t_indices, o.t_values, Wk, k, k, out, k, 1);
}

int sparse_ah(planc_ctx *ctx, const SparseOps &o, long m, long n, const long *indptr, const int *indices,
              const float *values, const float *H, int k, float *out) {
  float *Hk = nullptr;
  int rc = kmajor(ctx, kSparseSlot0 + 1, H, n, n, k, &Hk);
  if (rc) return rc;
  return spmm_csr(ctx, m, indptr, indices, values, Hk, k, k, out, 1, m);
}

}  // namespace planc
