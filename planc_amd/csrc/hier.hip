// Hierarchical rank-2 NMF (hiernmf/hiernmf.cpp:65-184, hiernmf/node.hpp:46-188)
// with all nodes of a tree level processed together.
//
// A level is the block-diagonal matrix diag(A(:, cols_1), ..., A(:, cols_t))
// of its nodes' column subsets.  It is stored twice on the device:
//  * block CSR over compacted rows: row r = (node t, original row i) for the
//    rows where A(:, cols_t) has entries, columns = level column ids;
//  * level CSC: column lc = original column perm[lc], entries' rows mapped to
//    compacted row ids.
// Both keep the original entry order (rows ascending within a column,
// columns ascending within a row), so every sum is taken in a fixed order.
//
// Rank-2 NMF of every node (DistAUNMF::computeNMF at p = 1 with DistR2's
// update, distnmf/aunmf.hpp:495-657, distnmf/distr2.hpp:35-61): H given W,
// then W given H, 2 x 2 normal equations in FP64.  W keeps only the rows of
// A(:, cols_t) with entries: an empty row of the node's matrix has A H = 0,
// so DistR2 sets that row of W to 0 at the first W update and it stays 0;
// it contributes to W^T W only at iteration 0, whose Gram is taken over all
// m rows of the initial W (drawn from the node's Philox stream).
// The iteration loop of a level is captured once into a hipGraph and
// launched as one graph.
#include <algorithm>
#include <cstring>
#include <map>
#include <queue>
#include <vector>

#include "common.h"
#include "philox.h"

// After common.h: hipCUB picks its backend from the platform macro that
// hip_runtime.h defines.
#include <hipcub/device/device_radix_sort.hpp>
#include <hipcub/device/device_scan.hpp>

namespace planc {

namespace {

// ---------------------------------------------------------------------------
// Total of an exclusive scan: out[n-1] + in[n-1].
__global__ void scan_total(const int *in, const long *out, long n, long *total) {
  if (threadIdx.x == 0) *total = out[n - 1] + in[n - 1];
}

// ---------------------------------------------------------------------------
// Level construction.
__global__ void level_col_len(const long *perm, long nlev, const long *cptr0, int *len) {
  long lc = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (lc >= nlev) return;
  long c = perm[lc];
  len[lc] = (int)(cptr0[c + 1] - cptr0[c]);
}

// Row of every CSR entry (binary search of indptr, one thread per entry).
__global__ void entry_rows(long m, long nnz, const long *rptr0, int *row0) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  long lo = 0, hi = m - 1;  // largest i with rptr0[i] <= e
  while (lo < hi) {
    long mid = (lo + hi + 1) >> 1;
    if (rptr0[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  row0[e] = (int)lo;
}

// Sort key of every entry: its node at this level, nn for entries whose
// column belongs to no node (they sort last and are dropped).
__global__ void level_keys(long nnz, const int *cidx0, const int *colmap, const int *node_of_col,
                           unsigned nn, unsigned *key, int *ent) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  int lc = colmap[cidx0[e]];
  key[e] = lc < 0 ? nn : (unsigned)node_of_col[lc];
  ent[e] = (int)e;
}

// After the stable sort by node, entries are ordered by (node, row, original
// position): position p starts a block-CSR row when its (node, row) differs
// from position p-1.  Also writes the level CSR columns and values.
__global__ void level_run_flags(long lnnz, const unsigned *key, const int *ent, const int *row0,
                                const int *cidx0, const float *val0, const int *colmap, int *flag,
                                int *csr_col, float *csr_val) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= lnnz) return;
  int e = ent[p];
  int f = 1;
  if (p > 0) {
    int e1 = ent[p - 1];
    f = key[p] != key[p - 1] || row0[e] != row0[e1];
  }
  flag[p] = f;
  csr_col[p] = colmap[cidx0[e]];
  csr_val[p] = val0[e];
}

// Compacted row of every level entry, and each row's start, node and
// original row.
__global__ void level_run_rows(long lnnz, const unsigned *key, const int *ent, const int *row0,
                               const int *flag, const long *rstart, int *rid_of_entry, long *row_ptr,
                               int *row_node, int *row_orig) {
  long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= lnnz) return;
  int e = ent[p];
  long r = rstart[p] + flag[p] - 1;
  rid_of_entry[e] = (int)r;
  if (flag[p]) {
    row_ptr[r] = p;
    row_node[r] = (int)key[p];
    row_orig[r] = row0[e];
  }
}

// Level CSC fill, one thread per level column; csc2csr maps a CSC entry of A
// to its CSR entry.
__global__ void level_csc_fill(const long *perm, long nlev, const long *cptr0, const int *csc2csr,
                               const float *cval0, const int *rid_of_entry, const long *col_ptr,
                               int *csc_row, float *csc_val) {
  long lc = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (lc >= nlev) return;
  long c = perm[lc];
  long o = col_ptr[lc];
  for (long e = cptr0[c]; e < cptr0[c + 1]; ++e) {
    csc_row[o] = rid_of_entry[csc2csr[e]];
    csc_val[o] = cval0[e];
    ++o;
  }
}

// ---------------------------------------------------------------------------
// Rank-2 normal equations (DistR2::update, distnmf/distr2.hpp:35-61): LU
// with partial pivoting of the 2x2 Gram (arma::solve fast), then columns with
// a negative entry replaced by the better single-variable solution.
__device__ inline void r2_solve(const double *G, double r0, double r1, double &x0, double &x1) {
  double a = G[0], c = G[1], b = G[2], d = G[3];
  bool sw = fabs(c) > fabs(a);
  double u11 = sw ? c : a, u12 = sw ? d : b;
  double l21 = (sw ? a : c) / u11;
  double u22 = (sw ? b : d) - l21 * u12;
  double z0 = sw ? r1 : r0, z1 = (sw ? r0 : r1) - l21 * z0;
  double g1 = z1 / u22;
  double g0 = (z0 - u12 * g1) / u11;
  if (g0 < 0 || g1 < 0) {
    double u = r0 / a, v = r1 / d;
    if (sqrt(a) * u >= sqrt(d) * v)
      v = 0;
    else
      u = 0;
    g0 = u;
    g1 = v;
  }
  x0 = g0;
  x1 = g1;
}

// One rank-2 update of n columns: X(:, j) = r2_solve(G, R(:, j)).
__global__ void r2_update_kernel(const double *G, const float *R, long r_js, long r_cs, long n, float *X,
                                 long x_js, long x_cs) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double x0, x1;
  r2_solve(G, (double)R[j * r_js], (double)R[j * r_js + r_cs], x0, x1);
  X[j * x_js] = (float)x0;
  X[j * x_js + x_cs] = (float)x1;
}

struct Chunk {
  int node;
  long b, e;
};

// Partial Grams of 2-column factors over chunks of contiguous rows.
__global__ void gram2_chunks(const Chunk *ch, const double *X, double *part) {
  __shared__ double red[3][4];
  Chunk c = ch[blockIdx.x];
  double s00 = 0, s01 = 0, s11 = 0;
  for (long r = c.b + threadIdx.x; r < c.e; r += blockDim.x) {
    double x0 = X[2 * r], x1 = X[2 * r + 1];
    s00 += x0 * x0;
    s01 += x0 * x1;
    s11 += x1 * x1;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s00 += __shfl_xor(s00, off, 64);
    s01 += __shfl_xor(s01, off, 64);
    s11 += __shfl_xor(s11, off, 64);
  }
  int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = s00;
    red[1][w] = s01;
    red[2][w] = s11;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t0 = 0, t1 = 0, t2 = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
      t0 += red[0][q];
      t1 += red[1][q];
      t2 += red[2][q];
    }
    part[3 * blockIdx.x] = t0;
    part[3 * blockIdx.x + 1] = t1;
    part[3 * blockIdx.x + 2] = t2;
  }
}

// Node Gram (column-major 2x2) from its chunks [cb[t], cb[t+1]).
__global__ void gram2_finalize(const double *part, const int *cb, int nn, double *G) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nn) return;
  double a = 0, b = 0, c = 0;
  for (int q = cb[t]; q < cb[t + 1]; ++q) {
    a += part[3 * q];
    b += part[3 * q + 1];
    c += part[3 * q + 2];
  }
  G[4 * t] = a;
  G[4 * t + 1] = b;
  G[4 * t + 2] = b;
  G[4 * t + 3] = c;
}

// Gram over all m rows of the initial W of every node (iteration 0).
__global__ void init_gram_chunks(const Chunk *ch, const unsigned long long *seeds, long m,
                                 double *part) {
  __shared__ double red[3][4];
  Chunk c = ch[blockIdx.x];
  uint64_t sd = seeds[c.node];
  double s00 = 0, s01 = 0, s11 = 0;
  for (long i = c.b + threadIdx.x; i < c.e; i += blockDim.x) {
    double x0 = planc_rng::u01(planc_rng::word(sd, 0, (uint64_t)i));
    double x1 = planc_rng::u01(planc_rng::word(sd, 0, (uint64_t)(i + m)));
    s00 += x0 * x0;
    s01 += x0 * x1;
    s11 += x1 * x1;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s00 += __shfl_xor(s00, off, 64);
    s01 += __shfl_xor(s01, off, 64);
    s11 += __shfl_xor(s11, off, 64);
  }
  int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = s00;
    red[1][w] = s01;
    red[2][w] = s11;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t0 = 0, t1 = 0, t2 = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
      t0 += red[0][q];
      t1 += red[1][q];
      t2 += red[2][q];
    }
    part[3 * blockIdx.x] = t0;
    part[3 * blockIdx.x + 1] = t1;
    part[3 * blockIdx.x + 2] = t2;
  }
}

// Initial factors from each node's Philox stream: W(i, c) word i + m c of
// stream 0, H(j, c) word j + n_t c of stream 1 (j local column).
__global__ void init_factors(long nrows, const int *row_node, const int *row_orig, long nlev,
                             const int *node_of_col, const long *node_col0,
                             const unsigned long long *seeds, long m, double *W, double *H) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nrows) {
    uint64_t sd = seeds[row_node[e]];
    long i = row_orig[e];
    W[2 * e] = planc_rng::u01(planc_rng::word(sd, 0, (uint64_t)i));
    W[2 * e + 1] = planc_rng::u01(planc_rng::word(sd, 0, (uint64_t)(i + m)));
  } else if (e < nrows + nlev) {
    long lc = e - nrows;
    int t = node_of_col[lc];
    long j = lc - node_col0[t], nt = node_col0[t + 1] - node_col0[t];
    uint64_t sd = seeds[t];
    H[2 * lc] = planc_rng::u01(planc_rng::word(sd, 1, (uint64_t)j));
    H[2 * lc + 1] = planc_rng::u01(planc_rng::word(sd, 1, (uint64_t)(j + nt)));
  }
}

// H update: WtA(:, lc) over the column's entries, then the R2 rule with the
// node's W^T W.
// Sums over the entries b, b+S, b+2S, ... < e1 of val * X[idx], taken in
// that order; four entries' index/value loads are issued before their
// gathers so each step waits on one memory round trip, not four.
__device__ inline void dot2_range(long b, long e1, long S, const int *__restrict__ idx,
                                  const float *__restrict__ val, const double2 *__restrict__ X, double &s0,
                                  double &s1) {
  long e = b;
  for (; e + 3 * S < e1; e += 4 * S) {
    int i0 = idx[e], i1 = idx[e + S], i2 = idx[e + 2 * S], i3 = idx[e + 3 * S];
    double v0 = val[e], v1 = val[e + S], v2 = val[e + 2 * S], v3 = val[e + 3 * S];
    double2 x0 = X[i0], x1 = X[i1], x2 = X[i2], x3 = X[i3];
    s0 += v0 * x0.x;
    s1 += v0 * x0.y;
    s0 += v1 * x1.x;
    s1 += v1 * x1.y;
    s0 += v2 * x2.x;
    s1 += v2 * x2.y;
    s0 += v3 * x3.x;
    s1 += v3 * x3.y;
  }
  for (; e < e1; e += S) {
    double v = val[e];
    double2 x = X[idx[e]];
    s0 += v * x.x;
    s1 += v * x.y;
  }
}
__device__ inline double dot1_range(long b, long e1, long S, const int *__restrict__ idx,
                                    const float *__restrict__ val, const double *__restrict__ X) {
  double s = 0;
  long e = b;
  for (; e + 3 * S < e1; e += 4 * S) {
    int i0 = idx[e], i1 = idx[e + S], i2 = idx[e + 2 * S], i3 = idx[e + 3 * S];
    double v0 = val[e], v1 = val[e + S], v2 = val[e + 2 * S], v3 = val[e + 3 * S];
    double x0 = X[i0], x1 = X[i1], x2 = X[i2], x3 = X[i3];
    s += v0 * x0;
    s += v1 * x1;
    s += v2 * x2;
    s += v3 * x3;
  }
  for (; e < e1; e += S) s += (double)val[e] * X[idx[e]];
  return s;
}

__global__ void h_update(long nlev, const long *col_ptr, const int *csc_row, const float *csc_val,
                         const int *node_of_col, const double *W, const double *Gw, double *H) {
  long lc = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (lc >= nlev) return;
  double s0 = 0, s1 = 0;
  dot2_range(col_ptr[lc], col_ptr[lc + 1], 1, csc_row, csc_val, (const double2 *)W, s0, s1);
  double x0, x1;
  r2_solve(Gw + 4 * node_of_col[lc], s0, s1, x0, x1);
  H[2 * lc] = x0;
  H[2 * lc + 1] = x1;
}

// W update: (A H)(r, :) over the compacted row's entries, then the R2 rule
// with the node's H^T H.  G lanes per row (a sub-wave group or a whole
// wave, G in {4, 16, 64}, picked per level by timing); rows longer than `longlen` are
// left to the segmented long-row kernels -- the power-law row
// popularity of term-document matrices puts a few very long rows next to
// millions of short ones.
template <int G>
__device__ inline double group_sum(double v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
  return v;
}

template <int G>
__global__ void w_update_short(long nrows, const long *row_ptr, const int *csr_col,
                               const float *csr_val, const int *row_node, const double *H,
                               const double *Gh, double *W, long longlen) {
  long g = ((long)blockIdx.x * blockDim.x + threadIdx.x) / G;
  int lane = threadIdx.x % G;
  if (g >= nrows) return;
  long b = row_ptr[g], e1 = row_ptr[g + 1];
  if (e1 - b > longlen) return;
  double s0 = 0, s1 = 0;
  dot2_range(b + lane, e1, G, csr_col, csr_val, (const double2 *)H, s0, s1);
  s0 = group_sum<G>(s0);
  s1 = group_sum<G>(s1);
  if (lane == 0) {
    double x0, x1;
    r2_solve(Gh + 4 * row_node[g], s0, s1, x0, x1);
    W[2 * g] = x0;
    W[2 * g + 1] = x1;
  }
}

__device__ inline double block_sum256(double v, double *red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

// Long rows are cut into segments of at most SEG_LEN entries (a block
// each, partial sums to seg_part); the per-row finish adds a row's segment
// partials in segment order, so the sums are taken in a fixed order.
constexpr long SEG_LEN = 4096;

__global__ void w_update_seg(const int *seg_row, const long *seg_b, const int *list, const long *row_ptr,
                             const int *csr_col, const float *csr_val, const double *H, double *seg_part) {
  __shared__ double red[4];
  long r = list[seg_row[blockIdx.x]];
  long b = seg_b[blockIdx.x];
  long e1 = min(b + SEG_LEN, row_ptr[r + 1]);
  double s0 = 0, s1 = 0;
  dot2_range(b + threadIdx.x, e1, blockDim.x, csr_col, csr_val, (const double2 *)H, s0, s1);
  s0 = block_sum256(s0, red);
  s1 = block_sum256(s1, red);
  if (threadIdx.x == 0) {
    seg_part[2 * blockIdx.x] = s0;
    seg_part[2 * blockIdx.x + 1] = s1;
  }
}
__global__ void w_update_long_fin(long nlong, const int *list, const long *seg_off, const double *seg_part,
                                  const int *row_node, const double *Gh, double *W) {
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nlong) return;
  long r = list[j];
  double s0 = 0, s1 = 0;
  for (long q = seg_off[j]; q < seg_off[j + 1]; ++q) {
    s0 += seg_part[2 * q];
    s1 += seg_part[2 * q + 1];
  }
  double x0, x1;
  r2_solve(Gh + 4 * row_node[r], s0, s1, x0, x1);
  W[2 * r] = x0;
  W[2 * r + 1] = x1;
}

__global__ void flag_long(const long *row_ptr, long nrows, long longlen, int *flag) {
  long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nrows) flag[r] = (row_ptr[r + 1] - row_ptr[r]) > longlen;
}

__global__ void compact_list(const int *flag, const long *pos, long n, int *list) {
  long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n && flag[r]) list[pos[r]] = (int)r;
}

__global__ void long_seg_count(const int *list, const long *row_ptr, long nlong, int *cnt) {
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nlong) return;
  long r = list[j];
  cnt[j] = (int)((row_ptr[r + 1] - row_ptr[r] + SEG_LEN - 1) / SEG_LEN);
}

__global__ void long_seg_fill(const int *list, const long *row_ptr, long nlong, const long *seg_off,
                              int *seg_row, long *seg_b) {
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nlong) return;
  long r = list[j];
  long q = seg_off[j];
  for (long b = row_ptr[r]; b < row_ptr[r + 1]; b += SEG_LEN, ++q) {
    seg_row[q] = (int)j;
    seg_b[q] = b;
  }
}

// Power iteration on A_t^T A_t (hiernmf/matutils.hpp:37-120): z = A q / s,
// q = A^T z, s = ||q|| per node.
__global__ void pow_init(long nlev, const int *node_of_col, const long *node_col0,
                         const unsigned long long *seeds, double *q) {
  long lc = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (lc >= nlev) return;
  int t = node_of_col[lc];
  q[lc] = planc_rng::normal(seeds[t], 2, (uint64_t)(lc - node_col0[t]));
}

template <int G>
__global__ void pow_rows_short(long nrows, const long *row_ptr, const int *csr_col,
                               const float *csr_val, const int *row_node, const double *q,
                               const double *sig, double *z, long longlen) {
  long g = ((long)blockIdx.x * blockDim.x + threadIdx.x) / G;
  int lane = threadIdx.x % G;
  if (g >= nrows) return;
  long b = row_ptr[g], e1 = row_ptr[g + 1];
  if (e1 - b > longlen) return;
  double s = dot1_range(b + lane, e1, G, csr_col, csr_val, q);
  s = group_sum<G>(s);
  if (lane == 0) z[g] = s / sig[row_node[g]];
}

__global__ void pow_rows_seg(const int *seg_row, const long *seg_b, const int *list, const long *row_ptr,
                             const int *csr_col, const float *csr_val, const double *q, double *seg_part) {
  __shared__ double red[4];
  long r = list[seg_row[blockIdx.x]];
  long b = seg_b[blockIdx.x];
  long e1 = min(b + SEG_LEN, row_ptr[r + 1]);
  double s = dot1_range(b + threadIdx.x, e1, blockDim.x, csr_col, csr_val, q);
  s = block_sum256(s, red);
  if (threadIdx.x == 0) seg_part[blockIdx.x] = s;
}
__global__ void pow_rows_long_fin(long nlong, const int *list, const long *seg_off, const double *seg_part,
                                  const int *row_node, const double *sig, double *z) {
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nlong) return;
  long r = list[j];
  double s = 0;
  for (long q = seg_off[j]; q < seg_off[j + 1]; ++q) s += seg_part[q];
  z[r] = s / sig[row_node[r]];
}

__global__ void pow_cols(long nlev, const long *col_ptr, const int *csc_row, const float *csc_val,
                         const double *z, double *q) {
  long lc = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (lc >= nlev) return;
  q[lc] = dot1_range(col_ptr[lc], col_ptr[lc + 1], 1, csc_row, csc_val, z);
}

__global__ void sumsq_chunks(const Chunk *ch, const double *x, double *part) {
  __shared__ double red[4];
  Chunk c = ch[blockIdx.x];
  double s = 0;
  for (long r = c.b + threadIdx.x; r < c.e; r += blockDim.x) s += x[r] * x[r];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += red[q];
    part[blockIdx.x] = t;
  }
}

__global__ void norm_finalize(const double *part, const int *cb, int nn, double *sig) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nn) return;
  double s = 0;
  for (int q = cb[t]; q < cb[t + 1]; ++q) s += part[q];
  sig[t] = sqrt(s);
}

__global__ void fill_ones(double *x, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) x[t] = 1.0;
}

__global__ void split_mask(long nlev, const double *H, unsigned char *mask) {
  long lc = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (lc < nlev) mask[lc] = H[2 * lc] >= H[2 * lc + 1];
}

__global__ void scatter_w(long nrows, const int *row_orig, const double *Wc, long m, int slot_node,
                          const int *row_node, double *Wfull) {
  long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows || row_node[r] != slot_node) return;
  long i = row_orig[r];
  Wfull[i] = Wc[2 * r];
  Wfull[i + m] = Wc[2 * r + 1];
}

__global__ void subset_count(long m, const long *indptr, const int *indices, const int *colmap, int *len) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  int c = 0;
  for (long e = indptr[i]; e < indptr[i + 1]; ++e) c += colmap[indices[e]] >= 0;
  len[i] = c;
}

__global__ void subset_fill(long m, const long *indptr, const int *indices, const float *data,
                            const int *colmap, const long *optr, int *oidx, float *oval) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  long o = optr[i];
  for (long e = indptr[i]; e < indptr[i + 1]; ++e) {
    int lc = colmap[indices[e]];
    if (lc < 0) continue;
    oidx[o] = lc;
    oval[o] = data[e];
    ++o;
  }
}

inline unsigned nb(long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// Device buffer owned by the level engine.
template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  int get(planc_ctx *ctx, size_t want) {
    if (want == 0) want = 1;
    if (n >= want) return PLANC_OK;
    if (p) {
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
    if (hipMalloc(&p, sizeof(T) * want) != hipSuccess) return set_error(ctx, PLANC_ENOMEM, "hiernmf: device memory");
    n = want;
    return PLANC_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

struct Engine {
  planc_ctx *ctx;
  long m = 0, n = 0, nnz = 0;
  DBuf<long> rptr0, cptr0;
  DBuf<int> cidx0, ridx0;
  DBuf<float> val0, cval0;
  // level
  DBuf<long> perm, col_ptr, row_ptr, rstart, bsum, node_col0;
  DBuf<int> colmap, node_of_col, flag, row_node, row_orig, row_len, col_len, csr_col, csc_row,
      cb_rows, cb_cols, cb_init, lflag, llist, row0, csc2csr, ent, ent_s, rid_of_entry, lcnt, seg_row;
  DBuf<long> seg_off, seg_b;
  DBuf<double> seg_part;
  DBuf<unsigned> key, key_s;
  DBuf<unsigned char> temp;
  DBuf<long> lpos;
  DBuf<double> Wscratch;
  DBuf<float> csr_val, csc_val;
  DBuf<double> W, H, Gw, Gh, part, q, z, sig;
  DBuf<unsigned long long> seeds;
  DBuf<Chunk> ch_rows, ch_cols, ch_init;
  DBuf<unsigned char> mask;

  ~Engine() {
    rptr0.release(); cptr0.release(); cidx0.release(); ridx0.release(); val0.release();
    cval0.release(); perm.release(); col_ptr.release(); row_ptr.release(); rstart.release();
    bsum.release(); node_col0.release(); colmap.release(); node_of_col.release();
    flag.release(); row_node.release(); row_orig.release(); row_len.release();
    col_len.release(); csr_col.release(); csc_row.release(); cb_rows.release();
    cb_cols.release(); cb_init.release(); lflag.release(); llist.release(); row0.release();
    csc2csr.release(); ent.release(); ent_s.release(); rid_of_entry.release(); key.release();
    key_s.release(); temp.release(); csr_val.release(); csc_val.release(); W.release();
    H.release(); Gw.release(); Gh.release(); part.release(); q.release(); z.release();
    sig.release(); seeds.release(); ch_rows.release(); ch_cols.release(); ch_init.release();
    mask.release(); lpos.release(); Wscratch.release(); lcnt.release(); seg_row.release();
    seg_off.release(); seg_b.release(); seg_part.release();
  }

  // Launches of the row kernels with G lanes per row.
  void launch_w_short(int G, long nrows, double *Wout, long longlen) {
    long thr = nrows * G;
    unsigned nbk = (unsigned)((thr + 255) / 256);
    if (nbk == 0) return;
    if (G == 64)
      w_update_short<64><<<nbk, 256, 0, ctx->stream>>>(nrows, row_ptr.p, csr_col.p, csr_val.p, row_node.p, H.p, Gh.p, Wout, longlen);
    else if (G == 4)
      w_update_short<4><<<nbk, 256, 0, ctx->stream>>>(nrows, row_ptr.p, csr_col.p, csr_val.p, row_node.p, H.p, Gh.p, Wout, longlen);
    else
      w_update_short<16><<<nbk, 256, 0, ctx->stream>>>(nrows, row_ptr.p, csr_col.p, csr_val.p, row_node.p, H.p, Gh.p, Wout, longlen);
  }
  void launch_pow_short(int G, long nrows, long longlen) {
    long thr = nrows * G;
    unsigned nbk = (unsigned)((thr + 255) / 256);
    if (nbk == 0) return;
    if (G == 64)
      pow_rows_short<64><<<nbk, 256, 0, ctx->stream>>>(nrows, row_ptr.p, csr_col.p, csr_val.p, row_node.p, q.p, sig.p, z.p, longlen);
    else if (G == 4)
      pow_rows_short<4><<<nbk, 256, 0, ctx->stream>>>(nrows, row_ptr.p, csr_col.p, csr_val.p, row_node.p, q.p, sig.p, z.p, longlen);
    else
      pow_rows_short<16><<<nbk, 256, 0, ctx->stream>>>(nrows, row_ptr.p, csr_col.p, csr_val.p, row_node.p, q.p, sig.p, z.p, longlen);
  }
  // Lanes per row for this level: the fastest of {4, 16, 64} over a few
  // timed launches of the W update into a scratch buffer.
  int tune_group(long nrows, long longlen) {
    if (nrows == 0) return 16;
    if (Wscratch.get(ctx, 2 * nrows)) return 16;
    const int cand[3] = {4, 16, 64};
    float best = 1e30f;
    int bestg = 16;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return 16;
    if (hipEventCreate(&e1) != hipSuccess) {
      (void)hipEventDestroy(e0);
      return 16;
    }
    for (int c : cand) {
      launch_w_short(c, nrows, Wscratch.p, longlen);
      (void)hipEventRecord(e0, ctx->stream);
      for (int rep = 0; rep < 3; ++rep) launch_w_short(c, nrows, Wscratch.p, longlen);
      (void)hipEventRecord(e1, ctx->stream);
      (void)hipEventSynchronize(e1);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) {
        best = ms;
        bestg = c;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return bestg;
  }

  // out[i] = sum_{j < i} in[j] (64-bit sums); returns the total (synchronises).
  int scan(const int *in, long cnt, long *out, long *total) {
    if (cnt <= 0) {
      *total = 0;
      return PLANC_OK;
    }
    size_t bytes = 0;
    PLANC_HIP_CHECK(ctx, hipcub::DeviceScan::ExclusiveScan(nullptr, bytes, in, out, hipcub::Sum(), 0L, cnt,
                                                           ctx->stream));
    int rc = temp.get(ctx, bytes);
    if (rc) return rc;
    if ((rc = bsum.get(ctx, 1))) return rc;
    PLANC_HIP_CHECK(ctx, hipcub::DeviceScan::ExclusiveScan(temp.p, bytes, in, out, hipcub::Sum(), 0L, cnt,
                                                           ctx->stream));
    scan_total<<<1, 64, 0, ctx->stream>>>(in, out, cnt, bsum.p);
    PLANC_LAUNCH_CHECK(ctx);
    PLANC_HIP_CHECK(ctx, hipMemcpyAsync(total, bsum.p, sizeof(long), hipMemcpyDeviceToHost, ctx->stream));
    PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return PLANC_OK;
  }
  // Stable sort of (key, ent) pairs by the low `bits` bits of key.
  int sort_pairs(long cnt, int bits) {
    size_t bytes = 0;
    PLANC_HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, key.p, key_s.p, ent.p, ent_s.p,
                                                            cnt, 0, bits, ctx->stream));
    int rc = temp.get(ctx, bytes);
    if (rc) return rc;
    PLANC_HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortPairs(temp.p, bytes, key.p, key_s.p, ent.p, ent_s.p,
                                                            cnt, 0, bits, ctx->stream));
    return PLANC_OK;
  }
};

// Chunks of at most `chunk` consecutive items per node (node t owns
// [off[t], off[t+1])); cb[t] is the first chunk of node t.
void make_chunks(const std::vector<long> &off, long chunk, std::vector<Chunk> &ch, std::vector<int> &cb) {
  ch.clear();
  cb.assign(off.size(), 0);
  for (size_t t = 0; t + 1 < off.size(); ++t) {
    cb[t] = (int)ch.size();
    for (long b = off[t]; b < off[t + 1]; b += chunk) ch.push_back({(int)t, b, std::min(off[t + 1], b + chunk)});
  }
  cb[off.size() - 1] = (int)ch.size();
}

}  // namespace

}  // namespace planc

using namespace planc;

extern "C" {

int planc_hiernmf_run(planc_ctx *ctx, const planc_hier_desc *desc, const long *h_indptr,
                      const int *h_indices, const float *h_data, planc_hier_node *h_nodes,
                      int *h_num_nodes, long *h_cols, double *h_H, double *h_W) {
  if (!ctx || !desc || !h_indptr || !h_indices || !h_data || !h_nodes || !h_num_nodes || !h_cols)
    return PLANC_EINVAL;
  const long m = desc->m, n = desc->n;
  const int depth = desc->depth, iters = desc->iters;
  // Node indices 2i+1 / 2i+2 stay below 2^63 up to level 61; the complete
  // tree's table has 2^depth - 1 entries.
  if (m <= 0 || n <= 0 || depth < 1 || depth > (desc->nodes > 0 ? 62 : 24) || iters < 0 ||
      (desc->nodes > 0 && depth < 2))
    return set_error(ctx, PLANC_EINVAL, "planc_hiernmf_run: bad arguments");
  const long nnz = h_indptr[m];
  if (nnz < 0 || nnz >= (1L << 31)) return set_error(ctx, PLANC_EINVAL, "planc_hiernmf_run: nnz");
  for (long i = 0; i < m; ++i)
    if (h_indptr[i + 1] < h_indptr[i]) return set_error(ctx, PLANC_EINVAL, "planc_hiernmf_run: indptr");
  for (long e = 0; e < nnz; ++e)
    if (h_indices[e] < 0 || h_indices[e] >= n) return set_error(ctx, PLANC_EINVAL, "planc_hiernmf_run: column index");
  PLANC_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  Engine E;
  E.ctx = ctx;
  E.m = m;
  E.n = n;
  E.nnz = nnz;
  int rc;
#define RC(x)          \
  do {                 \
    rc = (x);          \
    if (rc) return rc; \
  } while (0)
  // CSC of A on the host (counting sort, rows ascending within a column).
  std::vector<long> cptr(n + 1, 0);
  for (long e = 0; e < nnz; ++e) cptr[h_indices[e] + 1]++;
  for (long j = 0; j < n; ++j) cptr[j + 1] += cptr[j];
  std::vector<int> ridx(nnz), c2r(nnz);
  std::vector<float> cval(nnz);
  {
    std::vector<long> cur(cptr.begin(), cptr.end() - 1);
    for (long i = 0; i < m; ++i)
      for (long e = h_indptr[i]; e < h_indptr[i + 1]; ++e) {
        long p = cur[h_indices[e]]++;
        ridx[p] = (int)i;
        cval[p] = h_data[e];
        c2r[p] = (int)e;
      }
  }
  RC(E.rptr0.get(ctx, m + 1));
  RC(E.cptr0.get(ctx, n + 1));
  RC(E.cidx0.get(ctx, nnz));
  RC(E.ridx0.get(ctx, nnz));
  RC(E.val0.get(ctx, nnz));
  RC(E.cval0.get(ctx, nnz));
  RC(E.csc2csr.get(ctx, nnz));
  RC(E.row0.get(ctx, nnz));
  RC(E.key.get(ctx, nnz));
  RC(E.key_s.get(ctx, nnz));
  RC(E.ent.get(ctx, nnz));
  RC(E.ent_s.get(ctx, nnz));
  RC(E.rid_of_entry.get(ctx, nnz));
  RC(E.flag.get(ctx, nnz));
  RC(E.rstart.get(ctx, nnz));
  auto up = [&](void *d, const void *h, size_t b) {
    return hipMemcpyAsync(d, h, b, hipMemcpyHostToDevice, ctx->stream);
  };
  PLANC_HIP_CHECK(ctx, up(E.rptr0.p, h_indptr, sizeof(long) * (m + 1)));
  PLANC_HIP_CHECK(ctx, up(E.cptr0.p, cptr.data(), sizeof(long) * (n + 1)));
  if (nnz) {
    PLANC_HIP_CHECK(ctx, up(E.cidx0.p, h_indices, sizeof(int) * nnz));
    PLANC_HIP_CHECK(ctx, up(E.ridx0.p, ridx.data(), sizeof(int) * nnz));
    PLANC_HIP_CHECK(ctx, up(E.val0.p, h_data, sizeof(float) * nnz));
    PLANC_HIP_CHECK(ctx, up(E.cval0.p, cval.data(), sizeof(float) * nnz));
    PLANC_HIP_CHECK(ctx, up(E.csc2csr.p, c2r.data(), sizeof(int) * nnz));
    entry_rows<<<nb(nnz), 256, 0, ctx->stream>>>(m, nnz, E.rptr0.p, E.row0.p);
    PLANC_LAUNCH_CHECK(ctx);
  }

  // Tree state.
  struct HNode {
    long index;
    int level;
    std::vector<long> cols;
    double sigma = 0, score = 0;
    int lvalid = 0, rvalid = 0;
    long lchild = -1, rchild = -1;
    long out_slot = -1;  // column offset in h_cols / h_H (-1: not recorded yet)
    bool split = false;  // rank-2 NMF done, children created
    int lidx = 0;        // factor column of the left child (1 when the sides were swapped)
    long parent = -1;
  };
  std::vector<HNode> nodes;
  nodes.push_back(HNode{0, 0, {}});
  nodes[0].cols.resize(n);
  for (long j = 0; j < n; ++j) nodes[0].cols[j] = j;
  long cols_used = 0;
  const long cols_cap = n * (long)depth;
  const unsigned long long seed0 = (unsigned long long)(long long)desc->initseed;
  std::map<long, std::vector<double>> wstore;  // node id -> its rank-2 W (m x 2), when h_W is asked for

  auto run_batch = [&](const std::vector<long> &level_nodes, bool sigma_on, bool split_on,
                       std::vector<long> *children) -> int {
    if (level_nodes.empty()) return PLANC_OK;
    const int nn = (int)level_nodes.size();
    // Level columns: concatenation of the nodes' (sorted) column lists.
    std::vector<long> perm, noff(nn + 1, 0);
    for (int t = 0; t < nn; ++t) {
      const auto &c = nodes[level_nodes[t]].cols;
      perm.insert(perm.end(), c.begin(), c.end());
      noff[t + 1] = noff[t] + (long)c.size();
    }
    const long nlev = (long)perm.size();
    std::vector<int> colmap(n, -1), nodeof(nlev);
    for (int t = 0; t < nn; ++t)
      for (long lc = noff[t]; lc < noff[t + 1]; ++lc) {
        colmap[perm[lc]] = (int)lc;
        nodeof[lc] = t;
      }
    std::vector<unsigned long long> seeds(nn);
    for (int t = 0; t < nn; ++t)  // set_seed(initseed * index + rank), hiernmf/node.hpp:123
      seeds[t] = seed0 * (unsigned long long)nodes[level_nodes[t]].index;
    RC(E.perm.get(ctx, nlev));
    RC(E.colmap.get(ctx, n));
    RC(E.node_of_col.get(ctx, nlev));
    RC(E.node_col0.get(ctx, nn + 1));
    RC(E.seeds.get(ctx, nn));
    PLANC_HIP_CHECK(ctx, up(E.perm.p, perm.data(), sizeof(long) * nlev));
    PLANC_HIP_CHECK(ctx, up(E.colmap.p, colmap.data(), sizeof(int) * n));
    PLANC_HIP_CHECK(ctx, up(E.node_of_col.p, nodeof.data(), sizeof(int) * nlev));
    PLANC_HIP_CHECK(ctx, up(E.node_col0.p, noff.data(), sizeof(long) * (nn + 1)));
    PLANC_HIP_CHECK(ctx, up(E.seeds.p, seeds.data(), sizeof(unsigned long long) * nn));
    // Level CSC pointers.
    RC(E.col_len.get(ctx, nlev));
    RC(E.col_ptr.get(ctx, nlev + 1));
    level_col_len<<<nb(nlev), 256, 0, ctx->stream>>>(E.perm.p, nlev, E.cptr0.p, E.col_len.p);
    long lnnz = 0;
    RC(E.scan(E.col_len.p, nlev, E.col_ptr.p, &lnnz));
    PLANC_HIP_CHECK(ctx, up(E.col_ptr.p + nlev, &lnnz, sizeof(long)));
    // Block-CSR rows: a stable sort of the entries by node keeps (row,
    // original position) order inside each node; runs of equal (node, row)
    // are the compacted rows, numbered node-major.
    long nrows = 0;
    RC(E.csr_col.get(ctx, lnnz));
    RC(E.csr_val.get(ctx, lnnz));
    RC(E.csc_row.get(ctx, lnnz));
    RC(E.csc_val.get(ctx, lnnz));
    if (lnnz > 0) {
      int bits = 1;
      while (bits < 31 && (1L << bits) <= (long)nn) ++bits;
      level_keys<<<nb(nnz), 256, 0, ctx->stream>>>(nnz, E.cidx0.p, E.colmap.p, E.node_of_col.p,
                                                   (unsigned)nn, E.key.p, E.ent.p);
      PLANC_LAUNCH_CHECK(ctx);
      RC(E.sort_pairs(nnz, bits));
      level_run_flags<<<nb(lnnz), 256, 0, ctx->stream>>>(lnnz, E.key_s.p, E.ent_s.p, E.row0.p, E.cidx0.p,
                                                          E.val0.p, E.colmap.p, E.flag.p, E.csr_col.p,
                                                          E.csr_val.p);
      PLANC_LAUNCH_CHECK(ctx);
      RC(E.scan(E.flag.p, lnnz, E.rstart.p, &nrows));
    }
    RC(E.row_node.get(ctx, nrows));
    RC(E.row_orig.get(ctx, nrows));
    RC(E.row_ptr.get(ctx, nrows + 1));
    if (lnnz > 0)
      level_run_rows<<<nb(lnnz), 256, 0, ctx->stream>>>(lnnz, E.key_s.p, E.ent_s.p, E.row0.p, E.flag.p,
                                                         E.rstart.p, E.rid_of_entry.p, E.row_ptr.p,
                                                         E.row_node.p, E.row_orig.p);
    PLANC_HIP_CHECK(ctx, up(E.row_ptr.p + nrows, &lnnz, sizeof(long)));
    level_csc_fill<<<nb(nlev), 256, 0, ctx->stream>>>(E.perm.p, nlev, E.cptr0.p, E.csc2csr.p, E.cval0.p,
                                                     E.rid_of_entry.p, E.col_ptr.p, E.csc_row.p,
                                                     E.csc_val.p);
    PLANC_LAUNCH_CHECK(ctx);
    // Rows longer than LONGLEN are cut into SEG_LEN segments, a block each
    // (lists built on the device).
    const long LONGLEN = 2048;
    RC(E.lflag.get(ctx, nrows ? nrows : 1));
    RC(E.lpos.get(ctx, nrows ? nrows : 1));
    long nlong = 0;
    if (nrows) {
      flag_long<<<nb(nrows), 256, 0, ctx->stream>>>(E.row_ptr.p, nrows, LONGLEN, E.lflag.p);
      RC(E.scan(E.lflag.p, nrows, E.lpos.p, &nlong));
    }
    RC(E.llist.get(ctx, nlong ? nlong : 1));
    RC(E.lcnt.get(ctx, nlong + 1));
    RC(E.seg_off.get(ctx, nlong + 1));
    long nseg = 0;
    if (nlong) {
      compact_list<<<nb(nrows), 256, 0, ctx->stream>>>(E.lflag.p, E.lpos.p, nrows, E.llist.p);
      long_seg_count<<<nb(nlong), 256, 0, ctx->stream>>>(E.llist.p, E.row_ptr.p, nlong, E.lcnt.p);
      PLANC_LAUNCH_CHECK(ctx);
      RC(E.scan(E.lcnt.p, nlong, E.seg_off.p, &nseg));
      PLANC_HIP_CHECK(ctx, up(E.seg_off.p + nlong, &nseg, sizeof(long)));
    }
    RC(E.seg_row.get(ctx, nseg));
    RC(E.seg_b.get(ctx, nseg));
    RC(E.seg_part.get(ctx, 2 * nseg));
    if (nlong)
      long_seg_fill<<<nb(nlong), 256, 0, ctx->stream>>>(E.llist.p, E.row_ptr.p, nlong, E.seg_off.p, E.seg_row.p,
                                                        E.seg_b.p);
    PLANC_LAUNCH_CHECK(ctx);
    // Node row ranges (rows of node t are contiguous: t-major flag scan).
    std::vector<int> rnode(nrows);
    if (nrows)
      PLANC_HIP_CHECK(ctx, hipMemcpyAsync(rnode.data(), E.row_node.p, sizeof(int) * nrows,
                                          hipMemcpyDeviceToHost, ctx->stream));
    PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<long> roff(nn + 1, 0);
    for (long r = 0; r < nrows; ++r) roff[rnode[r] + 1]++;
    for (int t = 0; t < nn; ++t) roff[t + 1] += roff[t];
    std::vector<Chunk> chr, chc, chi;
    std::vector<int> cbr, cbc, cbi;
    make_chunks(roff, 4096, chr, cbr);
    make_chunks(noff, 4096, chc, cbc);
    std::vector<long> ioff(nn + 1);
    for (int t = 0; t <= nn; ++t) ioff[t] = (long)t * m;
    make_chunks(ioff, 16384, chi, cbi);
    // init chunks address rows 0..m-1 of each node: rebase to local rows.
    for (auto &c : chi) {
      long base = (long)c.node * m;
      c.b -= base;
      c.e -= base;
    }
    RC(E.ch_rows.get(ctx, chr.size()));
    RC(E.ch_cols.get(ctx, chc.size()));
    RC(E.ch_init.get(ctx, chi.size()));
    RC(E.cb_rows.get(ctx, nn + 1));
    RC(E.cb_cols.get(ctx, nn + 1));
    RC(E.cb_init.get(ctx, nn + 1));
    if (!chr.empty()) PLANC_HIP_CHECK(ctx, up(E.ch_rows.p, chr.data(), sizeof(Chunk) * chr.size()));
    if (!chc.empty()) PLANC_HIP_CHECK(ctx, up(E.ch_cols.p, chc.data(), sizeof(Chunk) * chc.size()));
    PLANC_HIP_CHECK(ctx, up(E.ch_init.p, chi.data(), sizeof(Chunk) * chi.size()));
    PLANC_HIP_CHECK(ctx, up(E.cb_rows.p, cbr.data(), sizeof(int) * (nn + 1)));
    PLANC_HIP_CHECK(ctx, up(E.cb_cols.p, cbc.data(), sizeof(int) * (nn + 1)));
    PLANC_HIP_CHECK(ctx, up(E.cb_init.p, cbi.data(), sizeof(int) * (nn + 1)));
    size_t np = std::max(std::max(chr.size(), chc.size()), chi.size());
    RC(E.part.get(ctx, 3 * np + 3));
    RC(E.W.get(ctx, 2 * (nrows ? nrows : 1)));
    RC(E.H.get(ctx, 2 * nlev));
    RC(E.Gw.get(ctx, 4 * nn));
    RC(E.Gh.get(ctx, 4 * nn));
    RC(E.sig.get(ctx, nn));
    RC(E.q.get(ctx, nlev));
    RC(E.z.get(ctx, nrows ? nrows : 1));
    RC(E.mask.get(ctx, nlev));

    // Lanes per CSR row of the row kernels (W update and the power
    // iteration's row pass), timed once per batch (a choice from the mean
    // row length was measured slower on the frontier's small batches).
    int Gb = -1;
    auto lanes = [&]() {
      if (Gb < 0) Gb = E.tune_group(nrows, LONGLEN);
      return Gb;
    };
    // Sigma of this level's nodes (children created by the previous level):
    // 100-step power iteration of Node::compute_sigma (hiernmf/node.hpp:66-69).
    std::vector<double> sig(nn, 0.0);
    if (sigma_on) {
      pow_init<<<nb(nlev), 256, 0, ctx->stream>>>(nlev, E.node_of_col.p, E.node_col0.p, E.seeds.p, E.q.p);
      sumsq_chunks<<<(unsigned)chc.size(), 256, 0, ctx->stream>>>(E.ch_cols.p, E.q.p, E.part.p);
      norm_finalize<<<nb(nn), 256, 0, ctx->stream>>>(E.part.p, E.cb_cols.p, nn, E.sig.p);
      const int Gp = lanes();
      for (int it = 0; it < iters; ++it) {
        E.launch_pow_short(Gp, nrows, LONGLEN);
        if (nlong) {
          pow_rows_seg<<<(unsigned)nseg, 256, 0, ctx->stream>>>(E.seg_row.p, E.seg_b.p, E.llist.p, E.row_ptr.p,
                                                                E.csr_col.p, E.csr_val.p, E.q.p, E.seg_part.p);
          pow_rows_long_fin<<<nb(nlong, 64), 64, 0, ctx->stream>>>(nlong, E.llist.p, E.seg_off.p, E.seg_part.p,
                                                                   E.row_node.p, E.sig.p, E.z.p);
        }
        pow_cols<<<nb(nlev), 256, 0, ctx->stream>>>(nlev, E.col_ptr.p, E.csc_row.p, E.csc_val.p, E.z.p, E.q.p);
        sumsq_chunks<<<(unsigned)chc.size(), 256, 0, ctx->stream>>>(E.ch_cols.p, E.q.p, E.part.p);
        norm_finalize<<<nb(nn), 256, 0, ctx->stream>>>(E.part.p, E.cb_cols.p, nn, E.sig.p);
      }
      PLANC_LAUNCH_CHECK(ctx);
      if (iters > 0)
        PLANC_HIP_CHECK(ctx, hipMemcpyAsync(sig.data(), E.sig.p, sizeof(double) * nn,
                                            hipMemcpyDeviceToHost, ctx->stream));
      PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
      for (int t = 0; t < nn; ++t) nodes[level_nodes[t]].sigma = iters > 0 ? sig[t] : 0.0;
    }
    // Record the batch's nodes (columns once per node).
    for (int t = 0; t < nn; ++t) {
      HNode &nd = nodes[level_nodes[t]];
      if (nd.out_slot >= 0) continue;
      if (cols_used + (long)nd.cols.size() > cols_cap)
        return set_error(ctx, PLANC_EINVAL, "planc_hiernmf_run: column output capacity (n * depth) exceeded");
      nd.out_slot = cols_used;
      std::memcpy(h_cols + cols_used, nd.cols.data(), sizeof(long) * nd.cols.size());
      cols_used += (long)nd.cols.size();
    }
    if (!split_on) return PLANC_OK;  // sigma only (leaves of this pass)

    // Rank-2 NMF of every node of the level: the iteration loop is captured
    // into one hipGraph and launched once.
    init_factors<<<nb(nrows + nlev), 256, 0, ctx->stream>>>(nrows, E.row_node.p, E.row_orig.p, nlev,
                                                            E.node_of_col.p, E.node_col0.p, E.seeds.p,
                                                            m, E.W.p, E.H.p);
    init_gram_chunks<<<(unsigned)chi.size(), 256, 0, ctx->stream>>>(E.ch_init.p, E.seeds.p, m, E.part.p);
    gram2_finalize<<<nb(nn), 256, 0, ctx->stream>>>(E.part.p, E.cb_init.p, nn, E.Gw.p);
    PLANC_LAUNCH_CHECK(ctx);
    const int Gw = lanes();
    if (iters > 0) {
      hipGraphExec_t exec = nullptr;
      RC(capture_graph(ctx, [&]() -> int {
        for (int it = 0; it < iters; ++it) {
          if (it > 0) {
            if (!chr.empty())
              gram2_chunks<<<(unsigned)chr.size(), 256, 0, ctx->stream>>>(E.ch_rows.p, E.W.p, E.part.p);
            gram2_finalize<<<nb(nn), 256, 0, ctx->stream>>>(E.part.p, E.cb_rows.p, nn, E.Gw.p);
          }
          h_update<<<nb(nlev), 256, 0, ctx->stream>>>(nlev, E.col_ptr.p, E.csc_row.p, E.csc_val.p,
                                                     E.node_of_col.p, E.W.p, E.Gw.p, E.H.p);
          gram2_chunks<<<(unsigned)chc.size(), 256, 0, ctx->stream>>>(E.ch_cols.p, E.H.p, E.part.p);
          gram2_finalize<<<nb(nn), 256, 0, ctx->stream>>>(E.part.p, E.cb_cols.p, nn, E.Gh.p);
          E.launch_w_short(Gw, nrows, E.W.p, LONGLEN);
          if (nlong) {
            w_update_seg<<<(unsigned)nseg, 256, 0, ctx->stream>>>(E.seg_row.p, E.seg_b.p, E.llist.p, E.row_ptr.p,
                                                                  E.csr_col.p, E.csr_val.p, E.H.p, E.seg_part.p);
            w_update_long_fin<<<nb(nlong, 64), 64, 0, ctx->stream>>>(nlong, E.llist.p, E.seg_off.p, E.seg_part.p,
                                                                     E.row_node.p, E.Gh.p, E.W.p);
          }
        }
        PLANC_LAUNCH_CHECK(ctx);
        return PLANC_OK;
      }, &exec));
      hipError_t ce = hipGraphLaunch(exec, ctx->stream);
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipGraphExecDestroy(exec);
      if (ce != hipSuccess) return set_error(ctx, PLANC_EHIP, std::string("hiernmf launch: ") + hipGetErrorString(ce));
    }
    // Split rule H(:,0) >= H(:,1), larger side to the left child
    // (hiernmf/node.hpp:147-185).
    split_mask<<<nb(nlev), 256, 0, ctx->stream>>>(nlev, E.H.p, E.mask.p);
    std::vector<unsigned char> mk(nlev);
    PLANC_HIP_CHECK(ctx, hipMemcpyAsync(mk.data(), E.mask.p, nlev, hipMemcpyDeviceToHost, ctx->stream));
    std::vector<double> Hh;
    if (h_H) {
      Hh.resize(2 * nlev);
      PLANC_HIP_CHECK(ctx, hipMemcpyAsync(Hh.data(), E.H.p, sizeof(double) * 2 * nlev, hipMemcpyDeviceToHost, ctx->stream));
    }
    std::vector<double> Wc;
    std::vector<int> rorig;
    if (h_W) {
      Wc.resize(2 * nrows);
      rorig.resize(nrows);
      if (nrows) {
        PLANC_HIP_CHECK(ctx, hipMemcpyAsync(Wc.data(), E.W.p, sizeof(double) * 2 * nrows, hipMemcpyDeviceToHost, ctx->stream));
        PLANC_HIP_CHECK(ctx, hipMemcpyAsync(rorig.data(), E.row_orig.p, sizeof(int) * nrows, hipMemcpyDeviceToHost, ctx->stream));
      }
    }
    PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    for (int t = 0; t < nn; ++t) {
      long id = level_nodes[t];
      long slot = nodes[id].out_slot;  // H / W rows of this node
      if (h_H)
        for (long lc = noff[t]; lc < noff[t + 1]; ++lc) {
          h_H[2 * (slot + (lc - noff[t]))] = Hh[2 * lc];
          h_H[2 * (slot + (lc - noff[t])) + 1] = Hh[2 * lc + 1];
        }
      if (h_W) {
        std::vector<double> &Wf = wstore[id];
        Wf.assign(2 * (size_t)m, 0.0);
        for (long r = roff[t]; r < roff[t + 1]; ++r) {
          Wf[rorig[r]] = Wc[2 * r];
          Wf[rorig[r] + m] = Wc[2 * r + 1];
        }
      }
      std::vector<long> L, R;
      for (long lc = noff[t]; lc < noff[t + 1]; ++lc) (mk[lc] ? L : R).push_back(perm[lc]);
      const bool swapped = R.size() > L.size();
      if (swapped) std::swap(L, R);
      HNode &nd = nodes[id];
      nd.split = true;
      nd.lidx = swapped ? 1 : 0;
      nd.lvalid = !L.empty();
      nd.rvalid = !R.empty();
      long idx = nd.index;
      int lv = nd.level;
      if (nd.lvalid) {
        nodes.push_back(HNode{2 * idx + 1, lv + 1, L});
        nodes.back().parent = id;
        nodes[id].lchild = (long)nodes.size() - 1;
        if (children) children->push_back((long)nodes.size() - 1);
      }
      if (nodes[id].rvalid) {
        nodes.push_back(HNode{2 * idx + 2, lv + 1, R});
        nodes.back().parent = id;
        nodes[id].rchild = (long)nodes.size() - 1;
        if (children) children->push_back((long)nodes.size() - 1);
      }
    }
    return PLANC_OK;
  };

  // Node::compute_score (hiernmf/node.hpp:71-81) from the children's sigmas.
  auto score_of = [&](const HNode &nd) {
    const double ls = nd.lchild >= 0 ? nodes[nd.lchild].sigma : 0.0;
    const double rs = nd.rchild >= 0 ? nodes[nd.rchild].sigma : 0.0;
    if (nd.lvalid && nd.rvalid) return (ls + rs) - nd.sigma;
    if (nd.lvalid) return ls - nd.sigma;
    if (nd.rvalid) return rs - nd.sigma;
    return 0.0;
  };
  std::vector<long> out_order;  // the reported nodes, in the reference's queue order
  if (desc->nodes <= 0) {
    // Complete tree: every node of levels 0 .. depth-2 is split, level by
    // level; the last level only gets its sigma.
    std::vector<long> level_nodes = {0};
    for (int lev = 0; lev < depth && !level_nodes.empty(); ++lev) {
      std::vector<long> next;
      RC(run_batch(level_nodes, lev > 0, lev < depth - 1, &next));
      out_order.insert(out_order.end(), level_nodes.begin(), level_nodes.end());
      level_nodes.swap(next);
    }
  } else {
    // Best-first frontier with a node budget (hiernmf/hiernmf.cpp:138-147):
    // the root is split and accepted (its children split), then nodes - 2
    // times the frontier node of highest score is accepted (its children
    // split).  Accepting needs the children's rank-2 NMFs and the sigmas of
    // their children (for the children's scores), so the device works in
    // rounds: the host replays the reference's queue with what is known,
    // collects every split the replay is missing, and computes them (and
    // the sigmas of the children they create) as one batch each; the replay
    // that misses nothing is the result.  Ties in score go to the smaller
    // node index.  A node at level depth - 1 is not split (treated as having
    // no children).
    std::vector<long> kids;
    RC(run_batch({0}, false, true, &kids));
    std::vector<long> grand;
    RC(run_batch(kids, true, depth > 2, &grand));
    RC(run_batch(grand, true, false, nullptr));
    for (;;) {
      std::vector<long> need;
      auto want_split = [&](long id) {
        HNode &nd = nodes[id];
        if (nd.split || nd.level >= depth - 1) return true;
        if (std::find(need.begin(), need.end(), id) == need.end()) need.push_back(id);
        return false;
      };
      auto cmp = [&](long a, long b) {
        const double sa = nodes[a].score, sb = nodes[b].score;
        if (sa != sb) return sa < sb;
        return nodes[a].index > nodes[b].index;
      };
      std::priority_queue<long, std::vector<long>, decltype(cmp)> frontier(cmp);
      std::vector<long> order = {0};
      auto accept = [&](long id) {  // enqueue_children after accept(); false: a split is missing
        const HNode &nd = nodes[id];
        bool ok = true;
        for (long c : {nd.lchild, nd.rchild})
          if (c >= 0) ok = want_split(c) && ok;
        if (!ok) return false;
        for (long c : {nd.lchild, nd.rchild})
          if (c >= 0) {
            nodes[c].score = score_of(nodes[c]);
            frontier.push(c);
            order.push_back(c);
          }
        return true;
      };
      accept(0);
      for (long i = 2; i < desc->nodes && !frontier.empty(); ++i) {
        const long x = frontier.top();
        frontier.pop();
        accept(x);
      }
      if (need.empty()) {
        out_order = order;
        break;
      }
      std::vector<long> created;
      RC(run_batch(need, false, true, &created));
      RC(run_batch(created, true, false, nullptr));
    }
  }
#undef RC
  // The node table, in out_order.
  if (h_W)
    for (size_t t = 0; t < out_order.size(); ++t) {
      const auto it = wstore.find(out_order[t]);
      double *dst = h_W + t * 2 * (size_t)m;
      if (it != wstore.end())
        std::memcpy(dst, it->second.data(), sizeof(double) * 2 * (size_t)m);
      else
        std::fill(dst, dst + 2 * m, 0.0);
    }
  long cnt = 0;
  for (long id : out_order) {
    HNode &nd = nodes[id];
    nd.score = score_of(nd);
    planc_hier_node &o = h_nodes[cnt];
    o.index = nd.index;
    o.level = nd.level;
    o.ncols = (long)nd.cols.size();
    o.col_offset = nd.out_slot;
    o.sigma = nd.sigma;
    o.score = nd.score;
    o.lvalid = nd.lvalid;
    o.rvalid = nd.rvalid;
    o.split = nd.split ? 1 : 0;
    o.parent_side = -1;
    if (nd.parent >= 0) {
      const HNode &pa = nodes[nd.parent];
      o.parent_side = nd.index == 2 * pa.index + 1 ? pa.lidx : 1 - pa.lidx;
    }
    ++cnt;
  }
  *h_num_nodes = (int)cnt;
  return PLANC_OK;
}

int planc_r2_update(planc_ctx *ctx, const double *d_G, const float *d_R, long r_js, long r_cs, long n,
                    float *d_X, long x_js, long x_cs) {
  if (!ctx) return PLANC_EINVAL;
  if (n < 0 || !d_G || (n > 0 && (!d_R || !d_X)))
    return set_error(ctx, PLANC_EINVAL, "planc_r2_update: bad arguments");
  if (n == 0) return PLANC_OK;
  PLANC_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  r2_update_kernel<<<nb(n), 256, 0, ctx->stream>>>(d_G, d_R, r_js, r_cs, n, d_X, x_js, x_cs);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int planc_csr_col_subset(planc_ctx *ctx, long m, long n, const long *d_indptr, const int *d_indices,
                         const float *d_data, const long *h_cols, long ncols, long *d_out_indptr,
                         int *d_out_indices, float *d_out_data, long *h_out_nnz) {
  if (!ctx) return PLANC_EINVAL;
  if (m < 0 || n < 0 || ncols < 0 || !d_indptr || !d_out_indptr || !h_out_nnz || (ncols && !h_cols))
    return set_error(ctx, PLANC_EINVAL, "planc_csr_col_subset: bad arguments");
  std::vector<int> colmap(n > 0 ? n : 1, -1);
  for (long j = 0; j < ncols; ++j) {
    if (h_cols[j] < 0 || h_cols[j] >= n || (j > 0 && h_cols[j] <= h_cols[j - 1]))
      return set_error(ctx, PLANC_EINVAL, "planc_csr_col_subset: columns must be sorted, unique, in range");
    colmap[h_cols[j]] = (int)j;
  }
  Engine E;
  E.ctx = ctx;
  int rc;
  if ((rc = E.colmap.get(ctx, colmap.size()))) return rc;
  if ((rc = E.row_len.get(ctx, m + 1))) return rc;
  PLANC_HIP_CHECK(ctx, hipMemcpyAsync(E.colmap.p, colmap.data(), sizeof(int) * colmap.size(),
                                      hipMemcpyHostToDevice, ctx->stream));
  subset_count<<<nb(m), 256, 0, ctx->stream>>>(m, d_indptr, d_indices, E.colmap.p, E.row_len.p);
  PLANC_LAUNCH_CHECK(ctx);
  long total = 0;
  if ((rc = E.scan(E.row_len.p, m, d_out_indptr, &total))) return rc;
  PLANC_HIP_CHECK(ctx, hipMemcpyAsync(d_out_indptr + m, &total, sizeof(long), hipMemcpyHostToDevice, ctx->stream));
  if (total > 0 && d_out_indices && d_out_data) {
    subset_fill<<<nb(m), 256, 0, ctx->stream>>>(m, d_indptr, d_indices, d_data, E.colmap.p, d_out_indptr,
                                               d_out_indices, d_out_data);
    PLANC_LAUNCH_CHECK(ctx);
  }
  PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *h_out_nnz = total;
  return PLANC_OK;
}

}  // extern "C"
