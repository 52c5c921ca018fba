// NTF kernels besides the bulk contractions: projection contraction
// (multi-TTV of the dimension tree), Hadamard of Grams, and the FP64 terms
// of the CP relative error.
#include "common.h"
#include "rowupd.h"

namespace planc {

namespace {

constexpr int NTHR = 256;

__device__ inline double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

struct ContractParams {
  const float *P;  // E x k, column-major (ld ldp); E = prod d[0..nb)
  long ldp;
  int nb;          // number of modes in the block
  int tn;          // position of the target mode inside the block
  int d[8];
  const float *F[8];
  long ldf[8];
  int k;
  float *M;        // k x d[tn], column-major (ld ldo)
  long ldo;
  long inner, In, outer;
};

// Product of F_t(digit_t(idx), c) over block positions [t0, t1) with idx the
// mixed-radix index of those modes (lowest fastest).  idx < 2^32.
__device__ inline float weight(const ContractParams &p, unsigned idx, int t0, int t1, int c) {
  float w = 1.0f;
  for (int t = t0; t < t1; ++t) {
    unsigned nx = idx / (unsigned)p.d[t];
    unsigned dg = idx - nx * (unsigned)p.d[t];
    idx = nx;
    w *= p.F[t][dg + p.ldf[t] * c];
  }
  return w;
}

constexpr int WMAX = 4096;  // weights held in LDS per block

// inner == 1 (target mode first in the block): lanes walk consecutive output
// indices i (coalesced), the 16 wavefronts split the outer index b
// (b = wave, wave + 16, ...), summed in wave order.
__global__ __launch_bounds__(1024) void contract_outer_kernel(ContractParams p) {
  __shared__ float wo[WMAX];
  __shared__ float red[16][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = blockIdx.y;
  const long i = (long)blockIdx.x * 64 + lane;
  const bool cached = p.outer <= WMAX;
  const bool valid = i < p.In;
  const float *src = p.P + p.ldp * c + (valid ? i : 0);
  const long st = p.In * 16;
  // The first group of eight loads is issued before the weights are
  // computed (independent round trips overlap); same sums in the same order.
  float v0[8];
  const bool first = cached && valid && wave + 16 * 7 < p.outer;
  if (first) {
#pragma unroll
    for (int u = 0; u < 8; ++u) v0[u] = src[p.In * wave + st * u];
  }
  if (cached)
    for (long b = threadIdx.x; b < p.outer; b += blockDim.x) wo[b] = weight(p, (unsigned)b, p.tn + 1, p.nb, c);
  __syncthreads();
  float acc = 0.0f;
  if (valid) {
    long b = wave;
    if (cached) {
      if (first) {
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = fmaf(v0[u], wo[b + 16 * u], acc);
        b += 16 * 8;
      }
      for (; b + 16 * 7 < p.outer; b += 16 * 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = src[p.In * b + st * u];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = fmaf(v[u], wo[b + 16 * u], acc);
      }
      for (; b < p.outer; b += 16) acc = fmaf(src[p.In * b], wo[b], acc);
    } else {
      for (; b < p.outer; b += 16) acc = fmaf(src[p.In * b], weight(p, (unsigned)b, p.tn + 1, p.nb, c), acc);
    }
  }
  red[wave][lane] = acc;
  __syncthreads();
  if (wave == 0 && i < p.In) {
    float t = 0.0f;
#pragma unroll
    for (int w = 0; w < 16; ++w) t += red[w][lane];
    p.M[c + p.ldo * i] = t;
  }
}

// inner > 1, inner % 4 == 0, 16-byte aligned columns: a wavefront computes
// CW_OUT outputs (i, c), i = first + 4 r, with their loads issued together;
// lanes read 16 consecutive bytes of the contiguous inner index, weights
// from LDS (computed once per block for its 4 * CW_OUT outputs); per output
// the same fixed-order sums and FP64 wave reduction.
constexpr int CW_OUT = 4;
__global__ __launch_bounds__(256) void contract_wave_kernel(ContractParams p) {
  __shared__ __attribute__((aligned(16))) float wi[WMAX];
  __shared__ float wo[WMAX];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = blockIdx.y;
  const long i0 = (long)blockIdx.x * (4 * CW_OUT) + wave;
  for (long a = threadIdx.x; a < p.inner; a += blockDim.x) wi[a] = weight(p, (unsigned)a, 0, p.tn, c);
  for (long b = threadIdx.x; b < p.outer; b += blockDim.x) wo[b] = weight(p, (unsigned)b, p.tn + 1, p.nb, c);
  __syncthreads();
  if (i0 >= p.In) return;
  float acc[CW_OUT];
#pragma unroll
  for (int r = 0; r < CW_OUT; ++r) acc[r] = 0.0f;
  const float *Pc = p.P + p.ldp * c;
  for (long b = 0; b < p.outer; ++b) {
    float sb[CW_OUT];
#pragma unroll
    for (int r = 0; r < CW_OUT; ++r) sb[r] = 0.0f;
    for (long a = 4 * lane; a < p.inner; a += 256) {
      const floatx4 w = *(const floatx4 *)&wi[a];
      floatx4 v[CW_OUT];
#pragma unroll
      for (int r = 0; r < CW_OUT; ++r) {
        const long i = i0 + 4 * r;
        const long ic = i < p.In ? i : i0;  // rows past In re-read row i0 (not stored)
        v[r] = *(const floatx4 *)(Pc + p.inner * (ic + p.In * b) + a);
      }
#pragma unroll
      for (int r = 0; r < CW_OUT; ++r) {
        sb[r] = fmaf(v[r][0], w[0], sb[r]);
        sb[r] = fmaf(v[r][1], w[1], sb[r]);
        sb[r] = fmaf(v[r][2], w[2], sb[r]);
        sb[r] = fmaf(v[r][3], w[3], sb[r]);
      }
    }
#pragma unroll
    for (int r = 0; r < CW_OUT; ++r) acc[r] = fmaf(sb[r], wo[b], acc[r]);
  }
#pragma unroll
  for (int r = 0; r < CW_OUT; ++r) {
    const long i = i0 + 4 * r;
    const double t = wave_sum_d((double)acc[r]);
    if (lane == 0 && i < p.In) p.M[c + p.ldo * i] = (float)t;
  }
}

// inner > 1: one block per (tile of TI output indices, c); threads walk the
// contiguous inner index a, loop over b; each output is a deterministic
// block reduction.
constexpr int TI = 8;
__global__ __launch_bounds__(256) void contract_inner_kernel(ContractParams p) {
  __shared__ float wi[WMAX];
  __shared__ float wo[WMAX];
  __shared__ double red[TI][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = blockIdx.y;
  const long i0 = (long)blockIdx.x * TI;
  const bool ci = p.inner <= WMAX, co = p.outer <= WMAX;
  if (ci)
    for (long a = threadIdx.x; a < p.inner; a += blockDim.x) wi[a] = weight(p, (unsigned)a, 0, p.tn, c);
  if (co)
    for (long b = threadIdx.x; b < p.outer; b += blockDim.x) wo[b] = weight(p, (unsigned)b, p.tn + 1, p.nb, c);
  __syncthreads();
  float acc[TI];
#pragma unroll
  for (int t = 0; t < TI; ++t) acc[t] = 0.0f;
  const float *Pc = p.P + p.ldp * c;
  for (long b = 0; b < p.outer; ++b) {
    const float wb = co ? wo[b] : weight(p, (unsigned)b, p.tn + 1, p.nb, c);
    for (long a = threadIdx.x; a < p.inner; a += blockDim.x) {
      const float w = (ci ? wi[a] : weight(p, (unsigned)a, 0, p.tn, c)) * wb;
#pragma unroll
      for (int t = 0; t < TI; ++t) {
        const long i = i0 + t;
        if (i < p.In) acc[t] = fmaf(Pc[a + p.inner * (i + p.In * b)], w, acc[t]);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < TI; ++t) {
    double v = wave_sum_d((double)acc[t]);
    if (lane == 0) red[t][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < TI) {
    const long i = i0 + threadIdx.x;
    if (i < p.In) {
      double s = red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] + red[threadIdx.x][3];
      p.M[c + p.ldo * i] = (float)s;
    }
  }
}

struct GramSet {
  const double *g[8];
  int n;
};

__global__ void hadamard_kernel(GramSet gs, int skip, int kk, double *out) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kk) return;
  double v = 1.0;
  for (int i = 0; i < gs.n; ++i)
    if (i != skip) v *= gs.g[i][e];
  out[e] = v;
}

// sum_c lambda_c sum_i F(i, c) * M(c, i), two stages.
__global__ __launch_bounds__(NTHR) void lambda_dot_partial_kernel(const float *F, long I, long ldf,
                                                                  const float *M, long ldm, int k,
                                                                  const double *lambda,
                                                                  double *part) {
  __shared__ double red[NTHR / 64];
  double s = 0.0;
  long tot = I * k;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    long i = e % I;
    int c = (int)(e / I);
    s += lambda[c] * (double)F[i + ldf * c] * (double)M[c + ldm * i];
  }
  s = wave_sum_d(s);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < NTHR / 64; ++w) t += red[w];
    part[blockIdx.x] = t;
  }
}

// rel = sqrt(max(||X||^2 - 2 <X, Xhat> + ||Xhat||^2, 0) / ||X||^2) with
// ||Xhat||^2 = lambda^T (Hadamard of all Grams) lambda.
__global__ void ntf_relerr_kernel(const double *normX2, const double *part, long nparts, GramSet gs,
                                  int k, const double *lambda, double *rel_out) {
  if (threadIdx.x >= 64) return;
  double t = 0.0;
  for (long b = threadIdx.x; b < nparts; b += 64) t += part[b];
  t = wave_sum_d(t);
  double xh = 0.0;
  for (int e = threadIdx.x; e < k * k; e += 64) {
    int a = e % k, b = e / k;
    double v = lambda[a] * lambda[b];
    for (int i = 0; i < gs.n; ++i) v *= gs.g[i][e];
    xh += v;
  }
  xh = wave_sum_d(xh);
  if (threadIdx.x == 0) {
    double e2 = *normX2 - 2.0 * t + xh;
    if (e2 < 0) e2 = 0;
    *rel_out = sqrt(e2 / *normX2);
  }
}

// lambda <- lambda .* norms (NCPFactors::normalize(), common/ncpfactors.hpp:329-339).
__global__ void lambda_mul_kernel(double *lambda, const double *norms, int k, int init) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k) return;
  lambda[c] = (init ? 1.0 : lambda[c]) * norms[c];
}

// out (k x I, ld ldo) = in^T for in I x k (ld ldi).
__global__ void transpose_kernel(const float *in, long I, long ldi, int k, float *out, long ldo) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= I * k) return;
  int c = (int)(e % k);
  long i = e / k;
  out[c + ldo * i] = in[i + ldi * c];
}

// ---------------------------------------------------------------------------
// One mode update of AUNTF: S = Hadamard of the other Grams, the HALS
// (ntf/ntfhals.hpp:13-29) or MU row update from this mode's MTTKRP,
// normalize(mode) (ntf/auntf.hpp:66-79: lambda = column 2-norms, columns
// divided by them), the FP64 Gram of the new factor and optionally the CP
// relative error of ntf/auntf.hpp:224-245 from this mode's MTTKRP.
//   mode_rows_kernel: a row group of MRT = 64 rows is handled by 4 waves, a
//     quad of lanes per row (each lane holds k/4 of the row, the column dot
//     products are quad sums); the group writes the unnormalised rows u and
//     accumulates the FP64 Gram of u and, per column, sum_i u(i, c) M(c, i)
//     (16-row groups over more CUs measured slower at 256 x 32).
//   NVB == 0 (any I): one group per block, the blocks stride over the rows
//     and store their partials; mode_finalize_kernel (second launch) sums
//     them in block order, derives the norms, divides F, writes lambda, the
//     normalised Gram and the error.
//   NVB >= 1 (I <= 64 NVB): one launch of a single block of NVB groups
//     (256 NVB threads) holding every row once; the groups leave their
//     partials in LDS, every thread sums them in group order (the same
//     order and grouping as the two-launch form, so the results are
//     identical), and the rows, still in registers, are divided by the
//     norms.  No cross-block traffic: at I <= 256 the update is a chain of
//     dependent round trips, and the second launch or a cross-block
//     exchange is a large part of it.
// kind: 0 MU, 1 HALS, 2 no row update (the factor was solved already, BPP).
struct ModeParams {
  float *F;
  long I, ldf;
  int k;
  const float *M;  // MTTKRP: element (x, i) at M[i * mrs + x * mcs]
  long mrs, mcs;
  const double *g[8];
  int ng, skip;
  double *gout, *lambda;
  int kind;
  float eps;
  const double *normX2;
  double *rel_out;
  double *part;  // NVB == 0: per block, the k*k Gram of u, then k column dots
  int *rel_counter;  // non-null: the error goes to rel_out[*rel_counter], which is incremented
};

template <int KM, int KIND, int MRT, int NVB>
__global__ __launch_bounds__(4 * MRT *(NVB > 0 ? NVB : 1)) void mode_rows_kernel(ModeParams p) {
  constexpr int MR = MRT;                      // rows per group
  constexpr int NTH = 4 * MR;                  // threads per group: a quad per row
  constexpr int NG1 = NVB > 0 ? NVB : 1;       // groups per block
  constexpr int NTT = NTH * NG1;               // threads per block
  constexpr int NW = NTH / 64;                 // waves per group
  constexpr int Q = KM / 4;                    // row values per lane
  constexpr int NT = KM / 16;                  // 16 x 16 tiles per Gram side
  constexpr int NUP = NT * (NT + 1) / 2;       // upper-triangle tiles
  constexpr int TPW = (NUP + NW - 1) / NW;     // tiles per wave
  constexpr int NDP = (4 * KM + NTH - 1) / NTH;  // (column, quarter) dot pairs per thread
  __shared__ __attribute__((aligned(16))) float St[KM][KM];  // St[x][t] = S(t, x)
  // Per group: u(row, c) at Us[c][row], M(c, row) at Ms[c][row]; the
  // one-block form then reuses a group's space for its partials (2 KM (MR +
  // 1) floats >= KM^2 + KM doubles).
  __shared__ __attribute__((aligned(16))) float UMs[NG1][2][KM][MR + 1];
  static_assert(MR + 1 >= KM + 1, "a group's partials must fit its row tiles");
  const int tid = threadIdx.x, vb = tid / NTH, gt = tid - vb * NTH;
  const int q = gt & 3, rr = gt >> 2;
  const int lane = tid & 63, wave = gt >> 6;
  float(&Us)[KM][MR + 1] = UMs[vb][0];
  float(&Ms)[KM][MR + 1] = UMs[vb][1];
  const int k = p.k;
  const long stride = (long)gridDim.x * MR;
  // Row values and MTTKRP values of rows r0 + rr (clamped to a valid row).
  auto load_rows = [&](long r0, float(&h)[Q], float(&mv)[Q]) {
    const long i = r0 + rr;
    const long ic = i < p.I ? i : p.I - 1;
#pragma unroll
    for (int u = 0; u < Q; ++u) {
      const int t = q * Q + u;
      const int tc = t < k ? t : k - 1;
      h[u] = t < k ? p.F[ic + p.ldf * tc] : 0.0f;
      mv[u] = t < k ? p.M[ic * p.mrs + tc * p.mcs] : 0.0f;
    }
  };
  // Upper-triangle tile index -> (ta, tb), ta <= tb.
  auto tile_of = [&](int t, int &ta, int &tb) {
    ta = 0;
    int rem = t;
    while (rem >= NT - ta) {
      rem -= NT - ta;
      ++ta;
    }
    tb = ta + rem;
  };
  float h[Q], mv[Q];
  long r0 = NVB > 0 ? (long)vb * MR : (long)blockIdx.x * MR;
  // S = Hadamard of the other Grams: every load of the block's entries is
  // issued before the first product (one memory round trip, not one per
  // entry), then the rows' loads, which stay in flight while S is built.
  constexpr int NE = (KM * KM + NTT - 1) / NTT;
  double dS[NE];  // FP64 Hadamard entries of this thread (the one-block error)
  {
    double gv[NE][4];
#pragma unroll
    for (int it = 0; it < NE; ++it) {
      const int e = tid + it * NTT;
      const int a = e % KM, b = e / KM;
      const bool in = e < KM * KM && a < k && b < k;
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int gg = g < p.skip ? g : g + 1;  // the g-th Gram other than skip
        gv[it][g] = (in && gg < p.ng) ? p.g[gg][a + (long)k * b] : 1.0;
      }
    }
    load_rows(r0, h, mv);
#pragma unroll
    for (int it = 0; it < NE; ++it) {
      const int e = tid + it * NTT;
      const int a = e % KM, b = e / KM;
      double d = gv[it][0];  // in Gram order, as the sequential product
      d *= gv[it][1];
      d *= gv[it][2];
      d *= gv[it][3];
      for (int gg = p.skip <= 4 ? 5 : 4; gg < p.ng; ++gg)  // more than five modes: the rest
        if (gg != p.skip) d *= (a < k && b < k) ? p.g[gg][a + (long)k * b] : 1.0;
      if (e < KM * KM) St[b][a] = (a < k && b < k) ? (float)d : 0.0f;
      dS[it] = d;
    }
  }
  doublex4 gacc[TPW];
#pragma unroll
  for (int j = 0; j < TPW; ++j) gacc[j] = doublex4{0.0, 0.0, 0.0, 0.0};
  double dacc[NDP];
#pragma unroll
  for (int j = 0; j < NDP; ++j) dacc[j] = 0.0;
  __syncthreads();
  // One pass per group in the one-block form (uniform barriers: every group
  // runs it); the blocks of the two-launch form stride over the rows.
  for (;;) {
    const long i = r0 + rr;
    const bool valid = i < p.I;
    quad_row_update<KM, false>(h, mv, St, q, k, KIND, p.eps);
#pragma unroll
    for (int u = 0; u < Q; ++u) {
      const int t = q * Q + u;
      Us[t][rr] = valid ? h[u] : 0.0f;
      Ms[t][rr] = valid ? mv[u] : 0.0f;
      if (NVB == 0 && valid && t < k) p.F[i + p.ldf * t] = h[u];
    }
    const bool more = NVB == 0 && r0 + stride < p.I;
    if (more) load_rows(r0 + stride, h, mv);  // next rows in flight
    __syncthreads();
    // Gram of this chunk's rows on FP64 MFMA: D(a, b) += sum_r U(r, a) U(r, b)
    // per 16 x 16 tile; lane (lane%16, lane/16) feeds A = U(r0 + lane/16, a0 +
    // lane%16) and B = U(r0 + lane/16, b0 + lane%16).  The row steps go
    // through two accumulators per tile (even and odd steps, summed at the
    // end) and the wave's tiles are interleaved, so consecutive MFMAs do not
    // wait on each other.
    {
      const float *ua[TPW], *ub[TPW];
      bool on[TPW];
#pragma unroll
      for (int j = 0; j < TPW; ++j) {
        const int t = wave + NW * j;
        on[j] = t < NUP;
        int ta = 0, tb = 0;
        if (on[j]) tile_of(t, ta, tb);
        ua[j] = &Us[16 * ta + (lane & 15)][0];
        ub[j] = &Us[16 * tb + (lane & 15)][0];
      }
      doublex4 g2[TPW];
#pragma unroll
      for (int j = 0; j < TPW; ++j) g2[j] = doublex4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int r = 0; r < MR / 4; r += 2) {
        const int ra = (lane >> 4) + 4 * r, rb = ra + 4;
#pragma unroll
        for (int j = 0; j < TPW; ++j)
          if (on[j]) {
            gacc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)ua[j][ra], (double)ub[j][ra], gacc[j], 0, 0, 0);
            g2[j] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)ua[j][rb], (double)ub[j][rb], g2[j], 0, 0, 0);
          }
      }
#pragma unroll
      for (int j = 0; j < TPW; ++j) gacc[j] += g2[j];
    }
    // Column dots sum_r U(r, c) M(c, r): pair (c, quarter) = gt + NTH j,
    // the four quarters of a column in consecutive lanes (quad sum below).
#pragma unroll
    for (int j = 0; j < NDP; ++j) {
      const int pi = gt + NTH * j;
      if (pi < 4 * KM) {
        const int c = pi >> 2, part = pi & 3;
        double s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int r = part; r < MR; r += 8) {
          s0 += (double)Us[c][r] * (double)Ms[c][r];
          s1 += (double)Us[c][r + 4] * (double)Ms[c][r + 4];
        }
        dacc[j] += s0 + s1;
      }
    }
    __syncthreads();
    if (!more) break;
    r0 += stride;
  }
  const int kk = k * k;
  // Partial (a, b) of this group: into global memory (two-launch form) or
  // into the group's LDS space (one-block form, free after the last barrier).
  double *dst = NVB > 0 ? reinterpret_cast<double *>(&UMs[vb][0][0][0])
                        : p.part + (long)blockIdx.x * (kk + k);
  // FP64 D layout: lane holds D(lane / 16 + 4 v, lane % 16), v = 0..3.
#pragma unroll
  for (int j = 0; j < TPW; ++j) {
    const int t = wave + NW * j;
    if (t < NUP) {
      int ta, tb;
      tile_of(t, ta, tb);
      const int b = 16 * tb + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int a = 16 * ta + (lane >> 4) + 4 * v;
        if (a < k && b < k && (ta != tb || a <= b)) {
          dst[a + k * b] = gacc[j][v];
          dst[b + k * a] = gacc[j][v];
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < NDP; ++j) {
    // fixed-order quad combine: ((d0 + d1) + (d2 + d3)); every lane of a
    // quad takes part (the shuffles need all of them).
    const int pi = gt + NTH * j;
    double d1 = __shfl_xor(dacc[j], 1, 64);
    double s01 = (tid & 1) ? d1 + dacc[j] : dacc[j] + d1;
    double s23 = __shfl_xor(s01, 2, 64);
    double tot = (tid & 2) ? s23 + s01 : s01 + s23;
    const int c = pi >> 2;
    if (pi < 4 * KM && (pi & 3) == 0 && c < k) dst[kk + c] = tot;
  }
  if constexpr (NVB > 0) {
    __syncthreads();
    // Group-order sums of the partials (0 + P0 + P1 + ..., the order of
    // mode_finalize_kernel), kept in registers until every group's space
    // has been read, then stored over group 0's.
    constexpr int NGS = (KM * KM + KM + NTT - 1) / NTT;
    double gs[NGS];
#pragma unroll
    for (int j = 0; j < NGS; ++j) {
      const int e = tid + NTT * j;
      double sj = 0.0;
      if (e < kk + k) {
#pragma unroll
        for (int g = 0; g < NVB; ++g) sj += reinterpret_cast<const double *>(&UMs[g][0][0][0])[e];
      }
      gs[j] = sj;
    }
    __syncthreads();
    __shared__ double nrm[KM];
    __shared__ double red[2][NTT / 64];
    double *Gu = reinterpret_cast<double *>(&UMs[0][0][0][0]);
#pragma unroll
    for (int j = 0; j < NGS; ++j) {
      const int e = tid + NTT * j;
      if (e < kk + k) Gu[e] = gs[j];
    }
    __syncthreads();
    if (tid < k) {
      nrm[tid] = sqrt(Gu[tid + k * tid]);
      p.lambda[tid] = nrm[tid];
    }
    __syncthreads();
    // This group's rows (still in registers) divided by the norms.
    {
      const long i = r0 + rr;
      if (i < p.I) {
#pragma unroll
        for (int u = 0; u < Q; ++u) {
          const int t = q * Q + u;
          if (t < k) p.F[i + p.ldf * t] = nrm[t] > 0 ? (float)((double)h[u] / nrm[t]) : h[u];
        }
      }
    }
    for (int e = tid; e < kk; e += NTT) {
      const int a = e % k, b = e / k;
      const double sa = nrm[a] > 0 ? nrm[a] : 1.0, sb = nrm[b] > 0 ? nrm[b] : 1.0;
      p.gout[e] = Gu[e] / (sa * sb);
    }
    if (p.rel_out) {
      // <X, model> = sum_c (u^T M)(c); ||model||^2 = sum (Gram u) .* S with
      // S the FP64 Hadamard of the other Grams built above.
      double t1 = 0.0, xh = 0.0;
      for (int c = tid; c < k; c += NTT) t1 += Gu[kk + c];
#pragma unroll
      for (int it = 0; it < NE; ++it) {
        const int e = tid + it * NTT;
        const int a = e % KM, b = e / KM;
        if (e < KM * KM && a < k && b < k) xh += Gu[a + k * b] * dS[it];
      }
      t1 = wave_sum_d(t1);
      xh = wave_sum_d(xh);
      if (lane == 0) {
        red[0][tid >> 6] = t1;
        red[1][tid >> 6] = xh;
      }
      __syncthreads();
      if (tid == 0) {
        t1 = 0.0;
        xh = 0.0;
        for (int w = 0; w < NTT / 64; ++w) {
          t1 += red[0][w];
          xh += red[1][w];
        }
        double e2 = *p.normX2 - 2.0 * t1 + xh;
        if (e2 < 0) e2 = 0;
        const double rel = sqrt(e2 / *p.normX2);
        if (p.rel_counter) {
          const int idx = *p.rel_counter;
          p.rel_out[idx] = rel;
          *p.rel_counter = idx + 1;
        } else {
          *p.rel_out = rel;
        }
      }
    }
  }
}

// Every block sums the block partials in block order (identical results in
// every block), derives the norms and divides its slice of F by them; block
// 0 also writes lambda, the Gram of the normalised factor and the error.
// Rolled loops keep this run-once kernel's code small (instruction fetch,
// not arithmetic, bounds a kernel this short).
__global__ __launch_bounds__(256) void mode_finalize_kernel(ModeParams p, int nblocks) {
  __shared__ double Gu[64 * 64 + 64];
  __shared__ double nrm[64];
  __shared__ double red[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int k = p.k, kk = k * k;
  const long tot = p.I * k;
  // This block's factor values, loaded before the partial sums are waited on.
  const long e0 = (long)blockIdx.x * 256 + tid;
  const long estride = (long)gridDim.x * 256;
  float f0 = 0.0f;
  if (e0 < tot) {
    const long c = e0 / p.I, i = e0 - c * p.I;
    f0 = p.F[i + p.ldf * c];
  }
  for (int e = tid; e < kk + k; e += 256) {
    double s = 0.0;
#pragma unroll 4
    for (int b = 0; b < nblocks; ++b) s += p.part[(long)b * (kk + k) + e];
    Gu[e] = s;
  }
  __syncthreads();
  if (tid < k) nrm[tid] = sqrt(Gu[tid + k * tid]);
  __syncthreads();
  for (long e = e0; e < tot; e += estride) {
    const long c = e / p.I, i = e - c * p.I;
    const float v = e == e0 ? f0 : p.F[i + p.ldf * c];
    if (nrm[c] > 0) p.F[i + p.ldf * c] = (float)((double)v / nrm[c]);
  }
  if (blockIdx.x != 0) return;
  if (tid < k) p.lambda[tid] = nrm[tid];
  for (int e = tid; e < kk; e += 256) {
    const int a = e % k, b = e / k;
    const double sa = nrm[a] > 0 ? nrm[a] : 1.0, sb = nrm[b] > 0 ? nrm[b] : 1.0;
    p.gout[e] = Gu[e] / (sa * sb);
  }
  if (!p.rel_out) return;
  double t1 = 0.0, xh = 0.0;
  for (int c = tid; c < k; c += 256) t1 += Gu[kk + c];
  for (int e = tid; e < kk; e += 256) {
    const int a = e % k, b = e / k;
    const double sa = nrm[a] > 0 ? nrm[a] : 1.0, sb = nrm[b] > 0 ? nrm[b] : 1.0;
    double v = nrm[a] * nrm[b] * (Gu[e] / (sa * sb));
#pragma unroll
    for (int g = 0; g < 8; ++g)
      if (g < p.ng && g != p.skip) v *= p.g[g][e];
    xh += v;
  }
  t1 = wave_sum_d(t1);
  xh = wave_sum_d(xh);
  __syncthreads();
  if (lane == 0) red[wave] = t1;
  __syncthreads();
  if (tid == 0) t1 = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  if (lane == 0) red[wave] = xh;
  __syncthreads();
  if (tid == 0) {
    xh = red[0] + red[1] + red[2] + red[3];
    double e2 = *p.normX2 - 2.0 * t1 + xh;
    if (e2 < 0) e2 = 0;
    const double rel = sqrt(e2 / *p.normX2);
    if (p.rel_counter) {
      const int idx = *p.rel_counter;
      p.rel_out[idx] = rel;
      *p.rel_counter = idx + 1;
    } else {
      *p.rel_out = rel;
    }
  }
}


}  // namespace

// M (k x I_n) from a projection P (E x k) over a block of modes with dims d,
// factors F, target position tn inside the block.
int contract_projection(planc_ctx *ctx, const float *P, long ldp, int nb, const long *d,
                        const float *const *F, const long *ldf, int tn, int k, float *M, long ldo) {
  if (nb < 1 || nb > 8 || tn < 0 || tn >= nb) return set_error(ctx, PLANC_EINVAL, "contract: bad block");
  ContractParams p{};
  p.P = P;
  p.ldp = ldp;
  p.nb = nb;
  p.tn = tn;
  p.k = k;
  p.M = M;
  p.ldo = ldo;
  p.inner = 1;
  p.outer = 1;
  for (int t = 0; t < nb; ++t) {
    p.d[t] = (int)d[t];
    p.F[t] = F[t];
    p.ldf[t] = ldf[t];
    if (t < tn) p.inner *= d[t];
    if (t > tn) p.outer *= d[t];
  }
  p.In = d[tn];
  if (p.In == 0) return PLANC_OK;
  long E = p.inner * p.In * p.outer;
  if (E >= (1L << 32)) return set_error(ctx, PLANC_EINVAL, "contract: block of 2^32 rows or more");
  if (p.inner == 1) {
    dim3 grid((unsigned)((p.In + 63) / 64), (unsigned)k);
    contract_outer_kernel<<<grid, 1024, 0, ctx->stream>>>(p);
  } else if (p.inner % 4 == 0 && p.inner <= WMAX && p.outer <= WMAX && ldp % 4 == 0 &&
             ((uintptr_t)P % 16) == 0) {
    dim3 grid((unsigned)((p.In + 4 * CW_OUT - 1) / (4 * CW_OUT)), (unsigned)k);
    contract_wave_kernel<<<grid, 256, 0, ctx->stream>>>(p);
  } else {
    dim3 grid((unsigned)((p.In + TI - 1) / TI), (unsigned)k);
    contract_inner_kernel<<<grid, 256, 0, ctx->stream>>>(p);
  }
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int hadamard_grams(planc_ctx *ctx, const double *const *grams, int n, int skip, int k, double *out) {
  if (n < 1 || n > 8) return set_error(ctx, PLANC_EINVAL, "hadamard: 1..8 grams");
  GramSet gs{};
  gs.n = n;
  for (int i = 0; i < n; ++i) gs.g[i] = grams[i];
  hadamard_kernel<<<(k * k + 255) / 256, 256, 0, ctx->stream>>>(gs, skip, k * k, out);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int ntf_relerr(planc_ctx *ctx, const double *normX2, const float *F, long I, long ldf, const float *M,
               long ldm, int k, const double *lambda, const double *const *grams, int n,
               double *rel_out) {
  long nb = (I * k + NTHR * 4 - 1) / (NTHR * 4);
  if (nb > 1024) nb = 1024;
  if (nb < 1) nb = 1;
  double *part = (double *)ws_get(ctx, 3, sizeof(double) * (size_t)nb);
  if (!part) return set_error(ctx, PLANC_ENOMEM, "ntf_relerr: workspace");
  lambda_dot_partial_kernel<<<(unsigned)nb, NTHR, 0, ctx->stream>>>(F, I, ldf, M, ldm, k, lambda, part);
  PLANC_LAUNCH_CHECK(ctx);
  GramSet gs{};
  gs.n = n;
  for (int i = 0; i < n; ++i) gs.g[i] = grams[i];
  ntf_relerr_kernel<<<1, 64, 0, ctx->stream>>>(normX2, part, nb, gs, k, lambda, rel_out);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int lambda_mul(planc_ctx *ctx, double *lambda, const double *norms, int k, int init) {
  lambda_mul_kernel<<<(k + 63) / 64, 64, 0, ctx->stream>>>(lambda, norms, k, init);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int transpose_to_k_major(planc_ctx *ctx, const float *in, long I, long ldi, int k, float *out, long ldo) {
  long tot = I * k;
  if (tot == 0) return PLANC_OK;
  transpose_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, ctx->stream>>>(in, I, ldi, k, out, ldo);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

}  // namespace planc

namespace planc {

bool mode_update_fits(long I, int k) { return k >= 1 && k <= 64 && I >= 1 && I < (1L << 31); }

int mode_update(planc_ctx *ctx, int kind, float eps, float *F, long I, long ldf, int k, const float *M,
                long mrs, long mcs, const double *const *grams, int ng, int skip, double *gout,
                double *lambda, const double *normX2, double *rel_out, int *rel_counter) {
  if (!mode_update_fits(I, k) || ng < 1 || ng > 8 || skip < 0 || skip >= ng)
    return set_error(ctx, PLANC_EINVAL, "mode_update: bad arguments");
  ModeParams p{};
  p.F = F;
  p.I = I;
  p.ldf = ldf;
  p.k = k;
  p.M = M;
  p.mrs = mrs;
  p.mcs = mcs;
  for (int q = 0; q < ng; ++q) p.g[q] = grams[q];
  p.ng = ng;
  p.skip = skip;
  p.gout = gout;
  p.lambda = lambda;
  p.kind = kind;
  p.eps = eps;
  p.normX2 = normX2;
  p.rel_out = rel_out;
  p.rel_counter = rel_counter;
  const int mr = 64;
  const long groups = (I + mr - 1) / mr;
  // One block of `groups` row groups: up to 4 (1024 threads; LDS 2 k x 65
  // floats per group) for k <= 48, up to 2 for k <= 64.
  const char *oenv = getenv("PLANC_MODE_ONE_BLOCK");  // 0: always the two-launch form
  const bool one = !(oenv && oenv[0] == '0') && groups <= (k <= 48 ? 4 : 2);
  if (one) {
#define ONE(KMV, G)                                                                     \
  do {                                                                                  \
    if (kind == 0) mode_rows_kernel<KMV, 0, 64, G><<<1, 256 * G, 0, ctx->stream>>>(p);      \
    else if (kind == 1) mode_rows_kernel<KMV, 1, 64, G><<<1, 256 * G, 0, ctx->stream>>>(p); \
    else mode_rows_kernel<KMV, 2, 64, G><<<1, 256 * G, 0, ctx->stream>>>(p);                \
  } while (0)
#define ONE_K(KMV)                    \
  do {                                \
    if (groups == 1) ONE(KMV, 1);     \
    else if (groups == 2) ONE(KMV, 2); \
    else if (groups == 3) ONE(KMV, 3); \
    else ONE(KMV, 4);                 \
  } while (0)
    if (k <= 16) ONE_K(16);
    else if (k <= 32) ONE_K(32);
    else if (k <= 48) ONE_K(48);
    else if (groups == 1) ONE(64, 1);
    else ONE(64, 2);
#undef ONE_K
#undef ONE
    PLANC_LAUNCH_CHECK(ctx);
    return PLANC_OK;
  }
  long nb = groups;
  if (nb > ctx->num_cu) nb = ctx->num_cu;
  p.part = (double *)ws_get(ctx, 18, sizeof(double) * (size_t)nb * (k * k + k));
  if (!p.part) return set_error(ctx, PLANC_ENOMEM, "mode_update: workspace");
#define ROWS(KMV)                                                                    \
  do {                                                                               \
    if (kind == 0) mode_rows_kernel<KMV, 0, 64, 0><<<(unsigned)nb, 256, 0, ctx->stream>>>(p);      \
    else if (kind == 1) mode_rows_kernel<KMV, 1, 64, 0><<<(unsigned)nb, 256, 0, ctx->stream>>>(p); \
    else mode_rows_kernel<KMV, 2, 64, 0><<<(unsigned)nb, 256, 0, ctx->stream>>>(p);                \
  } while (0)
  if (k <= 16) ROWS(16);
  else if (k <= 32) ROWS(32);
  else if (k <= 48) ROWS(48);
  else ROWS(64);
#undef ROWS
  PLANC_LAUNCH_CHECK(ctx);
  long fb = (I * k + 255) / 256;
  if (fb > 1024) fb = 1024;
  mode_finalize_kernel<<<(unsigned)fb, 256, 0, ctx->stream>>>(p, (int)nb);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

}  // namespace planc
