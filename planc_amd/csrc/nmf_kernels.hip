// Small-k kernels of the NMF/NTF loops: Grams (FP64 accumulation), the fused
// MU and HALS updates, column normalisation, FP64 reductions for the error
// terms and the 16-bit conversion of A.
//
// Row-parallel updates keep one row of the factor (k <= 64 values) in
// registers and the k x k Gram in LDS, so an update costs one read and one
// write of the factor plus one read of the numerator.
#include "common.h"
#include "rowupd.h"

namespace planc {

namespace {

constexpr int NTHR = 256;

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Deterministic block sum (fixed shuffle pattern, waves summed in order).
__device__ inline double block_sum(double v, double *red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; ++w) s += red[w];
  return s;
}

// ---------------------------------------------------------------------------
// Read-bandwidth probe: 16-byte grid-stride loads, one per iteration (the
// pattern that measured fastest on MI355X), summed so the loads stay live.
__global__ __launch_bounds__(NTHR) void read_probe_kernel(const floatx4 *p, long n4, float *out) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long stride = (long)gridDim.x * blockDim.x;
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  for (long i = gid; i < n4; i += stride) acc += p[i];
  out[gid] = acc[0] + acc[1] + acc[2] + acc[3];
}

// ---------------------------------------------------------------------------
// G = X^T X on FP64 MFMA (v_mfma_f64_16x16x4f64).  A block is four groups
// of four waves (1024 threads: four waves per SIMD, so one wave's dependent
// MFMAs and LDS reads overlap another's); at step s group g takes the
// 64-row chunk (4 s + g) gridDim.x + blockIdx.x through its own LDS tile
// (the next chunk's loads in flight during the current chunk's MFMAs), so a
// small X still spreads over all blocks.  Wave w of a
// group owns the upper-triangle 16 x 16 tiles w, w+4, ... of the KM x KM
// Gram; its tiles are interleaved and each goes through two accumulators
// (even and odd row steps), so consecutive MFMAs are independent.  The four
// groups' accumulators are combined in LDS as (g0 + g2) + (g1 + g3) and the
// block writes one full KM x KM partial (mirrored), summed by
// gram_reduce_kernel.
constexpr int GCR = 64;
constexpr int GGROUPS = 4;

template <int KM>
__global__ __launch_bounds__(NTHR * GGROUPS) void gram_mfma_kernel(const float *X, long rows, int k, long ldx,
                                                                   double *part) {
  constexpr int NT = KM / 16, NUP = NT * (NT + 1) / 2, TPW = (NUP + 3) / 4;
  constexpr int PER = GCR * KM / NTHR;  // chunk values per thread
  constexpr int XS = KM * (GCR + 1);    // floats of one group's tile
  static_assert(GGROUPS * XS * 4 >= 2 * NTHR * TPW * 32, "the combine buffer must fit the tiles");
  __shared__ __attribute__((aligned(32))) float Xall[GGROUPS * XS];
  const int tid = threadIdx.x, grp = tid / NTHR, gt = tid - grp * NTHR;
  const int lane = tid & 63, wave = gt >> 6;
  float(*Xs)[GCR + 1] = reinterpret_cast<float(*)[GCR + 1]>(Xall + grp * XS);
  auto tile_of = [&](int t, int &ta, int &tb) {
    ta = 0;
    int rem = t;
    while (rem >= NT - ta) {
      rem -= NT - ta;
      ++ta;
    }
    tb = ta + rem;
  };
  auto load = [&](long r0, float(&v)[PER]) {
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int e = gt + NTHR * j, r = e % GCR, c = e / GCR;
      const long gr = r0 + r;
      v[j] = (gr < rows && c < k) ? X[gr + ldx * c] : 0.0f;
    }
  };
  const float *ua[TPW], *ub[TPW];
  bool on[TPW];
#pragma unroll
  for (int j = 0; j < TPW; ++j) {
    const int t = wave + 4 * j;
    on[j] = t < NUP;
    int ta = 0, tb = 0;
    if (on[j]) tile_of(t, ta, tb);
    ua[j] = &Xs[16 * ta + (lane & 15)][0];
    ub[j] = &Xs[16 * tb + (lane & 15)][0];
  }
  doublex4 acc[TPW], acc2[TPW];
#pragma unroll
  for (int j = 0; j < TPW; ++j) {
    acc[j] = doublex4{0.0, 0.0, 0.0, 0.0};
    acc2[j] = doublex4{0.0, 0.0, 0.0, 0.0};
  }
  // Every group runs the same number of chunk steps (the barriers are
  // block-wide); a group past the last chunk loads zeros.
  const long cstride = (long)gridDim.x * GGROUPS * GCR;
  const long nsteps = (rows + cstride - 1) / cstride;
  float v[PER];
  long r0 = ((long)grp * gridDim.x + blockIdx.x) * GCR;
  load(r0, v);
  for (long st = 0; st < nsteps; ++st, r0 += cstride) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int e = gt + NTHR * j;
      Xs[e / GCR][e % GCR] = v[j];
    }
    __syncthreads();
    if (st + 1 < nsteps) load(r0 + cstride, v);
#pragma unroll
    for (int r = 0; r < GCR / 4; r += 2) {
      const int ra = (lane >> 4) + 4 * r, rb = ra + 4;
#pragma unroll
      for (int j = 0; j < TPW; ++j)
        if (on[j]) {
          acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)ua[j][ra], (double)ub[j][ra], acc[j], 0, 0, 0);
          acc2[j] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)ua[j][rb], (double)ub[j][rb], acc2[j], 0, 0, 0);
        }
    }
  }
#pragma unroll
  for (int j = 0; j < TPW; ++j) acc[j] += acc2[j];
  // Combine the groups: (g0 + g2) + (g1 + g3), through the tile space.
  doublex4 *cb = reinterpret_cast<doublex4 *>(Xall);  // [2][NTHR][TPW]
  auto slot = [&](int half, int j) -> doublex4 & { return cb[(half * NTHR + gt) * TPW + j]; };
  __syncthreads();
  if (grp >= 2)
#pragma unroll
    for (int j = 0; j < TPW; ++j) slot(grp - 2, j) = acc[j];
  __syncthreads();
  if (grp < 2)
#pragma unroll
    for (int j = 0; j < TPW; ++j) acc[j] += slot(grp, j);
  __syncthreads();
  if (grp == 1)
#pragma unroll
    for (int j = 0; j < TPW; ++j) slot(0, j) = acc[j];
  __syncthreads();
  if (grp != 0) return;
#pragma unroll
  for (int j = 0; j < TPW; ++j) acc[j] += slot(0, j);
  // FP64 D layout: lane holds D(lane / 16 + 4 v, lane % 16).
  double *dst = part + (long)blockIdx.x * KM * KM;
#pragma unroll
  for (int j = 0; j < TPW; ++j) {
    const int t = wave + 4 * j;
    if (t < NUP) {
      int ta, tb;
      tile_of(t, ta, tb);
      const int b = 16 * tb + (lane & 15);
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        const int a = 16 * ta + (lane >> 4) + 4 * w;
        if (ta != tb || a <= b) {
          dst[a + KM * b] = acc[j][w];
          dst[b + KM * a] = acc[j][w];
        }
      }
    }
  }
}

// Block of 256 threads per GR_E = 16 consecutive entries e of the kp x kp
// partial: thread (e, j) sums partials j, j + 16, ... (up to GR_U of them
// with every load in flight at once), and the 16 sums of an entry are added
// in j order.
constexpr int GR_E = 16, GR_J = NTHR / GR_E, GR_U = 16;
constexpr int gram_reduce_blocks(int kp) { return (kp * kp + GR_E - 1) / GR_E; }
__global__ __launch_bounds__(NTHR) void gram_reduce_kernel(const double *part, long nparts, int k, int kp,
                                                           double *G) {
  __shared__ double red[GR_J][GR_E];
  const int el = threadIdx.x % GR_E, j = threadIdx.x / GR_E;
  const long e = (long)blockIdx.x * GR_E + el;
  const long kk = (long)kp * kp;
  double s = 0.0;
  if (e < kk) {
    double v[GR_U];
#pragma unroll
    for (int u = 0; u < GR_U; ++u) {
      const long b = j + (long)GR_J * u;
      v[u] = b < nparts ? part[b * kk + e] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < GR_U; ++u) s += v[u];
    for (long b = j + (long)GR_J * GR_U; b < nparts; b += GR_J) s += part[b * kk + e];  // longer lists
  }
  red[j][el] = s;
  __syncthreads();
  if (j == 0 && e < kk) {
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < GR_J; ++q) t += red[q][el];
    const int a = (int)(e % kp), b = (int)(e / kp);
    if (a < k && b < k) G[a + (long)k * b] = t;
  }
}

// ---------------------------------------------------------------------------
// Loads the k x k FP64 matrix G into LDS as floats, transposed and zero
// padded: Gt[x][t] = G(t, x), so the column G(:, x) that a row update
// contracts with is one contiguous row of 16-byte words (uniform address:
// one LDS broadcast per word for the whole wave).  NTH = blockDim.x: every
// thread's loads are issued before the first store (one memory round trip).
template <int KM, int NTH>
__device__ inline void load_gram(float (&Gt)[KM][KM], const double *G, int k) {
  constexpr int PER = (KM * KM + NTH - 1) / NTH;
  double v[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int e = threadIdx.x + NTH * j, t = e % KM, x = e / KM;
    v[j] = (e < KM * KM && t < k && x < k) ? G[t + (size_t)k * x] : 0.0;
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int e = threadIdx.x + NTH * j;
    if (e < KM * KM) Gt[e / KM][e % KM] = (float)v[j];
  }
  __syncthreads();
}

// s + sum_t h[t] * G(t, x), in t order, from the transposed LDS tile.
template <int KM>
__device__ inline float dot_gcol(const float (&h)[KM], const float (&Gt)[KM][KM], int x, float s) {
  const floatx4 *g = (const floatx4 *)&Gt[x][0];
#pragma unroll
  for (int q = 0; q < KM / 4; ++q) {
    const floatx4 v = g[q];
    s = fmaf(h[4 * q + 0], v[0], s);
    s = fmaf(h[4 * q + 1], v[1], s);
    s = fmaf(h[4 * q + 2], v[2], s);
    s = fmaf(h[4 * q + 3], v[3], s);
  }
  return s;
}

// MU (kind 0: H <- H .* N ./ (H*G + eps)) and HALS rows (kind 1,
// nmf/hals.hpp:67-75, ntf/ntfhals.hpp:13-29: columns in order, each using
// the already-updated columns before it), N(j, c) at N[j*nrs + c*ncs].  A
// quad of lanes per row (quad_row_update), 64 rows per block.
// KIND is a template argument so that each kernel carries only its own
// rule's registers (the HALS sweep keeps more Gram rows in flight).
template <int KM, int KIND>
__global__ __launch_bounds__(NTHR) void rows_quad_kernel(float *H, long n, int k, long ldh,
                                                         const float *N, long nrs, long ncs,
                                                         const double *G, float eps) {
  constexpr int kind = KIND;
  constexpr int Q = KM / 4;
  __shared__ __attribute__((aligned(16))) float Gs[KM][KM];
  load_gram<KM, NTHR>(Gs, G, k);
  const int q = threadIdx.x & 3;
  const long j = (long)blockIdx.x * (NTHR / 4) + (threadIdx.x >> 2);
  if (j >= n) return;
  float h[Q], nv[Q];
#pragma unroll
  for (int u = 0; u < Q; ++u) {
    const int t = q * Q + u;
    const int tc = t < k ? t : k - 1;
    h[u] = t < k ? H[j + ldh * tc] : 0.0f;
    nv[u] = t < k ? N[j * nrs + tc * ncs] : 0.0f;
  }
  quad_row_update<KM, false>(h, nv, Gs, q, k, kind, eps);
#pragma unroll
  for (int u = 0; u < Q; ++u) {
    const int t = q * Q + u;
    if (t < k) H[j + ldh * t] = h[u];
  }
}

// Row update fused with the FP64 Gram of the updated rows and, optionally,
// the FP64 dot <updated rows, N> (the error's cross term): blocks of RG
// groups of 256 threads; a group takes a 64-row chunk per step (chunk (grp
// gridDim.x + blockIdx.x) + step RG gridDim.x), loads the chunk of H and of
// N with 16-byte loads into LDS tiles (column c of the chunk contiguous:
// coalesced for a column-major H / N, and for a row-major N, NROW), updates
// the rows in the quad layout (quad_row_update, the rule of
// rows_quad_kernel), stores them back through the tile with 16-byte stores
// and accumulates the tile's Gram on FP64 MFMA as gram_mfma_kernel does.
// The groups' accumulators are combined in LDS; each block writes one
// KM x KM partial (mirrored, summed by gram_reduce_kernel) and one dot
// partial.  The factor makes one round trip through HBM instead of three
// (update, Gram, dot).  Requires ldh % 4 == 0, ldn % 4 == 0 and 16-byte
// aligned H and N.
constexpr int RG = 2;  // groups per block (1 block of 512 threads per CU: the LDS tiles)

template <int KM, int KIND, int NROW>
__global__ __launch_bounds__(NTHR * RG) void rows_gram_kernel(float *H, long n, int k, long ldh,
                                                              const float *N, long ldn, const double *G,
                                                              float eps, double *gpart, double *dpart) {
  constexpr int Q = KM / 4;
  constexpr int NT = KM / 16, NUP = NT * (NT + 1) / 2, TPW = (NUP + 3) / 4;
  constexpr int PER = GCR * KM / 4 / NTHR;  // 16-byte pieces of a tile per thread
  constexpr int XS = KM * (GCR + 1);        // floats of one tile
  static_assert(2 * RG * XS * 4 >= NTHR * TPW * 32, "the combine buffer must fit the tiles");
  // KIND 1: the Gram transposed for quad_row_update (Gs[x][t] = G(t, x)).
  // KIND 0: G (symmetric) with rows padded to KM + 16 floats, read as the
  // MFMA B operand G(4 s + l / 16, 16 tj + l % 16): the four k rows of a
  // read land 16 banks apart (two-way at most, no wider conflict).
  constexpr int GP = KIND == 0 ? KM + 16 : KM;
  __shared__ __attribute__((aligned(16))) float Gs[KM][GP];
  __shared__ __attribute__((aligned(32))) float tiles[RG][2][XS];
  __shared__ double red[NTHR * RG / 64];
  const int tid = threadIdx.x, grp = tid / NTHR, gt = tid - grp * NTHR;
  const int lane = tid & 63, wave = gt >> 6;
  const int q = gt & 3, rr = gt >> 2;
  float(*Hs)[GCR + 1] = reinterpret_cast<float(*)[GCR + 1]>(&tiles[grp][0][0]);
  float(*Ns)[GCR + 1] = reinterpret_cast<float(*)[GCR + 1]>(&tiles[grp][1][0]);
  auto tile_of = [&](int t, int &ta, int &tb) {
    ta = 0;
    int rem = t;
    while (rem >= NT - ta) {
      rem -= NT - ta;
      ++ta;
    }
    tb = ta + rem;
  };
  // Piece p = gt + NTHR j of a column-major tile: column p / 16, rows 4 (p %
  // 16) .. + 3; of a row-major one: row p / (KM / 4), columns 4 (p % (KM /
  // 4)) .. + 3.
  auto load_cm = [&](const float *X, long ldx, long r0, floatx4 (&v)[PER]) {
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int p = gt + NTHR * j, c = p >> 4, r4 = 4 * (p & 15);
      const long gr = r0 + r4;
      floatx4 x = {0.f, 0.f, 0.f, 0.f};
      if (c < k) {
        if (gr + 3 < n) {
          x = *(const floatx4 *)(X + gr + ldx * c);
        } else {
#pragma unroll
          for (int u = 0; u < 4; ++u) x[u] = gr + u < n ? X[gr + u + ldx * c] : 0.0f;
        }
      }
      v[j] = x;
    }
  };
  auto load_rm = [&](const float *X, long ldx, long r0, floatx4 (&v)[PER]) {
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int p = gt + NTHR * j, r = p / (KM / 4), c4 = 4 * (p % (KM / 4));
      const long gr = r0 + r;
      floatx4 x = {0.f, 0.f, 0.f, 0.f};
      if (gr < n) {
        if (c4 + 3 < k) {
          x = *(const floatx4 *)(X + gr * ldx + c4);
        } else {
#pragma unroll
          for (int u = 0; u < 4; ++u) x[u] = c4 + u < k ? X[gr * ldx + c4 + u] : 0.0f;
        }
      }
      v[j] = x;
    }
  };
  auto put_cm = [&](float(*T)[GCR + 1], const floatx4 (&v)[PER]) {
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int p = gt + NTHR * j, c = p >> 4, r4 = 4 * (p & 15);
#pragma unroll
      for (int u = 0; u < 4; ++u) T[c][r4 + u] = v[j][u];
    }
  };
  auto put_rm = [&](float(*T)[GCR + 1], const floatx4 (&v)[PER]) {
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int p = gt + NTHR * j, r = p / (KM / 4), c4 = 4 * (p % (KM / 4));
#pragma unroll
      for (int u = 0; u < 4; ++u) T[c4 + u][r] = v[j][u];
    }
  };
  const float *ua[TPW], *ub[TPW];
  bool on[TPW];
#pragma unroll
  for (int j = 0; j < TPW; ++j) {
    const int t = wave + 4 * j;
    on[j] = t < NUP;
    int ta = 0, tb = 0;
    if (on[j]) tile_of(t, ta, tb);
    ua[j] = &Hs[16 * ta + (lane & 15)][0];
    ub[j] = &Hs[16 * tb + (lane & 15)][0];
  }
  doublex4 acc[TPW];
#pragma unroll
  for (int j = 0; j < TPW; ++j) acc[j] = doublex4{0.0, 0.0, 0.0, 0.0};
  double dsum = 0.0;
  // Every group runs the same number of steps (block-wide barriers); a group
  // past the last chunk works on zero rows (no stores, zero contributions).
  // The next chunk's loads are issued after the row update (fewer registers
  // live through it) and stay in flight through the stores and the Gram.
  const long cstride = (long)gridDim.x * RG * GCR;
  const long nsteps = (n + cstride - 1) / cstride;
  long r0 = ((long)grp * gridDim.x + blockIdx.x) * GCR;
  floatx4 vh[PER], vn[PER];
  load_cm(H, ldh, r0, vh);  // in flight with the Gram's loads
  if (NROW) load_rm(N, ldn, r0, vn);
  else load_cm(N, ldn, r0, vn);
  if constexpr (KIND == 0) {
    constexpr int GPER = (KM * KM + NTHR * RG - 1) / (NTHR * RG);
    double gv[GPER];
#pragma unroll
    for (int j = 0; j < GPER; ++j) {
      const int e = tid + NTHR * RG * j, t = e % KM, x = e / KM;
      gv[j] = (e < KM * KM && t < k && x < k) ? G[t + (size_t)k * x] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < GPER; ++j) {
      const int e = tid + NTHR * RG * j;
      if (e < KM * KM) Gs[e / KM][e % KM] = (float)gv[j];
    }
    __syncthreads();
  } else {
    load_gram<KM, NTHR * RG>(reinterpret_cast<float(&)[KM][KM]>(Gs), G, k);
  }
  for (long st = 0; st < nsteps; ++st, r0 += cstride) {
    __syncthreads();
    put_cm(Hs, vh);
    if (NROW) put_rm(Ns, vn);
    else put_cm(Ns, vn);
    __syncthreads();
    if constexpr (KIND == 0) {
      // MU: the denominators H G of the chunk on FP32 MFMA.  Wave w owns
      // rows 16 w .. 16 w + 15 of the tile (it reads and rewrites only
      // those): D(i, j) = sum_t H(16 w + i, t) G(t, 16 tj + j), lane l
      // supplying H(16 w + l % 16, 4 s + l / 16) and G(4 s + l / 16, 16 tj +
      // l % 16) per 4-k step s, and holding D(4 (l / 16) + v, l % 16).
      floatx4 den[NT];
#pragma unroll
      for (int tj = 0; tj < NT; ++tj) den[tj] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int st4 = 0; st4 < KM / 4; ++st4) {
        const int kq = 4 * st4 + (lane >> 4);
        const float a = Hs[kq][16 * wave + (lane & 15)];
#pragma unroll
        for (int tj = 0; tj < NT; ++tj)
          den[tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, Gs[kq][16 * tj + (lane & 15)], den[tj], 0, 0, 0);
      }
#pragma unroll
      for (int tj = 0; tj < NT; ++tj) {
        const int c = 16 * tj + (lane & 15);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = 16 * wave + 4 * (lane >> 4) + v;
          const float hv0 = Hs[c][r], nv0 = Ns[c][r];
          const float hv = (r0 + r < n && c < k) ? (hv0 * nv0) / (den[tj][v] + eps) : 0.0f;
          Hs[c][r] = hv;
          dsum += (double)hv * (double)nv0;
        }
      }
    } else {
      float h[Q], nv[Q];
#pragma unroll
      for (int u = 0; u < Q; ++u) {
        h[u] = Hs[q * Q + u][rr];
        nv[u] = Ns[q * Q + u][rr];
      }
      quad_row_update<KM, false>(h, nv, reinterpret_cast<const float(&)[KM][KM]>(Gs), q, k, KIND, eps);
      const bool valid = r0 + rr < n;
#pragma unroll
      for (int u = 0; u < Q; ++u) {
        const float hv = valid ? h[u] : 0.0f;  // rows past n stay zero (HALS floors them)
        Hs[q * Q + u][rr] = hv;
        dsum += (double)hv * (double)nv[u];
      }
    }
    __syncthreads();
    if (st + 1 < nsteps) {
      load_cm(H, ldh, r0 + cstride, vh);
      if (NROW) load_rm(N, ldn, r0 + cstride, vn);
      else load_cm(N, ldn, r0 + cstride, vn);
    }
    // Store the updated chunk (column pieces of 16 bytes).
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int p = gt + NTHR * j, c = p >> 4, r4 = 4 * (p & 15);
      const long gr = r0 + r4;
      if (c < k) {
        const floatx4 x = {Hs[c][r4], Hs[c][r4 + 1], Hs[c][r4 + 2], Hs[c][r4 + 3]};
        if (gr + 3 < n) {
          *(floatx4 *)(H + gr + ldh * c) = x;
        } else {
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (gr + u < n) H[gr + u + ldh * c] = x[u];
        }
      }
    }
    {
      doublex4 acc2[TPW];  // odd row steps (independent consecutive MFMAs)
#pragma unroll
      for (int j = 0; j < TPW; ++j) acc2[j] = doublex4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int r = 0; r < GCR / 4; r += 2) {
        const int ra = (lane >> 4) + 4 * r, rb = ra + 4;
#pragma unroll
        for (int j = 0; j < TPW; ++j)
          if (on[j]) {
            acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)ua[j][ra], (double)ub[j][ra], acc[j], 0, 0, 0);
            acc2[j] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)ua[j][rb], (double)ub[j][rb], acc2[j], 0, 0, 0);
          }
      }
#pragma unroll
      for (int j = 0; j < TPW; ++j) acc[j] += acc2[j];
    }
  }
  // Dot partial of the block (fixed order: waves in order).
  if (dpart) {
    const double bs = block_sum(dsum, red);
    if (tid == 0) dpart[blockIdx.x] = bs;
  }
  // Combine the groups in order (g0 + g1 (+ ...)) through the tile space.
  doublex4 *cb = reinterpret_cast<doublex4 *>(&tiles[0][0][0]);  // [NTHR][TPW]
  for (int g = 1; g < RG; ++g) {
    __syncthreads();
    if (grp == g)
#pragma unroll
      for (int j = 0; j < TPW; ++j) cb[gt * TPW + j] = acc[j];
    __syncthreads();
    if (grp == 0)
#pragma unroll
      for (int j = 0; j < TPW; ++j) acc[j] += cb[gt * TPW + j];
  }
  if (grp != 0) return;
  double *dst = gpart + (long)blockIdx.x * KM * KM;
#pragma unroll
  for (int j = 0; j < TPW; ++j) {
    const int t = wave + 4 * j;
    if (t < NUP) {
      int ta, tb;
      tile_of(t, ta, tb);
      const int b = 16 * tb + (lane & 15);
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        const int a = 16 * ta + (lane >> 4) + 4 * w;
        if (ta != tb || a <= b) {
          dst[a + KM * b] = acc[j][w];
          dst[b + KM * a] = acc[j][w];
        }
      }
    }
  }
}

// Column-sum exchange of a cooperative launch (all blocks co-resident).
// Every slot of a column starts as PART_EMPTY (set before the launch); thread
// 0 of each block writes its partial into its own slot with an agent-scope
// atomic store.  Wave 0 of every block polls the column's slots with
// agent-scope atomic loads until none is empty, summing the values of the
// same poll in a fixed order (identical in every block): the partials
// travel with the readiness test, one round trip after the last store.  The
// wait is bounded: on time-out *err is set, later exchanges return at once
// and the caller reports the failure.
constexpr unsigned long long PART_EMPTY = ~0ull;  // a NaN no arithmetic produces (all payload bits set)

__device__ inline bool column_exchange(double bs, double *part_col, int *err, double &total) {
  __shared__ int s_ok;
  __shared__ double s_tot;
  unsigned long long *slot = reinterpret_cast<unsigned long long *>(part_col);
  if (threadIdx.x == 0)
    __hip_atomic_store(&slot[blockIdx.x], __builtin_bit_cast(unsigned long long, bs), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  if (threadIdx.x < 64) {
    const int nb = (int)gridDim.x, lane = (int)threadIdx.x;
    int ok = 1;
    double tot = 0.0;
    for (int spin = 0;; ++spin) {
      int mine = 1;
      double t = 0.0;
      for (int b = lane; b < nb; b += 64) {
        const unsigned long long v = __hip_atomic_load(&slot[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        mine &= v != PART_EMPTY;
        t += __builtin_bit_cast(double, v);
      }
      if (__all(mine)) {
        tot = t;
        break;
      }
      if (spin >= (1 << 20) || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
        if (lane == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    tot = wave_sum(tot);
    if (lane == 0) {
      s_ok = ok;
      s_tot = tot;
    }
  }
  __syncthreads();
  total = s_tot;
  const bool ok = s_ok != 0;
  __syncthreads();
  return ok;
}

__global__ void zero_int_kernel(int *p) { *p = 0; }
__global__ void fill_u64_kernel(unsigned long long *p, long n, unsigned long long v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// HALS over the columns of W with per-column normalisation (nmf/hals.hpp:
// 89-102).  Persistent cooperative kernel: each thread owns RPT rows of W in
// registers; one column_exchange per column x publishes the column's sum of
// squares (block partials summed in the same order by every block, so every
// block derives the identical norm).  part: k x gridDim.x slots, PART_EMPTY
// at launch; sync[1] the time-out flag.
// BS: threads per block; 1024-thread blocks make four times fewer exchange
// participants.
template <int KM, int RPT, int BS>
__global__ __launch_bounds__(BS) void hals_cols_coop_kernel(float *W, long m, int k, long ldw,
                                                            const float *AH, long ldah,
                                                            const double *G, double *part, int *sync) {
  __shared__ __attribute__((aligned(16))) float Gs[KM][KM];
  __shared__ double red[BS / 64];
  load_gram<KM, BS>(Gs, G, k);
  const long total = (long)gridDim.x * blockDim.x;
  const long gtid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const float floor_v = 1e-17f;
  float w[RPT][KM];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    long i = gtid + r * total;
#pragma unroll
    for (int t = 0; t < KM; ++t) w[r][t] = (i < m && t < k) ? W[i + ldw * t] : 0.0f;
  }
  // AH(i, x) of the next column is loaded one column ahead, so its latency
  // hides behind the current column's barrier instead of stalling the
  // column update.
  float ahn[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    long i = gtid + r * total;
    ahn[r] = i < m ? AH[i] : 0.0f;
  }
  // The column loop is not unrolled (it holds a grid barrier); registers are
  // only indexed by the unrolled t loops, w[r][x] is read and written through
  // selects.
  for (int x = 0; x < k; ++x) {
    float u[RPT];
    double lsum = 0.0;
    const float gxx = Gs[x][x];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      long i = gtid + r * total;
      u[r] = 0.0f;
      float wx = 0.0f;
#pragma unroll
      for (int t = 0; t < KM; ++t) wx = (t == x) ? w[r][t] : wx;
      const float s = dot_gcol<KM>(w[r], Gs, x, 0.0f);
      const float ax = ahn[r];
      if (i < m && x + 1 < k) ahn[r] = AH[i + ldah * (x + 1)];
      if (i < m) {
        float v = (wx * gxx) + (ax - s);
        v = (v < floor_v) ? floor_v : v;
        u[r] = v;
        lsum += (double)v * (double)v;
      }
    }
    double bs = block_sum(lsum, red);
    double tot = 0.0;
    if (!column_exchange(bs, part + (size_t)x * gridDim.x, sync + 1, tot)) return;
    double nrm = sqrt(tot);
    if (nrm != 0.0) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        float nv = (float)((double)u[r] / nrm);
#pragma unroll
        for (int t = 0; t < KM; ++t) w[r][t] = (t == x) ? nv : w[r][t];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    long i = gtid + r * total;
    if (i < m) {
#pragma unroll
      for (int t = 0; t < KM; ++t)
        if (t < k) W[i + ldw * t] = w[r][t];
    }
  }
}

// Fallback for any m: one column per pair of launches.  Step 1 computes the
// column update for every row into W(:,x) and block sums of squares.
template <int KM>
__global__ __launch_bounds__(NTHR) void hals_col_step_kernel(float *W, long m, int k, long ldw,
                                                             const float *AH, long ldah,
                                                             const double *G, int x,
                                                             double *part) {
  __shared__ __attribute__((aligned(16))) float Gs[KM][KM];
  __shared__ double red[NTHR / 64];
  load_gram<KM, NTHR>(Gs, G, k);
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v2 = 0.0;
  if (i < m) {
    float s = 0.0f;
    for (int t = 0; t < k; ++t) s = fmaf(W[i + ldw * t], Gs[x][t], s);
    float v = (W[i + ldw * x] * Gs[x][x]) + (AH[i + ldah * x] - s);
    v = (v < 1e-17f) ? 1e-17f : v;
    W[i + ldw * x] = v;
    v2 = (double)v * v;
  }
  double bs = block_sum(v2, red);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

__global__ void hals_col_scale_kernel(float *W, long m, long ldw, int x, const double *part,
                                      long nparts) {
  __shared__ double nrm_s;
  if (threadIdx.x < 64) {
    double t = 0.0;
    for (long b = threadIdx.x; b < nparts; b += 64) t += part[b];
    t = wave_sum(t);
    if (threadIdx.x == 0) nrm_s = sqrt(t);
  }
  __syncthreads();
  double nrm = nrm_s;
  if (nrm == 0.0) return;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (long)gridDim.x * blockDim.x)
    W[i + ldw * x] = (float)((double)W[i + ldw * x] / nrm);
}

// ---------------------------------------------------------------------------
// Column sums of squares of X (rows x k): partial[chunk*k + c].
constexpr long COL_CHUNK = 1024;

__global__ __launch_bounds__(NTHR) void colsumsq_partial_kernel(const float *X, long rows, int k,
                                                                long ldx, double *part) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  long r0 = (long)blockIdx.x * COL_CHUNK;
  long r1 = min(rows, r0 + COL_CHUNK);
  for (int c = wave; c < k; c += NTHR / 64) {
    double s = 0.0;
    long r = r0 + lane;
    for (; r + 192 < r1; r += 256) {
      const double v0 = X[r + ldx * c], v1 = X[r + 64 + ldx * c];
      const double v2 = X[r + 128 + ldx * c], v3 = X[r + 192 + ldx * c];
      s += v0 * v0;
      s += v1 * v1;
      s += v2 * v2;
      s += v3 * v3;
    }
    for (; r < r1; r += 64) {
      double v = X[r + ldx * c];
      s += v * v;
    }
    s = wave_sum(s);
    if (lane == 0) part[(size_t)blockIdx.x * k + c] = s;
  }
}

// NCPFactors::normalize() of the initial factors (ntf/auntf.hpp:88) and
// their Grams for factors of at most INIT_ROWS rows and INIT_VALS values, in
// two launches for all modes.  init_gram_kernel, INIT_GS blocks per mode:
// every block reads the factor into LDS (column pitch I + 1, which spreads
// the columns of one row over the banks for even I), computes the column
// norms there exactly as col_norms computes them for one chunk (a wave per
// column, the same lane order and wave sum, so the same norms), divides its
// LDS copy by them as scale_cols does, and forms its share of the FP64 Gram
// of the normalised factor; block 0 of a mode stores the norms.
// init_scale_kernel then divides the factors in HBM by the norms (after
// every block has read them) and forms lambda.  This replaces a run-once
// prologue of 6 launches per mode, each a few microseconds of latency.
constexpr int INIT_ROWS = (int)COL_CHUNK;
constexpr int INIT_VALS = 16384;
constexpr int INIT_GS = 8;
struct InitFacs {
  float *f[8];
  double *g[8];
  int dim[8];
};

__global__ __launch_bounds__(NTHR) void init_gram_kernel(InitFacs fs, int k, double *norms) {
  __shared__ float Fs[INIT_VALS + 64];
  __shared__ double nrm[64];
  const int m = blockIdx.x, part = blockIdx.y, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const float *F = fs.f[m];
  const int I = fs.dim[m], P = I + 1, tot = I * k;
  // Sixteen loads in flight per thread (a loop of single loads waits a
  // memory round trip per element).
  for (int e0 = threadIdx.x; e0 < tot; e0 += 16 * NTHR) {
    float v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int e = e0 + u * NTHR;
      v[u] = e < tot ? F[e] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int e = e0 + u * NTHR;
      if (e < tot) Fs[e + e / I] = v[u];  // (e - I c) + P c
    }
  }
  __syncthreads();
  for (int c = wave; c < k; c += NTHR / 64) {
    const float *col = Fs + P * c;
    double s = 0.0;
    int r = lane;
    for (; r + 192 < I; r += 256) {
      const double v0 = col[r], v1 = col[r + 64];
      const double v2 = col[r + 128], v3 = col[r + 192];
      s += v0 * v0;
      s += v1 * v1;
      s += v2 * v2;
      s += v3 * v3;
    }
    for (; r < I; r += 64) {
      double v = col[r];
      s += v * v;
    }
    s = wave_sum(s);
    if (lane == 0) {
      const double nv = sqrt(0.0 + s);
      nrm[c] = nv;
      if (part == 0) norms[m * k + c] = nv;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < tot; e += NTHR) {
    const int c = e / I;
    const double nv = nrm[c];
    if (nv > 0) Fs[e + c] = (float)((double)Fs[e + c] / nv);
  }
  __syncthreads();
  double *G = fs.g[m];
  for (int e = part * NTHR + threadIdx.x; e < k * k; e += INIT_GS * NTHR) {
    const int a = e % k, b = e / k;
    if (a > b) continue;
    const float *fa = Fs + P * a, *fb = Fs + P * b;
    // Four interleaved partial sums: the LDS loads of several rows stay in
    // flight instead of one dependent FP64 chain per row.
    double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
    int r = 0;
#pragma unroll 4
    for (; r + 3 < I; r += 4) {
      t0 += (double)fa[r] * (double)fb[r];
      t1 += (double)fa[r + 1] * (double)fb[r + 1];
      t2 += (double)fa[r + 2] * (double)fb[r + 2];
      t3 += (double)fa[r + 3] * (double)fb[r + 3];
    }
    for (; r < I; ++r) t0 += (double)fa[r] * (double)fb[r];
    const double t = (t0 + t1) + (t2 + t3);
    G[a + k * b] = t;
    G[b + k * a] = t;
  }
}

// The factors divided by their norms (scale_cols' arithmetic), INIT_GS
// blocks per mode, each thread's loads issued together; block (0, 0) also
// forms lambda = the product of the modes' norms in mode order (lambda_mul's
// sequence: the first mode's norms, then multiplied by each further mode's).
__global__ __launch_bounds__(NTHR) void init_scale_kernel(InitFacs fs, int n, int k, const double *norms,
                                                         double *lambda) {
  const int m = blockIdx.x;
  float *F = fs.f[m];
  const int I = fs.dim[m], tot = I * k;
  const double *nm = norms + m * k;
  constexpr int U = INIT_VALS / (INIT_GS * NTHR);  // elements per thread at most
  float v[U];
  double nv[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int e = (blockIdx.y * U + u) * NTHR + threadIdx.x;
    v[u] = e < tot ? F[e] : 0.0f;
    nv[u] = e < tot ? nm[e / I] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int e = (blockIdx.y * U + u) * NTHR + threadIdx.x;
    if (e < tot && nv[u] > 0) F[e] = (float)((double)v[u] / nv[u]);
  }
  if (m == 0 && blockIdx.y == 0 && threadIdx.x < k) {
    const int c = threadIdx.x;
    double l = 1.0 * norms[c];
    for (int q = 1; q < n; ++q) l = l * norms[q * k + c];
    lambda[c] = l;
  }
}

// norms[c] = sqrt(sum of partials) in chunk order.
__global__ void colnorm_finalize_kernel(const double *part, long nchunks, int k, double *norms) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k) return;
  double s = 0.0;
  for (long b = 0; b < nchunks; ++b) s += part[(size_t)b * k + c];
  norms[c] = sqrt(s);
}

// X(:,c) /= norms[c] and, if Y, Y(:,c) *= norms[c]; columns with norm 0 untouched.
__global__ void scale_cols_kernel(float *X, long rows, long ldx, float *Y, long yrows, long ldy,
                                  int k, const double *norms) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long tot = rows * k, toty = Y ? yrows * k : 0;
  if (e < tot) {
    int c = (int)(e / rows);
    long r = e % rows;
    double nv = norms[c];
    if (nv > 0) X[r + ldx * c] = (float)((double)X[r + ldx * c] / nv);
  } else if (e < tot + toty) {
    long f = e - tot;
    int c = (int)(f / yrows);
    long r = f % yrows;
    double nv = norms[c];
    if (nv > 0) Y[r + ldy * c] = (float)((double)Y[r + ldy * c] * nv);
  }
}

// ---------------------------------------------------------------------------
// FP64 reductions: sum of X .* Y over a rows x cols pair of column-major
// matrices (Y may equal X for a sum of squares).  Two stages, fixed order.
__global__ __launch_bounds__(NTHR) void dot_partial_kernel(const float *X, long xrs, long xcs,
                                                           const float *Y, long yrs, long ycs,
                                                           long rows, long cols, double *part) {
  __shared__ double red[NTHR / 64];
  long tot = rows * cols;
  double s = 0.0;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    long r = e % rows, c = e / rows;
    s += (double)X[r * xrs + xcs * c] * (double)Y[r * yrs + ycs * c];
  }
  double bs = block_sum(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// Sum of n block partials by one block of NTHR threads: thread j adds
// partials j, j + NTHR, ... with all its loads in flight at once (PSUM_MAX
// per thread), then the fixed-order block sum.
constexpr int PSUM_MAX = 8;  // partials per thread: n <= NTHR * PSUM_MAX
__device__ inline double sum_parts(const double *part, long n, double *red) {
  double v[PSUM_MAX];
#pragma unroll
  for (int u = 0; u < PSUM_MAX; ++u) {
    const long b = threadIdx.x + (long)NTHR * u;
    v[u] = b < n ? part[b] : 0.0;
  }
  double t = 0.0;
#pragma unroll
  for (int u = 0; u < PSUM_MAX; ++u) t += v[u];
  for (long b = threadIdx.x + (long)NTHR * PSUM_MAX; b < n; b += NTHR) t += part[b];  // longer lists
  return block_sum(t, red);
}

__global__ __launch_bounds__(NTHR) void sum_partials_kernel(const double *part, long n, double *out,
                                                           int accumulate) {
  __shared__ double red[NTHR / 64];
  const double t = sum_parts(part, n, red);
  if (threadIdx.x == 0) *out = accumulate ? (*out + t) : t;
}

// NMF relative error from the trace identity (common/nmf.hpp:390-424):
// fit^2 = ||A||^2 - 2 <W, A H> + sum(WtW .* HtH); printObjective prints
// sqrt(fit^2)/||A||, or 1 when fit^2 <= 0 (common/nmf.hpp:462).  <W, A H>
// is *dotv, or (dotv null) the sum of the nparts block partials of a dot
// kernel, taken here as sum_partials_kernel takes it.
__global__ __launch_bounds__(NTHR) void nmf_relerr_kernel(const double *normA2, const double *dotv,
                                                          const double *WtW, const double *HtH, int k,
                                                          double *relerr_out, double *fit_out,
                                                          const double *part, long nparts) {
  __shared__ double red[NTHR / 64];
  // tr(WtW HtH): the loads of the first 16 NTHR products are issued before
  // the partial sum (one round trip for both at k <= 64).
  constexpr int TU = 16;
  double wv[TU], hv[TU];
#pragma unroll
  for (int u = 0; u < TU; ++u) {
    const int e = threadIdx.x + NTHR * u;
    const bool in = e < k * k;
    const int a = in ? e % k : 0, b = in ? e / k : 0;
    wv[u] = in ? WtW[a + (size_t)k * b] : 0.0;
    hv[u] = in ? HtH[b + (size_t)k * a] : 0.0;
  }
  const double dsum = dotv ? 0.0 : sum_parts(part, nparts, red);
  __syncthreads();  // red is reused below
  double tr = 0.0;
#pragma unroll
  for (int u = 0; u < TU; ++u) tr += wv[u] * hv[u];
  for (int e = threadIdx.x + NTHR * TU; e < k * k; e += NTHR) {
    const int a = e % k, b = e / k;
    tr += WtW[a + (size_t)k * b] * HtH[b + (size_t)k * a];
  }
  tr = block_sum(tr, red);
  if (threadIdx.x != 0) return;
  double fit = *normA2 - 2.0 * (dotv ? *dotv : dsum) + tr;
  double nA = sqrt(*normA2);
  double err = (fit > 0) ? sqrt(fit) : nA;
  *relerr_out = err / nA;
  if (fit_out) *fit_out = fit;
}

// out(a, b) = G(a, b) / (s_a s_b), s = norms (1 where a norm is 0): the
// Gram of X after its columns are divided by their norms.
__global__ void gram_rescale_kernel(const double *G, int k, const double *norms, double *out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= k * k) return;
  const int a = e % k, b = e / k;
  const double sa = norms[a] > 0 ? norms[a] : 1.0, sb = norms[b] > 0 ? norms[b] : 1.0;
  out[e] = G[e] / (sa * sb);
}

// norms[c] = sqrt(G(c, c)): the column 2-norms of X from its FP64 Gram.
__global__ void gram_norms_kernel(const double *G, int k, double *norms) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < k) norms[c] = sqrt(G[c + (size_t)k * c]);
}

// G += l2 * I + l1 * ones (applyReg, common/nmf.hpp:76-90; the reference
// keeps the values in float).
__global__ void apply_reg_kernel(double *G, int k, float l2, float l1) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= k * k) return;
  int a = e % k, b = e / k;
  double v = G[e];
  if (l2 > 0 && a == b) v += l2;
  if (l1 > 0) v += l1;
  G[e] = v;
}

// Symmetric regularisation (applySymmetricReg, common/nmf.hpp:123-129):
// G += sr * I and R(j, c) += sr * F(j, c), R's element (j, c) at
// R[j * rs + c * cs], F n x k column-major (ld n).
__global__ void add_diag_kernel(double *G, int k, double v) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < k) G[c + (long)k * c] += v;
}

__global__ void symm_rhs_kernel(float *R, long rs, long cs, const float *F, long n, int k, double sr) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * k) return;
  long j = e % n;
  long c = e / n;
  float *r = R + j * rs + c * cs;
  *r = (float)((double)*r + sr * (double)F[e]);
}

// ---------------------------------------------------------------------------
// 16-bit copies: per-column power-of-two scale then round.
__global__ __launch_bounds__(NTHR) void colmax_scale_kernel(const float *A, long rows, long lda,
                                                            float *scale) {
  __shared__ float red[NTHR / 64];
  long c = blockIdx.x;
  float mx = 0.0f;
  for (long r = threadIdx.x; r < rows; r += blockDim.x) mx = fmaxf(mx, fabsf(A[r + lda * c]));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = 0.0f;
    for (int w = 0; w < NTHR / 64; ++w) m = fmaxf(m, red[w]);
    float s = 1.0f;
    if (m > 0.0f && isfinite(m)) {
      int ex;
      frexpf(m, &ex);  // m = f * 2^ex, f in [0.5, 1)
      s = ldexpf(1.0f, ex);
    }
    scale[c] = s;
  }
}

template <int DT>
__global__ void convert_kernel(const float *A, long rows, long cols, long lda, uint16_t *out,
                               long ldo, const float *scale) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long tot = rows * cols;
  for (; e < tot; e += (long)gridDim.x * blockDim.x) {
    long r = e % rows, c = e / rows;
    float v = A[r + lda * c];
    if (scale) v = v / scale[c];
    out[r + ldo * c] = (DT == PLANC_BF16) ? float_to_bf16(v) : float_to_f16(v);
  }
}

// out (cols x rows, ld ldo) = in^T for a 16-bit in (rows x cols, ld ldi):
// 64 x 64 tiles through LDS, both sides coalesced.
__global__ __launch_bounds__(256) void transpose16_kernel(const uint16_t *in, long rows, long cols,
                                                          long ldi, uint16_t *out, long ldo) {
  __shared__ uint16_t tile[64][66];
  const long r0 = (long)blockIdx.x * 64, c0 = (long)blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int j = ty; j < 64; j += 4) {
    const long r = r0 + tx, c = c0 + j;
    tile[j][tx] = (r < rows && c < cols) ? in[r + ldi * c] : (uint16_t)0;
  }
  __syncthreads();
  for (int j = ty; j < 64; j += 4) {
    const long c = c0 + tx, r = r0 + j;
    if (r < rows && c < cols) out[c + ldo * r] = tile[tx][j];
  }
}

// The same with 16-byte global loads and stores (rows, cols, ldi, ldo all
// multiples of 8 and 16-byte aligned bases): a thread moves 8 consecutive
// elements of a source column in and 8 of a destination column out.
__global__ __launch_bounds__(256) void transpose16_vec_kernel(const uint16_t *in, long rows, long cols,
                                                              long ldi, uint16_t *out, long ldo) {
  __shared__ uint16_t tile[64][66];  // tile[c][r]
  const long r0 = (long)blockIdx.x * 64, c0 = (long)blockIdx.y * 64;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int q = threadIdx.x + 256 * j, c = q >> 3, rc = (q & 7) * 8;
    const long gr = r0 + rc, gc = c0 + c;
    uint4 v = {0u, 0u, 0u, 0u};
    if (gr < rows && gc < cols) v = *(const uint4 *)(in + gr + ldi * gc);
    uint16_t h[8];
    __builtin_memcpy(h, &v, 16);
#pragma unroll
    for (int u = 0; u < 8; ++u) tile[c][rc + u] = h[u];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int q = threadIdx.x + 256 * j, r = q >> 3, cc = (q & 7) * 8;
    const long gr = r0 + r, gc = c0 + cc;
    if (gr < rows && gc < cols) {
      uint16_t h[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) h[u] = tile[cc + u][r];
      *(uint4 *)(out + gc + ldo * gr) = __builtin_bit_cast(uint4, h);
    }
  }
}

// FP64 sum of X .* Y over n contiguous elements (16-byte loads), one partial
// per block.
__global__ __launch_bounds__(NTHR) void dot_flat_kernel(const float *X, const float *Y, long n,
                                                        double *part) {
  __shared__ double red[NTHR / 64];
  const long n4 = n / 4;
  double s = 0.0;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n4; e += stride) {
    const floatx4 x = ((const floatx4 *)X)[e], y = ((const floatx4 *)Y)[e];
    s += (double)x[0] * (double)y[0];
    s += (double)x[1] * (double)y[1];
    s += (double)x[2] * (double)y[2];
    s += (double)x[3] * (double)y[3];
  }
  if (blockIdx.x == 0 && threadIdx.x < n - 4 * n4) {
    const long e = 4 * n4 + threadIdx.x;
    s += (double)X[e] * (double)Y[e];
  }
  double bs = block_sum(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

template <int KM>
int launch_rows(planc_ctx *ctx, int kind, float *H, long n, int k, long ldh, const float *N,
                long nrs, long ncs, const double *G, float eps) {
  const long rows_per_block = NTHR / 4;
  unsigned nb = (unsigned)((n + rows_per_block - 1) / rows_per_block);
  if (nb == 0) return PLANC_OK;
  if (kind == 0)
    rows_quad_kernel<KM, 0><<<nb, NTHR, 0, ctx->stream>>>(H, n, k, ldh, N, nrs, ncs, G, eps);
  else
    rows_quad_kernel<KM, 1><<<nb, NTHR, 0, ctx->stream>>>(H, n, k, ldh, N, nrs, ncs, G, eps);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

template <int KM, int RPT, int BS>
int try_coop(planc_ctx *ctx, float *W, long m, int k, long ldw, const float *AH, long ldah,
             const double *G, bool *done) {
  *done = false;
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, (const void *)hals_cols_coop_kernel<KM, RPT, BS>, BS, 0);
  if (e != hipSuccess || per_cu <= 0) return PLANC_OK;
  long maxblocks = (long)per_cu * ctx->num_cu;
  long need = (m + (long)BS * RPT - 1) / ((long)BS * RPT);
  if (need > maxblocks || need <= 0) return PLANC_OK;
  unsigned nb = (unsigned)need;
  double *part = (double *)ws_get(ctx, 2, sizeof(double) * (size_t)KM * nb);
  int *sync = (int *)ws_get(ctx, 17, 2 * sizeof(int));
  if (!part || !sync) return set_error(ctx, PLANC_ENOMEM, "hals_cols: workspace");
  const long nslots = (long)k * nb;
  fill_u64_kernel<<<(unsigned)((nslots + 255) / 256), 256, 0, ctx->stream>>>((unsigned long long *)part, nslots,
                                                                             PART_EMPTY);
  PLANC_LAUNCH_CHECK(ctx);
  void *args[] = {&W, &m, &k, &ldw, &AH, &ldah, &G, &part, &sync};
  hipError_t le = hipLaunchCooperativeKernel((const void *)hals_cols_coop_kernel<KM, RPT, BS>, dim3(nb),
                                             dim3(BS), args, 0, ctx->stream);
  if (le != hipSuccess) {
    // The runtime can refuse a cooperative launch (e.g. while another
    // process holds queues on the device); nothing was launched, the
    // caller takes the two-launches-per-column path.
    (void)hipGetLastError();
    return PLANC_OK;
  }
  *done = true;
  return PLANC_OK;
}

template <int KM>
int hals_cols_impl(planc_ctx *ctx, float *W, long m, int k, long ldw, const float *AH, long ldah,
                   const double *G) {
  bool done = false;
  int rc = PLANC_OK;
  if constexpr (KM <= 48) {  // at k = 64 a 1024-thread block leaves too few registers per row
    rc = try_coop<KM, 1, 1024>(ctx, W, m, k, ldw, AH, ldah, G, &done);
    if (rc || done) return rc;
  }
  rc = try_coop<KM, 1, NTHR>(ctx, W, m, k, ldw, AH, ldah, G, &done);
  if (rc || done) return rc;
  if (KM <= 32) {
    rc = try_coop<KM, 2, NTHR>(ctx, W, m, k, ldw, AH, ldah, G, &done);
    if (rc || done) return rc;
  }
  // Fallback: two launches per column.
  unsigned nb = (unsigned)((m + NTHR - 1) / NTHR);
  double *part = (double *)ws_get(ctx, 2, sizeof(double) * (size_t)nb);
  if (!part) return set_error(ctx, PLANC_ENOMEM, "hals_cols: workspace");
  for (int x = 0; x < k; ++x) {
    hals_col_step_kernel<KM><<<nb, NTHR, 0, ctx->stream>>>(W, m, k, ldw, AH, ldah, G, x, part);
    PLANC_LAUNCH_CHECK(ctx);
    unsigned sb = nb < 1024 ? nb : 1024;
    hals_col_scale_kernel<<<sb, NTHR, 0, ctx->stream>>>(W, m, ldw, x, part, nb);
    PLANC_LAUNCH_CHECK(ctx);
  }
  return PLANC_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host entry points used by the C ABI and the loop drivers.

int read_bandwidth(planc_ctx *ctx, const void *buf, size_t bytes, int reps, double *gbps) {
  constexpr unsigned NB = 2048;
  float *out = (float *)ws_get(ctx, 3, sizeof(float) * NB * NTHR);
  if (!out) return set_error(ctx, PLANC_ENOMEM, "read_bandwidth: workspace");
  const long n4 = (long)(bytes / 16);
  hipEvent_t e0, e1;
  PLANC_HIP_CHECK(ctx, hipEventCreate(&e0));
  PLANC_HIP_CHECK(ctx, hipEventCreate(&e1));
  read_probe_kernel<<<NB, NTHR, 0, ctx->stream>>>((const floatx4 *)buf, n4, out);  // warm-up
  PLANC_HIP_CHECK(ctx, hipEventRecord(e0, ctx->stream));
  for (int r = 0; r < reps; ++r) read_probe_kernel<<<NB, NTHR, 0, ctx->stream>>>((const floatx4 *)buf, n4, out);
  PLANC_HIP_CHECK(ctx, hipEventRecord(e1, ctx->stream));
  PLANC_HIP_CHECK(ctx, hipEventSynchronize(e1));
  float ms = 0.0f;
  PLANC_HIP_CHECK(ctx, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *gbps = (double)n4 * 16.0 * reps / ((double)ms * 1e-3) / 1e9;
  return PLANC_OK;
}

int gram(planc_ctx *ctx, const float *X, long rows, int k, long ldx, double *G) {
  if (k <= 0 || k > 64 || rows < 0) return set_error(ctx, PLANC_EINVAL, "gram: k must be 1..64");
  const int kp = ((k + 15) / 16) * 16;
  long nchunks = (rows + GCR - 1) / GCR;
  if (nchunks == 0) nchunks = 1;
  long nb = nchunks < ctx->num_cu ? nchunks : ctx->num_cu;
  // One partial per block (each block accumulates its chunks; rows == 0
  // still writes zero partials).
  double *part = (double *)ws_get(ctx, 1, sizeof(double) * (size_t)nb * kp * kp);
  if (!part) return set_error(ctx, PLANC_ENOMEM, "gram: workspace");
  const unsigned nt = NTHR * GGROUPS;
  if (kp == 16)
    gram_mfma_kernel<16><<<(unsigned)nb, nt, 0, ctx->stream>>>(X, rows, k, ldx, part);
  else if (kp == 32)
    gram_mfma_kernel<32><<<(unsigned)nb, nt, 0, ctx->stream>>>(X, rows, k, ldx, part);
  else if (kp == 48)
    gram_mfma_kernel<48><<<(unsigned)nb, nt, 0, ctx->stream>>>(X, rows, k, ldx, part);
  else
    gram_mfma_kernel<64><<<(unsigned)nb, nt, 0, ctx->stream>>>(X, rows, k, ldx, part);
  PLANC_LAUNCH_CHECK(ctx);
  gram_reduce_kernel<<<(unsigned)gram_reduce_blocks(kp), NTHR, 0, ctx->stream>>>(part, nb, k, kp, G);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int row_update(planc_ctx *ctx, int kind, float *H, long n, int k, long ldh, const float *N, long nrs,
               long ncs, const double *G, float eps) {
  if (k <= 16) return launch_rows<16>(ctx, kind, H, n, k, ldh, N, nrs, ncs, G, eps);
  if (k <= 32) return launch_rows<32>(ctx, kind, H, n, k, ldh, N, nrs, ncs, G, eps);
  if (k <= 48) return launch_rows<48>(ctx, kind, H, n, k, ldh, N, nrs, ncs, G, eps);
  if (k <= 64) return launch_rows<64>(ctx, kind, H, n, k, ldh, N, nrs, ncs, G, eps);
  return set_error(ctx, PLANC_EUNSUPPORTED, "row update: k > 64");
}

template <int KM>
int launch_rows_gram(planc_ctx *ctx, int kind, float *H, long n, int k, long ldh, const float *N, long ldn,
                     int nrow, const double *G, float eps, double *gpart, long nb, double *dpart) {
#define RGK(KINDV, NROWV)                                                                                   \
  rows_gram_kernel<KM, KINDV, NROWV><<<(unsigned)nb, NTHR * RG, 0, ctx->stream>>>(H, n, k, ldh, N, ldn, G, \
                                                                                 eps, gpart, dpart)
  if (kind == 0 && nrow) RGK(0, 1);
  else if (kind == 0) RGK(0, 0);
  else if (nrow) RGK(1, 1);
  else RGK(1, 0);
#undef RGK
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int hals_cols(planc_ctx *ctx, float *W, long m, int k, long ldw, const float *AH, long ldah,
              const double *G) {
  if (m <= 0) return PLANC_OK;
  if (k <= 16) return hals_cols_impl<16>(ctx, W, m, k, ldw, AH, ldah, G);
  if (k <= 32) return hals_cols_impl<32>(ctx, W, m, k, ldw, AH, ldah, G);
  if (k <= 48) return hals_cols_impl<48>(ctx, W, m, k, ldw, AH, ldah, G);
  if (k <= 64) return hals_cols_impl<64>(ctx, W, m, k, ldw, AH, ldah, G);
  return set_error(ctx, PLANC_EUNSUPPORTED, "hals_cols: k > 64");
}

bool rows_gram_ok(const float *H, long ldh, const float *N, long nrs, long ncs, int k) {
  const bool nrow = ncs == 1;
  const long ldn = nrow ? nrs : ncs;
  return k >= 1 && k <= 64 && (nrow || nrs == 1) && ldh % 4 == 0 && ldn % 4 == 0 &&
         (uintptr_t)H % 16 == 0 && (uintptr_t)N % 16 == 0;
}

int row_update_gram(planc_ctx *ctx, int kind, float *H, long n, int k, long ldh, const float *N, long nrs,
                    long ncs, const double *G, float eps, double *Gout, double **dparts, long *ndp) {
  if (!rows_gram_ok(H, ldh, N, nrs, ncs, k))
    return set_error(ctx, PLANC_EINVAL, "row_update_gram: layout not supported");
  const bool nrow = ncs == 1;
  const long ldn = nrow ? nrs : ncs;
  const int kp = ((k + 15) / 16) * 16;
  long nchunks = (n + GCR - 1) / GCR;
  long nb = (nchunks + RG - 1) / RG;
  if (nb > ctx->num_cu) nb = ctx->num_cu;
  if (nb < 1) nb = 1;
  double *gpart = (double *)ws_get(ctx, 1, sizeof(double) * (size_t)nb * kp * kp);
  double *dpart = dparts ? (double *)ws_get(ctx, 3, sizeof(double) * (size_t)nb) : nullptr;
  if (!gpart || (dparts && !dpart)) return set_error(ctx, PLANC_ENOMEM, "row_update_gram: workspace");
  int rc;
  if (kp == 16) rc = launch_rows_gram<16>(ctx, kind, H, n, k, ldh, N, ldn, nrow, G, eps, gpart, nb, dpart);
  else if (kp == 32) rc = launch_rows_gram<32>(ctx, kind, H, n, k, ldh, N, ldn, nrow, G, eps, gpart, nb, dpart);
  else if (kp == 48) rc = launch_rows_gram<48>(ctx, kind, H, n, k, ldh, N, ldn, nrow, G, eps, gpart, nb, dpart);
  else rc = launch_rows_gram<64>(ctx, kind, H, n, k, ldh, N, ldn, nrow, G, eps, gpart, nb, dpart);
  if (rc) return rc;
  gram_reduce_kernel<<<(unsigned)gram_reduce_blocks(kp), NTHR, 0, ctx->stream>>>(gpart, nb, k, kp, Gout);
  PLANC_LAUNCH_CHECK(ctx);
  if (dparts) {
    *dparts = dpart;
    *ndp = nb;
  }
  return PLANC_OK;
}

// norms (device, k doubles) of the columns of X.
int col_norms(planc_ctx *ctx, const float *X, long rows, int k, long ldx, double *norms) {
  long nchunks = (rows + COL_CHUNK - 1) / COL_CHUNK;
  if (nchunks == 0) nchunks = 1;
  double *part = (double *)ws_get(ctx, 2, sizeof(double) * (size_t)nchunks * k);
  if (!part) return set_error(ctx, PLANC_ENOMEM, "col_norms: workspace");
  colsumsq_partial_kernel<<<(unsigned)nchunks, NTHR, 0, ctx->stream>>>(X, rows, k, ldx, part);
  PLANC_LAUNCH_CHECK(ctx);
  colnorm_finalize_kernel<<<(k + 63) / 64, 64, 0, ctx->stream>>>(part, nchunks, k, norms);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

bool init_normalize_fits(const long *dims, int n, int k) {
  if (n < 1 || n > 8 || k < 1 || k > 64) return false;
  for (int i = 0; i < n; ++i)
    if (dims[i] < 1 || dims[i] > INIT_ROWS || dims[i] * k > INIT_VALS) return false;
  return true;
}

int init_normalize(planc_ctx *ctx, float *const *F, const long *dims, int n, int k, double *const *G,
                   double *lambda) {
  if (!init_normalize_fits(dims, n, k)) return set_error(ctx, PLANC_EINVAL, "init_normalize: shape");
  double *norms = (double *)ws_get(ctx, 38, sizeof(double) * (size_t)n * k);
  if (!norms) return set_error(ctx, PLANC_ENOMEM, "init_normalize: workspace");
  InitFacs fs{};
  for (int i = 0; i < n; ++i) {
    fs.f[i] = F[i];
    fs.g[i] = G[i];
    fs.dim[i] = (int)dims[i];
  }
  init_gram_kernel<<<dim3(n, INIT_GS), NTHR, 0, ctx->stream>>>(fs, k, norms);
  PLANC_LAUNCH_CHECK(ctx);
  init_scale_kernel<<<dim3(n, INIT_GS), NTHR, 0, ctx->stream>>>(fs, n, k, norms, lambda);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int scale_cols(planc_ctx *ctx, float *X, long rows, long ldx, float *Y, long yrows, long ldy, int k,
               const double *norms) {
  long tot = rows * k + (Y ? yrows * k : 0);
  if (tot == 0) return PLANC_OK;
  scale_cols_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, ctx->stream>>>(X, rows, ldx, Y, yrows,
                                                                            ldy, k, norms);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

// *out = sum_{r,c} X(r,c) Y(r,c) in FP64 (accumulate: *out += ...), with
// X(r,c) = X[r*xrs + c*xcs] and likewise Y.
// The partial kernel of dot(): one FP64 partial per block into workspace
// slot 3 (*part, *nparts).
int dot_parts(planc_ctx *ctx, const float *X, long xrs, long xcs, const float *Y, long yrs, long ycs,
              long rows, long cols, double **part_out, long *nparts) {
  long tot = rows * cols;
  long nb = (tot + NTHR * 8 - 1) / (NTHR * 8);
  if (nb > 2048) nb = 2048;
  if (nb < 1) nb = 1;
  double *part = (double *)ws_get(ctx, 3, sizeof(double) * (size_t)nb);
  if (!part) return set_error(ctx, PLANC_ENOMEM, "dot: workspace");
  const bool flat = xrs == 1 && yrs == 1 && (cols == 1 || (xcs == rows && ycs == rows)) &&
                    (uintptr_t)X % 16 == 0 && (uintptr_t)Y % 16 == 0;
  if (flat)
    dot_flat_kernel<<<(unsigned)nb, NTHR, 0, ctx->stream>>>(X, Y, tot, part);
  else
    dot_partial_kernel<<<(unsigned)nb, NTHR, 0, ctx->stream>>>(X, xrs, xcs, Y, yrs, ycs, rows, cols,
                                                               part);
  PLANC_LAUNCH_CHECK(ctx);
  *part_out = part;
  *nparts = nb;
  return PLANC_OK;
}

int dot(planc_ctx *ctx, const float *X, long xrs, long xcs, const float *Y, long yrs, long ycs,
        long rows, long cols, double *out, int accumulate) {
  long tot = rows * cols;
  long nb = (tot + NTHR * 8 - 1) / (NTHR * 8);
  if (nb > 2048) nb = 2048;
  if (nb < 1) nb = 1;
  double *part = (double *)ws_get(ctx, 3, sizeof(double) * (size_t)nb);
  if (!part) return set_error(ctx, PLANC_ENOMEM, "dot: workspace");
  const bool flat = xrs == 1 && yrs == 1 && (cols == 1 || (xcs == rows && ycs == rows)) &&
                    (uintptr_t)X % 16 == 0 && (uintptr_t)Y % 16 == 0;
  if (flat)
    dot_flat_kernel<<<(unsigned)nb, NTHR, 0, ctx->stream>>>(X, Y, tot, part);
  else
    dot_partial_kernel<<<(unsigned)nb, NTHR, 0, ctx->stream>>>(X, xrs, xcs, Y, yrs, ycs, rows, cols,
                                                               part);
  PLANC_LAUNCH_CHECK(ctx);
  sum_partials_kernel<<<1, NTHR, 0, ctx->stream>>>(part, nb, out, accumulate);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int nmf_relerr(planc_ctx *ctx, const double *normA2, const double *dotv, const double *WtW,
               const double *HtH, int k, double *relerr_out, double *fit_out) {
  nmf_relerr_kernel<<<1, NTHR, 0, ctx->stream>>>(normA2, dotv, WtW, HtH, k, relerr_out, fit_out, nullptr, 0);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

// The NMF error of <W, A H> = sum of the dot partials of dot_parts(): the
// partial sum and the error in one launch.
int nmf_relerr_parts(planc_ctx *ctx, const double *normA2, const double *part, long nparts, const double *WtW,
                     const double *HtH, int k, double *relerr_out) {
  nmf_relerr_kernel<<<1, NTHR, 0, ctx->stream>>>(normA2, nullptr, WtW, HtH, k, relerr_out, nullptr, part, nparts);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

// Clears / reads the column-exchange time-out flag of the HALS column sweep.
int hals_sync_reset(planc_ctx *ctx) {
  int *sync = (int *)ws_get(ctx, 17, 2 * sizeof(int));
  if (!sync) return set_error(ctx, PLANC_ENOMEM, "hals_cols: workspace");
  zero_int_kernel<<<1, 1, 0, ctx->stream>>>(sync + 1);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int hals_sync_check(planc_ctx *ctx) {
  int *sync = (int *)ws_get(ctx, 17, 2 * sizeof(int));
  if (!sync) return set_error(ctx, PLANC_ENOMEM, "hals_cols: workspace");
  int flag = 0;
  PLANC_HIP_CHECK(ctx, hipMemcpyAsync(&flag, sync + 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  PLANC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (flag) return set_error(ctx, PLANC_EHIP, "HALS column sweep: column exchange timed out");
  return PLANC_OK;
}

int gram_rescale(planc_ctx *ctx, const double *G, int k, const double *norms, double *out) {
  gram_rescale_kernel<<<(k * k + 255) / 256, 256, 0, ctx->stream>>>(G, k, norms, out);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int gram_norms(planc_ctx *ctx, const double *G, int k, double *norms) {
  gram_norms_kernel<<<(k + 63) / 64, 64, 0, ctx->stream>>>(G, k, norms);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int apply_reg(planc_ctx *ctx, double *G, int k, const double *reg) {
  if (!reg || (!(reg[0] > 0) && !(reg[1] > 0))) return PLANC_OK;
  apply_reg_kernel<<<(k * k + 255) / 256, 256, 0, ctx->stream>>>(G, k, (float)reg[0], (float)reg[1]);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int symm_reg_apply(planc_ctx *ctx, double *G, int k, float *R, long rs, long cs, const float *F,
                   long n, double sr) {
  if (!(sr > 0)) return PLANC_OK;
  add_diag_kernel<<<(k + 63) / 64, 64, 0, ctx->stream>>>(G, k, sr);
  PLANC_LAUNCH_CHECK(ctx);
  long tot = n * k;
  if (tot > 0) {
    symm_rhs_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, ctx->stream>>>(R, rs, cs, F, n, k, sr);
    PLANC_LAUNCH_CHECK(ctx);
  }
  return PLANC_OK;
}

int convert16(planc_ctx *ctx, int dtype, const float *A, long rows, long cols, long lda, void *out,
              long ldo, float *colscale) {
  if (colscale && cols > 0) {
    colmax_scale_kernel<<<(unsigned)cols, NTHR, 0, ctx->stream>>>(A, rows, lda, colscale);
    PLANC_LAUNCH_CHECK(ctx);
  }
  long tot = rows * cols;
  if (tot == 0) return PLANC_OK;
  long nb = (tot + 255) / 256;
  if (nb > 65536) nb = 65536;
  if (dtype == PLANC_BF16)
    convert_kernel<PLANC_BF16><<<(unsigned)nb, 256, 0, ctx->stream>>>(A, rows, cols, lda, (uint16_t *)out,
                                                                       ldo, colscale);
  else
    convert_kernel<PLANC_F16><<<(unsigned)nb, 256, 0, ctx->stream>>>(A, rows, cols, lda, (uint16_t *)out,
                                                                      ldo, colscale);
  PLANC_LAUNCH_CHECK(ctx);
  return PLANC_OK;
}

int transpose16(planc_ctx *ctx, const void *in, long rows, long cols, long ldi, void *out, long ldo) {
  if (rows <= 0 || cols <= 0) return PLANC_OK;
  const long gx = (rows + 63) / 64;
  if (gx > 0x7fffffffL) return set_error(ctx, PLANC_EINVAL, "transpose16: too many rows");
  const bool vec = rows % 8 == 0 && cols % 8 == 0 && ldi % 8 == 0 && ldo % 8 == 0 &&
                   (uintptr_t)in % 16 == 0 && (uintptr_t)out % 16 == 0;
  // gridDim.y is at most 65535: wider matrices go in slabs of 65535 tiles
  // of 64 columns (slab starts are multiples of 64 columns, so the 16-byte
  // alignment of the vector form holds in every slab).
  const long slab = 65535L * 64;
  for (long c0 = 0; c0 < cols; c0 += slab) {
    const long cc = cols - c0 < slab ? cols - c0 : slab;
    const long gy = (cc + 63) / 64;
    const uint16_t *src = (const uint16_t *)in + ldi * c0;
    uint16_t *dst = (uint16_t *)out + c0;
    if (vec)
      transpose16_vec_kernel<<<dim3((unsigned)gx, (unsigned)gy), 256, 0, ctx->stream>>>(src, rows, cc,
                                                                                       ldi, dst, ldo);
    else
      transpose16_kernel<<<dim3((unsigned)gx, (unsigned)gy), 256, 0, ctx->stream>>>(src, rows, cc, ldi,
                                                                                  dst, ldo);
    PLANC_LAUNCH_CHECK(ctx);
  }
  return PLANC_OK;
}

}  // namespace planc
