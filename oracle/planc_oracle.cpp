// ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// FP64 CPU restatement of PLANC's shared-memory NMF/NTF hot path.  It is the
// checker for the HIP library, never part of the product: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  Each
// function cites the reference file:line whose behaviour it restates
// (paths relative to the PLANC source tree).  Armadillo is not available, so
// a minimal column-major matrix type stands in for arma::mat; every product
// is an FP64 GEMM (internal OpenMP-blocked kernel).
//
// Parity status of each function (how it is pinned) is listed in the header
// comment of oracle/oracle.py and in tests/test_oracle.py.

#include <dlfcn.h>
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

namespace orc {

// Constants of common/utils.h:46-49.
static const double EPS_1EM16 = 0.00000000000000001;  // literally 1e-17
static const double EPS = 0.000001;
static const double EPS_1EM12 = 1e-12;

// ----------------------------------------------------------------------------
// Column-major dense matrix (stand-in for arma::mat).
struct Mat {
  long r = 0, c = 0;
  std::vector<double> d;
  Mat() {}
  Mat(long r_, long c_) : r(r_), c(c_), d((size_t)r_ * c_, 0.0) {}
  double &operator()(long i, long j) { return d[i + r * j]; }
  double operator()(long i, long j) const { return d[i + r * j]; }
  double *col(long j) { return d.data() + r * j; }
  const double *col(long j) const { return d.data() + r * j; }
};

// ----------------------------------------------------------------------------
// C = alpha*op(A)*op(B) + beta*C, column-major, op = 'N' or 'T'.
// Blocked, packed, OpenMP over output tiles; splits K across threads when
// there are fewer tiles than threads (long-K products such as KRP^T * X).
static inline double opA(const double *A, long lda, char t, long i, long k) {
  return t == 'N' ? A[i + k * lda] : A[k + i * lda];
}

static void gemm_block(char ta, char tb, long i0, long i1, long j0, long j1,
                       long k0, long k1, const double *A, long lda,
                       const double *B, long ldb, double *C, long ldc) {
  const long KB = 256;
  std::vector<double> Ap, Bp;
  for (long kb = k0; kb < k1; kb += KB) {
    long ke = std::min(k1, kb + KB), kl = ke - kb;
    long mi = i1 - i0, nj = j1 - j0;
    long mi8 = (mi + 7) / 8 * 8, nj4 = (nj + 3) / 4 * 4;
    Ap.assign((size_t)kl * mi8, 0.0);
    Bp.assign((size_t)kl * nj4, 0.0);
    // Ap: for each row group of 8, kl x 8 (row index fastest).
    for (long g = 0; g < mi8 / 8; ++g)
      for (long kk = 0; kk < kl; ++kk)
        for (long u = 0; u < 8; ++u) {
          long i = i0 + g * 8 + u;
          Ap[(size_t)g * kl * 8 + kk * 8 + u] = (i < i1) ? opA(A, lda, ta, i, kb + kk) : 0.0;
        }
    for (long g = 0; g < nj4 / 4; ++g)
      for (long kk = 0; kk < kl; ++kk)
        for (long u = 0; u < 4; ++u) {
          long j = j0 + g * 4 + u;
          // op(B)(k, j)
          double v = 0.0;
          if (j < j1) v = (tb == 'N') ? B[(kb + kk) + j * ldb] : B[j + (kb + kk) * ldb];
          Bp[(size_t)g * kl * 4 + kk * 4 + u] = v;
        }
    for (long gi = 0; gi < mi8 / 8; ++gi)
      for (long gj = 0; gj < nj4 / 4; ++gj) {
        double acc[4][8];
        for (int a = 0; a < 4; ++a)
          for (int b = 0; b < 8; ++b) acc[a][b] = 0.0;
        const double *ap = &Ap[(size_t)gi * kl * 8];
        const double *bp = &Bp[(size_t)gj * kl * 4];
        for (long kk = 0; kk < kl; ++kk) {
          for (int a = 0; a < 4; ++a) {
            double bv = bp[kk * 4 + a];
            for (int b = 0; b < 8; ++b) acc[a][b] += ap[kk * 8 + b] * bv;
          }
        }
        for (int a = 0; a < 4; ++a) {
          long j = j0 + gj * 4 + a;
          if (j >= j1) continue;
          for (int b = 0; b < 8; ++b) {
            long i = i0 + gi * 8 + b;
            if (i < i1) C[i + j * ldc] += acc[a][b];
          }
        }
      }
  }
}

// Host BLAS resolved at run time for the CPU baseline only (MKL's or
// OpenBLAS's cblas_dgemm / cblas_sgemm through dlopen; CBLAS column-major =
// 102, NoTrans = 111, Trans = 112).  Unset by default and in every test, so
// the oracle's results come from the blocked GEMM below unless a caller asks.
typedef void (*cblas_dgemm_fn)(int, int, int, int, int, int, double, const double *, int,
                               const double *, int, double, double *, int);
typedef void (*cblas_sgemm_fn)(int, int, int, int, int, int, float, const float *, int,
                               const float *, int, float, float *, int);
static cblas_dgemm_fn g_dgemm = nullptr;
static cblas_sgemm_fn g_sgemm = nullptr;

static bool fits_int(long a, long b, long c, long d, long e, long f) {
  return a < INT_MAX && b < INT_MAX && c < INT_MAX && d < INT_MAX && e < INT_MAX && f < INT_MAX;
}

void gemm(char ta, char tb, long M, long N, long K, double alpha, const double *A, long lda,
          const double *B, long ldb, double beta, double *C, long ldc) {
  if (M <= 0 || N <= 0) return;
  if (g_dgemm && K > 0 && fits_int(M, N, K, lda, ldb, ldc)) {
    g_dgemm(102, ta == 'N' ? 111 : 112, tb == 'N' ? 111 : 112, (int)M, (int)N, (int)K, alpha, A,
            (int)lda, B, (int)ldb, beta, C, (int)ldc);
    return;
  }
  for (long j = 0; j < N; ++j)
    for (long i = 0; i < M; ++i) C[i + j * ldc] = (beta == 0.0) ? 0.0 : beta * C[i + j * ldc];
  if (K <= 0 || alpha == 0.0) return;
  const long MB = 128, NB = 64;
  long ti = (M + MB - 1) / MB, tj = (N + NB - 1) / NB, nt = ti * tj;
  int nth = omp_get_max_threads();
  // Product accumulated into a scratch then scaled by alpha.
  std::vector<double> T((size_t)M * N, 0.0);
  if (nt >= nth || K < 4096 || nth == 1) {
#pragma omp parallel for schedule(dynamic) collapse(2)
    for (long bi = 0; bi < ti; ++bi)
      for (long bj = 0; bj < tj; ++bj)
        gemm_block(ta, tb, bi * MB, std::min(M, (bi + 1) * MB), bj * NB, std::min(N, (bj + 1) * NB),
                   0, K, A, lda, B, ldb, T.data(), M);
  } else {
    // Split K into nth slices; reduce partials in fixed slice order.
    int ns = nth;
    std::vector<std::vector<double>> part(ns);
#pragma omp parallel for schedule(static)
    for (int s = 0; s < ns; ++s) {
      long k0 = K * s / ns, k1 = K * (s + 1) / ns;
      part[s].assign((size_t)M * N, 0.0);
      for (long bi = 0; bi < ti; ++bi)
        for (long bj = 0; bj < tj; ++bj)
          gemm_block(ta, tb, bi * MB, std::min(M, (bi + 1) * MB), bj * NB,
                     std::min(N, (bj + 1) * NB), k0, k1, A, lda, B, ldb, part[s].data(), M);
    }
    for (int s = 0; s < ns; ++s)
      for (size_t e = 0; e < T.size(); ++e) T[e] += part[s][e];
  }
  for (long j = 0; j < N; ++j)
    for (long i = 0; i < M; ++i) C[i + j * ldc] += alpha * T[i + j * M];
}

static Mat mul(const Mat &A, char ta, const Mat &B, char tb) {
  long M = ta == 'N' ? A.r : A.c, K = ta == 'N' ? A.c : A.r;
  long N = tb == 'N' ? B.c : B.r;
  Mat C(M, N);
  gemm(ta, tb, M, N, K, 1.0, A.d.data(), A.r, B.d.data(), B.r, 0.0, C.d.data(), M);
  return C;
}

static double fro2(const Mat &A) {
  double s = 0;
  for (double v : A.d) s += v * v;
  return s;
}

// ----------------------------------------------------------------------------
// Small dense solves: arma::solve(A, B, likely_sympd) = Cholesky, falling back
// to LU with partial pivoting when A is not positive definite.
static bool chol_inplace(std::vector<double> &L, int p) {
  for (int j = 0; j < p; ++j) {
    double s = L[j + p * j];
    for (int t = 0; t < j; ++t) s -= L[j + p * t] * L[j + p * t];
    if (!(s > 0.0)) return false;
    double ljj = std::sqrt(s);
    L[j + p * j] = ljj;
    for (int i = j + 1; i < p; ++i) {
      double v = L[i + p * j];
      for (int t = 0; t < j; ++t) v -= L[i + p * t] * L[j + p * t];
      L[i + p * j] = v / ljj;
    }
  }
  return true;
}

// Solves S X = B for X (p x nrhs, column-major, in B). S is p x p.
static void solve_sympd(const std::vector<double> &S, int p, std::vector<double> &B, long nrhs) {
  if (p == 0) return;
  std::vector<double> L = S;
  if (chol_inplace(L, p)) {
    for (long c = 0; c < nrhs; ++c) {
      double *b = &B[(size_t)c * p];
      for (int i = 0; i < p; ++i) {
        double v = b[i];
        for (int t = 0; t < i; ++t) v -= L[i + p * t] * b[t];
        b[i] = v / L[i + p * i];
      }
      for (int i = p - 1; i >= 0; --i) {
        double v = b[i];
        for (int t = i + 1; t < p; ++t) v -= L[t + p * i] * b[t];
        b[i] = v / L[i + p * i];
      }
    }
    return;
  }
  // LU with partial pivoting.
  std::vector<double> LU = S;
  std::vector<int> piv(p);
  for (int j = 0; j < p; ++j) {
    int pr = j;
    for (int i = j + 1; i < p; ++i)
      if (std::fabs(LU[i + p * j]) > std::fabs(LU[pr + p * j])) pr = i;
    piv[j] = pr;
    if (pr != j)
      for (int t = 0; t < p; ++t) std::swap(LU[j + p * t], LU[pr + p * t]);
    double d = LU[j + p * j];
    if (d == 0.0) continue;
    for (int i = j + 1; i < p; ++i) {
      LU[i + p * j] /= d;
      for (int t = j + 1; t < p; ++t) LU[i + p * t] -= LU[i + p * j] * LU[j + p * t];
    }
  }
  for (long c = 0; c < nrhs; ++c) {
    double *b = &B[(size_t)c * p];
    for (int j = 0; j < p; ++j)
      if (piv[j] != j) std::swap(b[j], b[piv[j]]);
    for (int i = 0; i < p; ++i)
      for (int t = 0; t < i; ++t) b[i] -= LU[i + p * t] * b[t];
    for (int i = p - 1; i >= 0; --i) {
      for (int t = i + 1; t < p; ++t) b[i] -= LU[i + p * t] * b[t];
      b[i] /= LU[i + p * i];
    }
  }
}

// One ADMM sweep of AOADMMNMF::computeNMF (nmf/aoadmm.hpp:111-141 for H,
// :157-185 for W) and NTFAOADMM::update (ntf/ntfaoadmm.hpp:34-72), shared:
// G (k x k, Gram + regularisation, without the rho shift) is shifted by
// rho = trace(G)/k (0.01 if not positive) and Cholesky-factored once; then
// at most 5 sweeps, each
//   T = (G + rho I)^-1 (R + rho (X + D)^T)    (R is k x n, X and D n x k)
//   X = max(T^T - D, 0);  D = D + X - T^T
// stopping when ||X - T^T|| < 0.01 ||X|| and ||X - X_prev|| < 0.01 ||D||.
// Returns false when G + rho I is not positive definite.
static bool admm_solve(const Mat &G0, const Mat &R, Mat &X, Mat &D) {
  const long k = G0.r, n = X.r;
  double tr = 0;
  for (long c = 0; c < k; ++c) tr += G0(c, c);
  double rho = tr / k;
  rho = rho > 0 ? rho : 0.01;
  std::vector<double> L(G0.d);
  for (long c = 0; c < k; ++c) L[c + k * c] += rho;
  if (!chol_inplace(L, (int)k)) return false;
  const double tol = 0.01;
  std::vector<double> T((size_t)k * n);
  Mat Xprev;
  for (int it = 0; it < 5; ++it) {
    Xprev = X;
    for (long j = 0; j < n; ++j) {
      double *b = &T[(size_t)k * j];
      for (long c = 0; c < k; ++c) b[c] = R(c, j) + rho * (X(j, c) + D(j, c));
      for (long i = 0; i < k; ++i) {  // forward: L y = b
        double v = b[i];
        for (long t = 0; t < i; ++t) v -= L[i + k * t] * b[t];
        b[i] = v / L[i + k * i];
      }
      for (long i = k - 1; i >= 0; --i) {  // backward: L^T x = y
        double v = b[i];
        for (long t = i + 1; t < k; ++t) v -= L[t + k * i] * b[t];
        b[i] = v / L[i + k * i];
      }
    }
    double r2 = 0, s2 = 0, x2 = 0, d2 = 0;
    for (long j = 0; j < n; ++j)
      for (long c = 0; c < k; ++c) {
        const double t = T[c + (size_t)k * j];
        double x = t - D(j, c);
        x = x > 0.0 ? x : 0.0;
        X(j, c) = x;
        D(j, c) = D(j, c) + x - t;
        r2 += (x - t) * (x - t);
        s2 += (x - Xprev(j, c)) * (x - Xprev(j, c));
        x2 += x * x;
        d2 += D(j, c) * D(j, c);
      }
    if (std::sqrt(r2) < tol * std::sqrt(x2) && std::sqrt(s2) < tol * std::sqrt(d2)) break;
  }
  return true;
}

// ----------------------------------------------------------------------------
// Block principal pivoting NNLS, multiple right-hand sides.
// Restates BPPNNLS::solveNNLSMultipleRHS (nnls/bppnnls.hpp:67-191) with
// solveNormalEqComb (:204-259) and the passive-set grouping of
// computeCorrelationScore / SortBooleanMatrix (:272-287,
// nnls/SortBooleanMatrix.hpp:10-45).  AtA is n x n, AtB n x nrhs, X out.
// Returns the iteration count, or -1 when MAX_ITERATIONS = 5n is exceeded
// (the reference calls exit(EXIT_FAILURE) there, :185-188).
static void solve_normal_eq_comb(const double *AtA, int n, const double *AtB, long ldb,
                                 const std::vector<long> &cols, const std::vector<uint8_t> &Pset,
                                 double *X, long ldx) {
  // Group columns with identical passive sets: lexicographic sort of the
  // boolean columns, as SortBooleanMatrix does.
  std::vector<long> idx(cols.begin(), cols.end());
  auto less = [&](long a, long b) {
    for (int i = 0; i < n; ++i) {
      uint8_t x = Pset[(size_t)a * n + i], y = Pset[(size_t)b * n + i];
      if (x < y) return true;
      if (x > y) return false;
    }
    return false;
  };
  std::sort(idx.begin(), idx.end(), less);
  size_t g0 = 0;
  while (g0 < idx.size()) {
    size_t g1 = g0 + 1;
    while (g1 < idx.size() && !less(idx[g0], idx[g1]) && !less(idx[g1], idx[g0])) ++g1;
    std::vector<int> P;
    for (int i = 0; i < n; ++i)
      if (Pset[(size_t)idx[g0] * n + i]) P.push_back(i);
    int p = (int)P.size();
    long nr = (long)(g1 - g0);
    std::vector<double> S((size_t)p * p), B((size_t)p * nr);
    for (int a = 0; a < p; ++a)
      for (int b = 0; b < p; ++b) S[a + p * b] = AtA[P[a] + (size_t)n * P[b]];
    for (long c = 0; c < nr; ++c)
      for (int a = 0; a < p; ++a) B[a + (size_t)p * c] = AtB[P[a] + ldb * idx[g0 + c]];
    solve_sympd(S, p, B, nr);
    for (long c = 0; c < nr; ++c) {
      double *x = X + ldx * idx[g0 + c];
      for (int i = 0; i < n; ++i) x[i] = 0.0;
      for (int a = 0; a < p; ++a) x[P[a]] = B[a + (size_t)p * c];
    }
    g0 = g1;
  }
}

int bppnnls(const double *AtA, int n, const double *AtB, long nrhs, double *X) {
  const long MAXIT = (long)n * 5;  // nnls/bppnnls.hpp:69
  std::vector<double> Y((size_t)n * nrhs);
  std::vector<uint8_t> Pset((size_t)n * nrhs, 0), NonOpt((size_t)n * nrhs), Infea((size_t)n * nrhs);
  std::vector<int> P(nrhs, 3), Ninf(nrhs, n + 1), NotGood(nrhs, 0);
  std::vector<uint8_t> NotOpt(nrhs, 0);
  for (long c = 0; c < nrhs; ++c)
    for (int i = 0; i < n; ++i) {
      X[i + (size_t)n * c] = 0.0;
      Y[i + (size_t)n * c] = -AtB[i + (size_t)n * c];  // AtA*0 - AtB (:73)
    }
  auto recompute = [&](long c) {
    int cnt = 0;
    for (int i = 0; i < n; ++i) {
      size_t e = i + (size_t)n * c;
      NonOpt[e] = (Y[e] < 0) && (Pset[e] == 0);
      Infea[e] = (X[e] < 0) && (Pset[e] != 0);
      cnt += NonOpt[e] + Infea[e];
    }
    NotGood[c] = cnt;
    NotOpt[c] = cnt > 0;
  };
  long numNonOpt = 0;
  for (long c = 0; c < nrhs; ++c) {
    recompute(c);
    numNonOpt += NotOpt[c];
  }
  long iter = 0;
  std::vector<long> nonopt_cols;
  while (numNonOpt > 0) {
    ++iter;
    if (MAXIT > 0 && iter > MAXIT) return -1;
    for (long c = 0; c < nrhs; ++c) {
      if (!NotOpt[c]) continue;
      bool c1 = NotGood[c] < Ninf[c];
      bool c2 = !c1 && NotGood[c] >= Ninf[c] && P[c] >= 1;
      if (c1 || c2) {
        if (c1) {
          P[c] = 3;
          Ninf[c] = NotGood[c];
        } else {
          P[c] -= 1;
        }
        for (int i = 0; i < n; ++i) {
          size_t e = i + (size_t)n * c;
          if (NonOpt[e]) Pset[e] = 1;
          if (Infea[e]) Pset[e] = 0;
        }
      } else {
        // Backup rule: flip the largest index that is non-optimal or infeasible.
        int rowidx = -1;
        for (int i = 0; i < n; ++i)
          if (NonOpt[i + (size_t)n * c] + Infea[i + (size_t)n * c]) rowidx = i;
        size_t e = rowidx + (size_t)n * c;
        Pset[e] = Pset[e] ? 0 : 1;
      }
    }
    nonopt_cols.clear();
    for (long c = 0; c < nrhs; ++c)
      if (NotOpt[c]) nonopt_cols.push_back(c);
    solve_normal_eq_comb(AtA, n, AtB, n, nonopt_cols, Pset, X, n);
    for (long c : nonopt_cols) {
      for (int i = 0; i < n; ++i) {
        double v = -AtB[i + (size_t)n * c];
        for (int t = 0; t < n; ++t) v += AtA[i + (size_t)n * t] * X[t + (size_t)n * c];
        Y[i + (size_t)n * c] = v;
      }
    }
    // fixAbsNumericalError on X and Y (:170-172); all columns.
    for (size_t e = 0; e < (size_t)n * nrhs; ++e) {
      if (std::fabs(X[e]) < EPS_1EM12) X[e] = 0.0;
      if (std::fabs(Y[e]) < EPS_1EM12) Y[e] = 0.0;
    }
    numNonOpt = 0;
    for (long c = 0; c < nrhs; ++c) {
      recompute(c);
      numNonOpt += NotOpt[c];
    }
  }
  return (int)iter;
}

// ----------------------------------------------------------------------------
// Shared-memory NMF (common/nmf.hpp, nmf/mu.hpp, nmf/hals.hpp, nmf/bppnmf.hpp).
struct NMFState {
  const Mat *A;
  Mat W, H;
  long m, n, k;
  double normA;
  double regW[2], regH[2];
  double symm_reg = -1.0;
  double fit_err_sq = 0, objective = 0;
};

// applyReg, common/nmf.hpp:76-90 (the reference holds reg in an FVEC, so the
// value is rounded to float).
static void apply_reg(const double *reg, Mat &G) {
  if (reg[0] > 0) {
    float l2 = (float)reg[0];
    for (long i = 0; i < G.r; ++i) G(i, i) += l2;
  }
  if (reg[1] > 0) {
    float l1 = (float)reg[1];
    for (double &v : G.d) v += l1;
  }
}

// normalize_by_W, common/nmf.hpp:152-161.
static void normalize_by_W(NMFState &s) {
  for (long c = 0; c < s.k; ++c) {
    double nrm = 0;
    for (long i = 0; i < s.m; ++i) nrm += s.W(i, c) * s.W(i, c);
    nrm = std::sqrt(nrm);
    if (nrm > 0) {
      for (long i = 0; i < s.m; ++i) s.W(i, c) /= nrm;
      for (long j = 0; j < s.n; ++j) s.H(j, c) *= nrm;
    }
  }
}

// computeObjectiveError, common/nmf.hpp:390-424 (recomputes A^T W).
static void objective_error(NMFState &s) {
  Mat AtW = mul(*s.A, 'T', s.W, 'N');
  Mat WtW = mul(s.W, 'T', s.W, 'N');
  Mat HtH = mul(s.H, 'T', s.H, 'N');
  double TrHtAtW = 0;
  for (long j = 0; j < s.n; ++j)
    for (long c = 0; c < s.k; ++c) TrHtAtW += s.H(j, c) * AtW(j, c);
  double TrWtWHtH = 0;
  for (long a = 0; a < s.k; ++a)
    for (long b = 0; b < s.k; ++b) TrWtWHtH += WtW(a, b) * HtH(b, a);
  double froW = 0, froH = 0;
  for (long c = 0; c < s.k; ++c) {
    froW += WtW(c, c);
    froH += HtH(c, c);
  }
  double l1W = 0, l1H = 0;
  for (long i = 0; i < s.m; ++i) {
    double r = 0;
    for (long c = 0; c < s.k; ++c) r += s.W(i, c);
    l1W += r * r;
  }
  for (long j = 0; j < s.n; ++j) {
    double r = 0;
    for (long c = 0; c < s.k; ++c) r += s.H(j, c);
    l1H += r * r;
  }
  s.fit_err_sq = s.normA * s.normA - 2 * TrHtAtW + TrWtWHtH;
  s.objective = s.fit_err_sq + s.regW[0] * froW + s.regH[0] * froH + s.regW[1] * l1W +
                s.regH[1] * l1H;
}

// printObjective's relative error, common/nmf.hpp:462,468.
static double rel_error(const NMFState &s) {
  double err = (s.fit_err_sq > 0) ? std::sqrt(s.fit_err_sq) : s.normA;
  return err / s.normA;
}

// HALSNMF::computeNMF, nmf/hals.hpp:50-116 (the constructor's normalize_by_W,
// :41,46, is applied by the caller before the loop).
static void hals_loop(NMFState &s, int iters, double *relerr) {
  std::vector<double> x;
  for (int it = 0; it < iters; ++it) {
    Mat WtA = mul(s.W, 'T', *s.A, 'N');  // k x n
    Mat WtW = mul(s.W, 'T', s.W, 'N');
    apply_reg(s.regH, WtW);
    x.resize(s.n);
    for (long c = 0; c < s.k; ++c) {
      for (long j = 0; j < s.n; ++j) {
        double hv = 0;
        for (long t = 0; t < s.k; ++t) hv += s.H(j, t) * WtW(t, c);
        double v = s.H(j, c) + (WtA(c, j) - hv);
        x[j] = (v < EPS_1EM16) ? EPS_1EM16 : v;
      }
      double nrm = 0;
      for (long j = 0; j < s.n; ++j) nrm += x[j] * x[j];
      if (std::sqrt(nrm) != 0)
        for (long j = 0; j < s.n; ++j) s.H(j, c) = x[j];
    }
    Mat AH = mul(*s.A, 'N', s.H, 'N');  // m x k
    Mat HtH = mul(s.H, 'T', s.H, 'N');
    apply_reg(s.regW, HtH);
    x.resize(s.m);
    for (long c = 0; c < s.k; ++c) {
      for (long i = 0; i < s.m; ++i) {
        double wv = 0;
        for (long t = 0; t < s.k; ++t) wv += s.W(i, t) * HtH(t, c);
        double v = (s.W(i, c) * HtH(c, c)) + (AH(i, c) - wv);
        x[i] = (v < EPS_1EM16) ? EPS_1EM16 : v;
      }
      double nrm = 0;
      for (long i = 0; i < s.m; ++i) nrm += x[i] * x[i];
      nrm = std::sqrt(nrm);
      if (nrm != 0)
        for (long i = 0; i < s.m; ++i) s.W(i, c) = x[i] / nrm;
    }
    normalize_by_W(s);
    objective_error(s);
    if (relerr) relerr[it] = rel_error(s);
  }
}

// BPPNMF::updateOtherGivenOneMultipleRHS, nmf/bppnmf.hpp:27-94, with the
// input passed untransposed: for 'W' the right-hand side Given^T*A^T is
// formed as (A*H)^T.
static int bpp_update(const NMFState &s, bool forW, const Mat &given, const double *reg, Mat &other) {
  const long CH = 2000;  // ONE_THREAD_MATRIX_SIZE, :16
  Mat GtG = mul(given, 'T', given, 'N');
  apply_reg(reg, GtG);
  Mat GtIn;  // k x ncols
  if (forW) {
    Mat AH = mul(*s.A, 'N', given, 'N');  // m x k
    GtIn = Mat(s.k, s.m);
    for (long i = 0; i < s.m; ++i)
      for (long c = 0; c < s.k; ++c) GtIn(c, i) = AH(i, c);
  } else {
    GtIn = mul(given, 'T', *s.A, 'N');  // k x n
  }
  if (s.symm_reg > 0) {  // applySymmetricReg, common/nmf.hpp:123-129
    for (long c = 0; c < s.k; ++c) GtG(c, c) += s.symm_reg;
    for (long j = 0; j < GtIn.c; ++j)
      for (long c = 0; c < s.k; ++c) GtIn(c, j) += s.symm_reg * given(j, c);
  }
  long ncols = GtIn.c;
  long nch = (ncols + CH - 1) / CH;
  int fail = 0;
  std::vector<int> rc(nch, 0);
#pragma omp parallel for schedule(dynamic)
  for (long ch = 0; ch < nch; ++ch) {
    long c0 = ch * CH, c1 = std::min(ncols, c0 + CH);
    std::vector<double> X((size_t)s.k * (c1 - c0));
    rc[ch] = bppnnls(GtG.d.data(), (int)s.k, GtIn.col(c0), c1 - c0, X.data());
    for (long j = c0; j < c1; ++j)
      for (long c = 0; c < s.k; ++c) other(j, c) = X[c + (size_t)s.k * (j - c0)];
  }
  for (long ch = 0; ch < nch; ++ch)
    if (rc[ch] < 0) fail = 1;
  return fail ? -1 : 0;
}

// ----------------------------------------------------------------------------
// Dense tensor, mode 0 fastest (common/tensor.hpp:23-28).
struct Tens {
  int N;
  std::vector<long> dims;
  const double *x;
  long numel() const {
    long p = 1;
    for (long d : dims) p *= d;
    return p;
  }
};

// krp_leave_out_one, common/ncpfactors.hpp:171-233: rows ordered with the
// lowest remaining mode fastest.  Out is (prod others) x k.
static void krp_leave_out_one(const std::vector<Mat> &F, int n, long k, Mat &out) {
  int N = (int)F.size();
  std::vector<int> modes;
  for (int i = 0; i < N; ++i)
    if (i != n) modes.push_back(i);
  long rows = 1;
  for (int i : modes) rows *= F[i].r;
  out = Mat(rows, k);
#pragma omp parallel for schedule(static)
  for (long r = 0; r < rows; ++r) {
    long rem = r;
    std::vector<long> dig(modes.size());
    for (size_t t = 0; t < modes.size(); ++t) {
      dig[t] = rem % F[modes[t]].r;
      rem /= F[modes[t]].r;
    }
    for (long c = 0; c < k; ++c) {
      double v = 1.0;
      // The reference multiplies from the highest mode down (matorder is reversed).
      for (int t = (int)modes.size() - 1; t >= 0; --t) v *= F[modes[t]](dig[t], c);
      out(r, c) = v;
    }
  }
}

// Tensor::mttkrp, common/tensor.hpp:246-327.  Out is k x I_n.
static void mttkrp_direct(const Tens &X, const Mat &krp, int n, long k, Mat &out) {
  long In = X.dims[n];
  out = Mat(k, In);
  if (n == 0) {
    // X_(0) (I0 x rest, column-major) * KRP  -> I0 x k, stored transposed.
    Mat T(In, k);
    long rest = X.numel() / In;
    gemm('N', 'N', In, k, rest, 1.0, X.x, In, krp.d.data(), krp.r, 0.0, T.d.data(), In);
    for (long i = 0; i < In; ++i)
      for (long c = 0; c < k; ++c) out(c, i) = T(i, c);
    return;
  }
  long ncols = 1, nmats = 1;
  for (int i = 0; i < n; ++i) ncols *= X.dims[i];
  for (int i = n + 1; i < X.N; ++i) nmats *= X.dims[i];
  // For each slab p: out(c, i_n) += sum_q X[q + ncols*(i_n + In*p)] * KRP(q + ncols*p, c).
  Mat T(In, k);
  for (long p = 0; p < nmats; ++p) {
    // slab as (ncols x In) column-major: element (q, i_n) at q + ncols*i_n.
    gemm('T', 'N', In, k, ncols, 1.0, X.x + (size_t)p * ncols * In, ncols,
         krp.d.data() + p * ncols, krp.r, 1.0, T.d.data(), In);
  }
  for (long i = 0; i < In; ++i)
    for (long c = 0; c < k; ++c) out(c, i) = T(i, c);
}

// Khatri-Rao product of the factors of modes [a, b) (lowest mode fastest).
static Mat krp_range(const std::vector<Mat> &F, int a, int b, long k) {
  long rows = 1;
  for (int i = a; i < b; ++i) rows *= F[i].r;
  Mat out(rows, k);
#pragma omp parallel for schedule(static)
  for (long r = 0; r < rows; ++r) {
    long rem = r;
    std::vector<long> dig(b - a);
    for (int t = a; t < b; ++t) {
      dig[t - a] = rem % F[t].r;
      rem /= F[t].r;
    }
    for (long c = 0; c < k; ++c) {
      double v = 1.0;
      for (int t = b - 1; t >= a; --t) v *= F[t](dig[t - a], c);
      out(r, c) = v;
    }
  }
  return out;
}

// Contract a projection over all modes of [a, b) except n, with the factors.
// Pr is (prod dims[a..b)) x k column-major (proj index: lowest mode fastest),
// or k x (prod) when kfirst.  Out k x I_n.
static void contract_projection(const Mat &Pr, bool kfirst, const std::vector<Mat> &F, int a,
                                int b, int n, long k, Mat &out) {
  long In = F[n].r;
  long total = 1;
  for (int i = a; i < b; ++i) total *= F[i].r;
  out = Mat(k, In);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < k; ++c) {
    std::vector<long> dig(b - a);
    for (long e = 0; e < total; ++e) {
      long rem = e;
      for (int t = a; t < b; ++t) {
        dig[t - a] = rem % F[t].r;
        rem /= F[t].r;
      }
      double w = kfirst ? Pr(c, e) : Pr(e, c);
      for (int t = a; t < b; ++t)
        if (t != n) w *= F[t](dig[t - a], c);
      out(c, dig[n - a]) += w;
    }
  }
}

// Dimension-tree MTTKRP (dimtree/ddt.hpp:134-243 with the shared-memory split
// s = N/2 of ntf/auntf.hpp:123): the right contraction X_(L x R)*KRP_R is formed
// at n = 0 and reused for n <= s; the left contraction KRP_L^T*X_(L x R) at
// n = s+1 and reused for n > s+1.  Mathematically equal to mttkrp_direct.
struct DimTree {
  int s;
  Mat P;  // L x k
  Mat Q;  // k x R
};

static void mttkrp_dimtree(const Tens &X, const std::vector<Mat> &F, int n, long k, DimTree &dt,
                           Mat &out) {
  int N = X.N, s = dt.s;
  long L = 1, R = 1;
  for (int i = 0; i <= s; ++i) L *= X.dims[i];
  for (int i = s + 1; i < N; ++i) R *= X.dims[i];
  if (n == 0) {
    Mat KR = krp_range(F, s + 1, N, k);
    dt.P = Mat(L, k);
    gemm('N', 'N', L, k, R, 1.0, X.x, L, KR.d.data(), R, 0.0, dt.P.d.data(), L);
  }
  if (n == s + 1) {
    Mat KL = krp_range(F, 0, s + 1, k);
    dt.Q = Mat(k, R);
    gemm('T', 'N', k, R, L, 1.0, KL.d.data(), L, X.x, L, 0.0, dt.Q.d.data(), k);
  }
  if (n <= s)
    contract_projection(dt.P, false, F, 0, s + 1, n, k, out);
  else
    contract_projection(dt.Q, true, F, s + 1, N, n, k, out);
}

// ----------------------------------------------------------------------------
// AUNTF (ntf/auntf.hpp) with the NCPFactors operations it uses.
struct NTFState {
  Tens X;
  std::vector<Mat> F;  // I_i x k, unit-norm columns after normalisation
  std::vector<double> lambda;
  long k;
  double normX2;
  bool dimtree;
  DimTree dt;
};

// NCPFactors::normalize(), common/ncpfactors.hpp:329-339.
static void ncp_normalize_all(NTFState &s) {
  std::fill(s.lambda.begin(), s.lambda.end(), 1.0);
  for (size_t i = 0; i < s.F.size(); ++i)
    for (long c = 0; c < s.k; ++c) {
      double nrm = 0;
      for (long r = 0; r < s.F[i].r; ++r) nrm += s.F[i](r, c) * s.F[i](r, c);
      nrm = std::sqrt(nrm);
      if (nrm > 0)
        for (long r = 0; r < s.F[i].r; ++r) s.F[i](r, c) /= nrm;
      s.lambda[c] *= nrm;
    }
}

// NCPFactors::normalize(mode), common/ncpfactors.hpp:346-351: lambda replaced.
static void ncp_normalize_mode(NTFState &s, int mode) {
  Mat &A = s.F[mode];
  for (long c = 0; c < s.k; ++c) {
    double nrm = 0;
    for (long r = 0; r < A.r; ++r) nrm += A(r, c) * A(r, c);
    nrm = std::sqrt(nrm);
    s.lambda[c] = nrm;
    if (nrm > 0)
      for (long r = 0; r < A.r; ++r) A(r, c) /= nrm;
  }
}

// gram_leave_out_one, common/ncpfactors.hpp:139-148.
static Mat gram_leave_out_one(const NTFState &s, int n) {
  Mat S(s.k, s.k);
  std::fill(S.d.begin(), S.d.end(), 1.0);
  for (size_t i = 0; i < s.F.size(); ++i) {
    if ((int)i == n) continue;
    Mat G = mul(s.F[i], 'T', s.F[i], 'N');
    for (size_t e = 0; e < S.d.size(); ++e) S.d[e] *= G.d[e];
  }
  return S;
}

// AUNTF::computeObjectiveError, ntf/auntf.hpp:200-246.
static double ntf_rel_error(const NTFState &s) {
  Mat krp;
  krp_leave_out_one(s.F, 0, s.k, krp);
  long I0 = s.X.dims[0], rest = krp.r;
  Mat U = s.F[0];
  for (long c = 0; c < s.k; ++c)
    for (long i = 0; i < I0; ++i) U(i, c) *= s.lambda[c];
  std::vector<double> low((size_t)I0 * rest);
  gemm('N', 'T', I0, rest, s.k, 1.0, U.d.data(), I0, krp.d.data(), rest, 0.0, low.data(), I0);
  double err = 0;
  long ne = s.X.numel();
#pragma omp parallel for reduction(+ : err) schedule(static)
  for (long e = 0; e < ne; ++e) {
    double d = s.X.x[e] - low[e];
    err += d * d;
  }
  return std::sqrt(err / s.normX2);
}

// NTF update rules: NTFHALS::update (ntf/ntfhals.hpp:13-29), NTFMU::update
// (ntf/ntfmu.hpp:11-16), NTFANLSBPP::update (ntf/ntfanlsbpp.hpp:15-57).
// M is k x I_n, S k x k; the factor is updated in place (I_n x k).
static int ntf_update(int algo, Mat &H, const Mat &M, const Mat &S, long k, Mat *dual = nullptr) {
  long In = H.r;
  if (algo == 4) {  // AO-ADMM (ntf/ntfaoadmm.hpp:23-77): dual of this mode in *dual
    if (!dual || !admm_solve(S, M, H, *dual)) return -1;
    // m_ncp_aux.normalize(mode): the dual's columns to unit norm (:73).
    for (long c = 0; c < k; ++c) {
      double nrm = 0;
      for (long r = 0; r < In; ++r) nrm += (*dual)(r, c) * (*dual)(r, c);
      nrm = std::sqrt(nrm);
      if (nrm > 0)
        for (long r = 0; r < In; ++r) (*dual)(r, c) /= nrm;
    }
    return 0;
  }
  if (algo == 1) {  // HALS
    std::vector<double> u(In);
    for (long i = 0; i < k; ++i) {
      double nrm = 0;
      for (long r = 0; r < In; ++r) {
        double hs = 0;
        for (long t = 0; t < k; ++t) hs += H(r, t) * S(t, i);
        double v = H(r, i) + (M(i, r) - hs);
        v = (v < EPS_1EM16) ? EPS_1EM16 : v;
        u[r] = v;
        nrm += v * v;
      }
      if (nrm > 0)
        for (long r = 0; r < In; ++r) H(r, i) = u[r];
    }
  } else if (algo == 0) {  // MU
    Mat T = mul(H, 'N', S, 'N');
    for (long r = 0; r < In; ++r)
      for (long c = 0; c < k; ++c) H(r, c) = (H(r, c) * M(c, r)) / (T(r, c) + EPS);
  } else if (algo == 2) {  // ANLS-BPP, chunks of 2000 right-hand sides
    const long CH = 2000;
    long nch = (In + CH - 1) / CH;
    std::vector<int> rc(nch, 0);
#pragma omp parallel for schedule(dynamic)
    for (long ch = 0; ch < nch; ++ch) {
      long c0 = ch * CH, c1 = std::min(In, c0 + CH);
      std::vector<double> X((size_t)k * (c1 - c0));
      rc[ch] = bppnnls(S.d.data(), (int)k, M.col(c0), c1 - c0, X.data());
      for (long r = c0; r < c1; ++r)
        for (long c = 0; c < k; ++c) H(r, c) = X[c + (size_t)k * (r - c0)];
    }
    for (long ch = 0; ch < nch; ++ch)
      if (rc[ch] < 0) return -1;
  } else {
    return -2;
  }
  return 0;
}

// ---- NTFNES (ntf/ntfnes.hpp) ------------------------------------------------
// Eigenvalues of a symmetric k x k matrix by cyclic Jacobi rotations
// (arma::eig_sym in the reference, ntf/ntfnes.hpp:130): the smallest and the
// largest.
static void sym_eig_minmax(const Mat &S, double &mn, double &mx) {
  const long k = S.r;
  std::vector<double> a(S.d);
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0, diag = 0;
    for (long p = 0; p < k; ++p)
      for (long q = 0; q < k; ++q) (p == q ? diag : off) += a[p + k * q] * a[p + k * q];
    if (off <= 1e-30 * diag || off == 0) break;
    for (long p = 0; p < k - 1; ++p)
      for (long q = p + 1; q < k; ++q) {
        const double apq = a[p + k * q];
        if (apq == 0) continue;
        const double app = a[p + k * p], aqq = a[q + k * q];
        const double theta = (aqq - app) / (2 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        const double c = 1 / std::sqrt(t * t + 1), sn = t * c;
        for (long r = 0; r < k; ++r) {  // columns p, q
          const double arp = a[r + k * p], arq = a[r + k * q];
          a[r + k * p] = c * arp - sn * arq;
          a[r + k * q] = sn * arp + c * arq;
        }
        for (long r = 0; r < k; ++r) {  // rows p, q
          const double apr = a[p + k * r], aqr = a[q + k * r];
          a[p + k * r] = c * apr - sn * aqr;
          a[q + k * r] = sn * apr + c * aqr;
        }
      }
  }
  mn = mx = a[0];
  for (long p = 1; p < k; ++p) {
    mn = std::min(mn, a[p + k * p]);
    mx = std::max(mx, a[p + k * p]);
  }
}

// The factors of NTFNES in its transposed NCPFactors (k x I_n), held here as
// I_n x k; lambda per NCPFactors.
struct NESState {
  std::vector<Mat> prox, prev;        // m_prox_t, m_prev_t (zero at construction, :175-176)
  std::vector<double> prox_l, prev_l;  // their lambdas (ones at construction)
  double delta1 = 1e-2, delta2 = 1e-2;
  int acc_start = 5, acc_exp = 3, acc_fails = 0, fail_limit = 5;  // :180-186
};

// normalize_rows(mode) of a transposed factor (common/ncpfactors.hpp:358-363).
static void normalize_rows_t(Mat &F, std::vector<double> &lambda) {
  for (long c = 0; c < F.c; ++c) {
    double nrm = 0;
    for (long r = 0; r < F.r; ++r) nrm += F(r, c) * F(r, c);
    nrm = std::sqrt(nrm);
    lambda[c] = nrm;
    if (nrm > 0)
      for (long r = 0; r < F.r; ++r) F(r, c) /= nrm;
  }
}

// NTFNES::update (ntf/ntfnes.hpp:119-166); the new factor (unnormalised)
// into H.  Returns -4 when the inner loop runs past 100000 steps (the
// reference has no bound).
static int nes_update(NESState &ns, int mode, const Mat &S, const Mat &M, long k, Mat &H) {
  if (mode == 0) ns.prev_l = ns.prox_l;
  ns.prev[mode] = ns.prox[mode];
  Mat Ht = ns.prox[mode], Htprev = Ht, acc = Ht;
  const long I = Ht.r;
  double L, mu;
  sym_eig_minmax(S, mu, L);
  const double q0 = L / mu;
  const double lambda = q0 > 1e6 ? 10 * mu : (q0 > 1e3 ? mu : mu / 10);  // get_lambda, :27-38
  const double q = (mu + lambda) / (L + lambda);
  Mat G = S;
  for (long c = 0; c < k; ++c) G(c, c) += lambda;
  Mat mod(I, k);
  for (long r = 0; r < I; ++r)
    for (long c = 0; c < k; ++c) mod(r, c) = -lambda * acc(r, c) - M(c, r);
  double alpha = 1, alpha_prev = 1, beta = 1;
  Mat grad(I, k);
  for (long it = 0;; ++it) {
    double absmax = 0, gmin = 0;
    bool first = true;
    for (long r = 0; r < I; ++r)
      for (long c = 0; c < k; ++c) {
        double g = mod(r, c);
        for (long t = 0; t < k; ++t) g += G(c, t) * acc(r, t);
        grad(r, c) = g;
        absmax = std::max(absmax, std::fabs(g * acc(r, c)));
        gmin = first ? g : std::min(gmin, g);
        first = false;
      }
    if (absmax <= ns.delta1 && gmin >= -ns.delta2) break;  // stop_iter, :56-66
    if (it >= 100000) return -4;
    Htprev = Ht;
    for (long e = 0; e < (long)Ht.d.size(); ++e) {
      const double v = acc.d[e] - grad.d[e] / (L + lambda);
      Ht.d[e] = v > 0.0 ? v : 0.0;
    }
    alpha_prev = alpha;
    {  // get_alpha, :40-54: root of x^2 + (alpha^2 - q) x - alpha^2 = 0
      const double b = alpha * alpha - q, c2 = -alpha * alpha;
      alpha = (-b + std::sqrt(b * b - 4 * c2)) / 2;
    }
    beta = (alpha_prev * (1 - alpha_prev)) / (alpha + alpha_prev * alpha_prev);
    for (long e = 0; e < (long)Ht.d.size(); ++e) acc.d[e] = Ht.d[e] + beta * (Ht.d[e] - Htprev.d[e]);
  }
  ns.prox[mode] = Ht;
  normalize_rows_t(ns.prox[mode], ns.prox_l);
  H = Ht;
  return 0;
}

// reset(new_factors, trans) (ntf/auntf.hpp:186-198): every mode set and
// normalised (normalize(mode) replacing lambda), then lambda taken over.
static void ntf_reset(NTFState &s, const std::vector<Mat> &F, const std::vector<double> &lambda) {
  for (size_t j = 0; j < F.size(); ++j) {
    s.F[j] = F[j];
    ncp_normalize_mode(s, (int)j);
  }
  s.lambda = lambda;
}

// NTFNES::accelerate (ntf/ntfnes.hpp:68-117) after iteration `iter` whose
// relative error is cur_err.
static void nes_accelerate(NTFState &s, NESState &ns, int iter, double cur_err) {
  if (iter <= ns.acc_start) return;
  const int N = (int)s.F.size();
  const long k = s.k;
  const double acc_step = std::pow(iter + 1, 1.0 / ns.acc_exp);
  std::vector<Mat> acc(N);
  std::vector<double> acc_l(k, 1.0);
  for (int mode = 0; mode < N; ++mode) {
    const Mat &F = s.F[mode];
    Mat A(F.r, k);
    for (long c = 0; c < k; ++c) {
      const double sc = mode == N - 1 ? s.lambda[c] : 1.0;
      const double sp = mode == N - 1 ? ns.prev_l[c] : 1.0;
      for (long r = 0; r < F.r; ++r) {
        const double cur = sc * F(r, c);
        A(r, c) = (cur - sp * ns.prev[mode](r, c)) * acc_step + cur;
      }
    }
    normalize_rows_t(A, acc_l);
    acc[mode] = A;
  }
  NTFState trial = s;  // computeObjectiveError(m_acc_t) resets the factors to acc
  ntf_reset(trial, acc, acc_l);
  const double acc_err = ntf_rel_error(trial);
  if (acc_err < cur_err) {
    s.F = trial.F;
    s.lambda = trial.lambda;
    for (int mode = 0; mode < N; ++mode) {
      ns.prox[mode] = acc[mode];
      normalize_rows_t(ns.prox[mode], ns.prox_l);
    }
  } else {
    ntf_reset(s, ns.prox, ns.prox_l);
    if (++ns.acc_fails > ns.fail_limit) {
      ns.acc_fails = 0;
      ns.acc_exp++;
    }
  }
}

// AUNTF::computeNTF, ntf/auntf.hpp:128-177.
static int ntf_loop(NTFState &s, int algo, int iters, double *relerr) {
  int N = s.X.N;
  Mat krp, M;
  std::vector<Mat> duals;  // AO-ADMM auxiliaries, zero at construction (:84-86)
  NESState ns;
  for (int j = 0; j < N; ++j) {
    duals.emplace_back(s.F[j].r, s.k);
    ns.prox.emplace_back(s.F[j].r, s.k);
    ns.prev.emplace_back(s.F[j].r, s.k);
  }
  ns.prox_l.assign(s.k, 1.0);
  ns.prev_l.assign(s.k, 1.0);
  for (int it = 0; it < iters; ++it) {
    for (int j = 0; j < N; ++j) {
      Mat S = gram_leave_out_one(s, j);
      // The reference forms the KRP even when the dimension tree is used
      // (ntf/auntf.hpp:138-139); kept so the CPU timing matches its work.
      krp_leave_out_one(s.F, j, s.k, krp);
      if (s.dimtree)
        mttkrp_dimtree(s.X, s.F, j, s.k, s.dt, M);
      else
        mttkrp_direct(s.X, krp, j, s.k, M);
      int rc;
      if (algo == 5) {
        Mat Hn;
        rc = nes_update(ns, j, S, M, s.k, Hn);
        if (rc == 0) s.F[j] = Hn;
      } else {
        rc = ntf_update(algo, s.F[j], M, S, s.k, &duals[j]);
      }
      if (rc < 0) return rc;
      ncp_normalize_mode(s, j);
    }
    // NTFNES turns acceleration on, which turns the error on (:188,
    // ntf/auntf.hpp:178-181).
    const double err = (relerr || algo == 5) ? ntf_rel_error(s) : 0.0;
    if (relerr) relerr[it] = err;
    if (algo == 5) nes_accelerate(s, ns, it, err);
  }
  return 0;
}

}  // namespace orc

using namespace orc;

// ============================================================================
// C entry points (loaded with ctypes by the tests and bench.py).
// The shared loop of oracle_nmf and oracle_nmf_sparse; Am is the (dense)
// input; sparse_build applies the BUILD_SPARSE start of BPPNMF.
static int nmf_core(int algo, Mat &Am, long m, long n, double *W, double *H, int k, int iters,
                    const double *regW, const double *regH, double symm_reg, double *relerr,
                    double *objective, bool sparse_build);

extern "C" {

void oracle_set_threads(int n) { omp_set_num_threads(n); }
int oracle_get_threads() { return omp_get_max_threads(); }

// C = op(A) * op(B), column-major.
void oracle_gemm(char ta, char tb, long M, long N, long K, const double *A, long lda,
                 const double *B, long ldb, double *C, long ldc) {
  gemm(ta, tb, M, N, K, 1.0, A, lda, B, ldb, 0.0, C, ldc);
}

int oracle_bppnnls(const double *AtA, int n, const double *AtB, long nrhs, double *X) {
  return bppnnls(AtA, n, AtB, nrhs, X);
}

// Runs `iters` iterations of the shared-memory NMF `algo` (0 MU, 1 HALS,
// 2 ANLS-BPP) from the given W (m x k) and H (n x k), both overwritten, as
// NMF(A, W, H) + computeNMF() does (common/nmf.hpp:206-222).  relerr[it] is
// printObjective's relative error after iteration it; objective[it] the
// objective.  Returns 0, or -1 when BPP does not converge.
static double g_last_loop_seconds = 0.0;

int oracle_nmf(int algo, const double *A, long m, long n, double *W, double *H, int k, int iters,
               const double *regW, const double *regH, double symm_reg, double *relerr,
               double *objective) {
  Mat Am(m, n);
  std::memcpy(Am.d.data(), A, sizeof(double) * m * n);
  return nmf_core(algo, Am, m, n, W, H, k, iters, regW, regH, symm_reg, relerr, objective, false);
}

// The sparse build of the shared-memory NMF (nmf/nmf.cpp:48-130 with
// BUILD_SPARSE), A given as CSR (m x n, sorted column ids).  The products
// with a sparse A sum the same non-zero terms as the dense products of its
// densified form, so the loop runs on that dense form; what the sparse build
// changes is BPPNMF's start (nmf/bppnmf.hpp:195-201): it constructs a
// HALSNMF from W and H and takes its factors without running it, so the
// factors enter the loop through HALSNMF's constructor normalize_by_W
// (nmf/hals.hpp:46).
int oracle_nmf_sparse(int algo, long m, long n, const long *indptr, const int *indices,
                      const double *values, double *W, double *H, int k, int iters,
                      const double *regW, const double *regH, double *relerr) {
  Mat Am(m, n);
  for (long i = 0; i < m; ++i)
    for (long e = indptr[i]; e < indptr[i + 1]; ++e) Am(i, indices[e]) += values[e];
  return nmf_core(algo, Am, m, n, W, H, k, iters, regW, regH, -1.0, relerr, nullptr, true);
}

// computeObjectiveError(A, W, H) for a sparse A (common/utils.hpp:190-241),
// A as CSC (col_ptrs n + 1, row_indices, values) as Armadillo's SpMat holds
// it: sqrt(nnzsse + ||Rw Rh^T||_F^2 - nnzwh) with Rw, Rh the R factors of
// economy QRs of W (m x k) and H (n x k), here by modified Gram-Schmidt
// (the R of any QR differs only in the signs of its rows, which the
// Frobenius norm of Rw Rh^T does not see).
static void mgs_r(const double *X, long rows, int k, std::vector<double> &R) {
  std::vector<double> Q(X, X + rows * k);
  R.assign((size_t)k * k, 0.0);
  for (int j = 0; j < k; ++j) {
    double *v = &Q[(size_t)rows * j];
    for (int i = 0; i < j; ++i) {
      const double *q = &Q[(size_t)rows * i];
      double r = 0;
      for (long t = 0; t < rows; ++t) r += q[t] * v[t];
      R[i + (size_t)k * j] = r;
      for (long t = 0; t < rows; ++t) v[t] -= r * q[t];
    }
    double nr = 0;
    for (long t = 0; t < rows; ++t) nr += v[t] * v[t];
    nr = std::sqrt(nr);
    R[j + (size_t)k * j] = nr;
    if (nr > 0)
      for (long t = 0; t < rows; ++t) v[t] /= nr;
  }
}

double oracle_sparse_fit_error(long m, long n, const long *col_ptrs, const int *row_indices,
                               const double *values, const double *W, const double *H, int k) {
  double nnzsse = 0, nnzwh = 0;
  for (long col = 0; col < n; ++col) {
    for (long ii = col_ptrs[col]; ii < col_ptrs[col + 1]; ++ii) {
      const long row = row_indices[ii];
      double t = 0;
      for (int c = 0; c < k; ++c) t += W[row + m * c] * H[col + n * c];
      nnzwh += t * t;
      nnzsse += (values[ii] - t) * (values[ii] - t);
    }
  }
  std::vector<double> Rw, Rh;
  mgs_r(W, m, k, Rw);
  mgs_r(H, n, k, Rh);
  double nrm2 = 0;
  for (int a = 0; a < k; ++a)
    for (int b = 0; b < k; ++b) {
      double s = 0;
      for (int c = 0; c < k; ++c) s += Rw[a + (size_t)k * c] * Rh[b + (size_t)k * c];
      nrm2 += s * s;
    }
  return std::sqrt(nnzsse + (nrm2 - nnzwh));
}

}  // extern "C"

static int nmf_core(int algo, Mat &Am, long m, long n, double *W, double *H, int k, int iters,
                    const double *regW, const double *regH, double symm_reg, double *relerr,
                    double *objective, bool sparse_build) {
  NMFState s;
  s.A = &Am;
  s.m = m;
  s.n = n;
  s.k = k;
  s.W = Mat(m, k);
  s.H = Mat(n, k);
  std::memcpy(s.W.d.data(), W, sizeof(double) * m * k);
  std::memcpy(s.H.d.data(), H, sizeof(double) * n * k);
  for (int i = 0; i < 2; ++i) {
    s.regW[i] = regW ? regW[i] : 0.0;
    s.regH[i] = regH ? regH[i] : 0.0;
  }
  s.symm_reg = symm_reg;
  s.normA = std::sqrt(fro2(Am));
  int rc = 0;
  // steady_clock around the iteration loop, as the reference's tic/toc
  // brackets computeNMF (common/utils.hpp:44-57); the input copy is outside.
  auto t0 = std::chrono::steady_clock::now();
  if (algo == 0) {
    for (int it = 0; it < iters; ++it) {
      // One MU iteration without the trailing normalisation.
      Mat AtW = mul(*s.A, 'T', s.W, 'N');
      Mat WtW = mul(s.W, 'T', s.W, 'N');
      apply_reg(s.regH, WtW);
      Mat HWtW = mul(s.H, 'N', WtW, 'N');
      for (size_t e = 0; e < s.H.d.size(); ++e)
        s.H.d[e] = (s.H.d[e] * AtW.d[e]) / (HWtW.d[e] + EPS_1EM16);
      Mat AH = mul(*s.A, 'N', s.H, 'N');
      Mat HtH = mul(s.H, 'T', s.H, 'N');
      apply_reg(s.regW, HtH);
      Mat WHtH = mul(s.W, 'N', HtH, 'N');
      for (size_t e = 0; e < s.W.d.size(); ++e)
        s.W.d[e] = (s.W.d[e] * AH.d[e]) / (WHtH.d[e] + EPS_1EM16);
      objective_error(s);
      if (relerr) relerr[it] = rel_error(s);
      if (objective) objective[it] = s.objective;
    }
    normalize_by_W(s);  // nmf/mu.hpp:90
  } else if (algo == 1) {
    normalize_by_W(s);  // HALSNMF constructor, nmf/hals.hpp:46
    for (int it = 0; it < iters; ++it) {
      hals_loop(s, 1, relerr ? relerr + it : nullptr);
      if (objective) objective[it] = s.objective;
    }
  } else if (algo == 2) {
    if (sparse_build) normalize_by_W(s);  // nmf/bppnmf.hpp:195-201 via nmf/hals.hpp:46
    for (int it = 0; it < iters; ++it) {
      if (bpp_update(s, true, s.H, s.regW, s.W) < 0) {
        rc = -1;
        break;
      }
      if (bpp_update(s, false, s.W, s.regH, s.H) < 0) {
        rc = -1;
        break;
      }
      objective_error(s);
      if (relerr) relerr[it] = rel_error(s);
      if (objective) objective[it] = s.objective;
    }
    if (rc == 0) normalize_by_W(s);  // nmf/bppnmf.hpp:236
  } else if (algo == 4) {
    // AOADMMNMF (nmf/aoadmm.hpp): constructor normalize_by_W (:78-85), duals
    // U (m x k) and V (n x k) start at zero and persist; H then W per
    // iteration; no final normalisation.
    normalize_by_W(s);
    Mat U(m, k), V(n, k);
    for (int it = 0; it < iters && rc == 0; ++it) {
      Mat WtA = mul(s.W, 'T', *s.A, 'N');  // k x n
      Mat WtW = mul(s.W, 'T', s.W, 'N');
      apply_reg(s.regH, WtW);
      if (!admm_solve(WtW, WtA, s.H, V)) rc = -1;
      Mat AH = mul(*s.A, 'N', s.H, 'N');  // m x k
      Mat HtH = mul(s.H, 'T', s.H, 'N');
      apply_reg(s.regW, HtH);
      Mat AHt(k, m);
      for (long i = 0; i < m; ++i)
        for (long c = 0; c < k; ++c) AHt(c, i) = AH(i, c);
      if (rc == 0 && !admm_solve(HtH, AHt, s.W, U)) rc = -1;
      objective_error(s);
      if (relerr) relerr[it] = rel_error(s);
      if (objective) objective[it] = s.objective;
    }
  } else if (algo == 7) {
    // GNSYMNMF (nmf/gnsym.hpp:178-260): A ~ H H^T for a square A; W is
    // cleared and stays empty.  Gauss-Newton step from at most k conjugate-
    // gradient iterations (tolerance 1e-4 on the squared residual), then the
    // projection max(H - D^T, 0); error by the fast formula (:145-163).
    if (m != n) return -3;
    const double cg_tol = 0.0001;
    const double sqnormA = s.normA * s.normA;
    Mat HtH = mul(s.H, 'T', s.H, 'N');
    Mat AHt = mul(s.H, 'T', *s.A, 'N');  // k x n
    for (int it = 0; it < iters; ++it) {
      Mat grad(k, n), Pk, Dt(k, n), Y(k, n);
      for (long j = 0; j < n; ++j)
        for (long c = 0; c < k; ++c) {
          double hh = 0;
          for (long t = 0; t < k; ++t) hh += HtH(c, t) * s.H(j, t);
          grad(c, j) = -2 * (AHt(c, j) - hh);
        }
      Pk = grad;
      double rsold = 0;
      for (double v : grad.d) rsold += v * v;
      if (rsold > cg_tol) {
        for (long cg = 0; cg < k; ++cg) {
          Mat XtX = mul(Pk, 'N', s.H, 'N');  // k x k
          for (long j = 0; j < n; ++j)
            for (long c = 0; c < k; ++c) {
              double a = 0, b = 0;
              for (long t = 0; t < k; ++t) {
                a += HtH(c, t) * Pk(t, j);
                b += XtX(c, t) * s.H(j, t);
              }
              Y(c, j) = 2 * (a + b);
            }
          double py = 0;
          for (size_t e = 0; e < Pk.d.size(); ++e) py += Pk.d[e] * Y.d[e];
          const double alpha = rsold / py;
          for (size_t e = 0; e < Dt.d.size(); ++e) Dt.d[e] += alpha * Pk.d[e];
          for (size_t e = 0; e < grad.d.size(); ++e) grad.d[e] -= alpha * Y.d[e];
          double rsnew = 0;
          for (double v : grad.d) rsnew += v * v;
          if (rsnew < cg_tol) break;
          const double beta = rsnew / rsold;
          rsold = rsnew;
          for (size_t e = 0; e < Pk.d.size(); ++e) Pk.d[e] = grad.d[e] + beta * Pk.d[e];
        }
      }
      for (long j = 0; j < n; ++j)
        for (long c = 0; c < k; ++c) {
          double v = s.H(j, c) - Dt(c, j);
          s.H(j, c) = v > 0.0 ? v : 0.0;
        }
      HtH = mul(s.H, 'T', s.H, 'N');
      AHt = mul(s.H, 'T', *s.A, 'N');
      double tAHHt = 0, tHH = 0;
      for (long j = 0; j < n; ++j)
        for (long c = 0; c < k; ++c) tAHHt += AHt(c, j) * s.H(j, c);
      for (long a = 0; a < k; ++a)
        for (long b = 0; b < k; ++b) tHH += HtH(a, b) * HtH(b, a);
      s.fit_err_sq = sqnormA - 2 * tAHHt + tHH;
      s.objective = s.fit_err_sq;
      if (relerr) relerr[it] = rel_error(s);
      if (objective) objective[it] = s.objective;
    }
  } else {
    return -2;
  }
  g_last_loop_seconds =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::memcpy(W, s.W.d.data(), sizeof(double) * m * k);
  std::memcpy(H, s.H.d.data(), sizeof(double) * n * k);
  return rc;
}

extern "C" {

// One call of the ADMM solve shared by AO-ADMM NMF and NTF: G k x k (the
// Gram with its regularisation, before the rho shift), R k x n, X and D n x k
// (updated; D persists between calls as the reference's dual does).
// Returns 0, or -1 when G + rho I is not positive definite.
int oracle_admm_solve(const double *G, int k, const double *R, long n, double *X, double *D) {
  Mat Gm(k, k), Rm(k, n), Xm(n, k), Dm(n, k);
  std::memcpy(Gm.d.data(), G, sizeof(double) * k * k);
  std::memcpy(Rm.d.data(), R, sizeof(double) * k * n);
  std::memcpy(Xm.d.data(), X, sizeof(double) * n * k);
  std::memcpy(Dm.d.data(), D, sizeof(double) * n * k);
  if (!admm_solve(Gm, Rm, Xm, Dm)) return -1;
  std::memcpy(X, Xm.d.data(), sizeof(double) * n * k);
  std::memcpy(D, Dm.d.data(), sizeof(double) * n * k);
  return 0;
}

// Smallest and largest eigenvalue of a symmetric k x k matrix (Jacobi).
void oracle_sym_eig_minmax(const double *S, int k, double *mn, double *mx) {
  Mat Sm(k, k);
  std::memcpy(Sm.d.data(), S, sizeof(double) * k * k);
  sym_eig_minmax(Sm, *mn, *mx);
}

// One NTFNES mode update with the proximal centre Hprox (I x k): S k x k, M
// k x I; the new (unnormalised) factor into H (I x k).  Returns 0 or -4.
int oracle_nes_mode_update(const double *S, int k, const double *M, long I, const double *Hprox,
                           double *H) {
  NESState ns;
  ns.prox.emplace_back(I, k);
  ns.prev.emplace_back(I, k);
  ns.prox_l.assign(k, 1.0);
  ns.prev_l.assign(k, 1.0);
  std::memcpy(ns.prox[0].d.data(), Hprox, sizeof(double) * I * k);
  Mat Sm(k, k), Mm(k, I), Hm;
  std::memcpy(Sm.d.data(), S, sizeof(double) * k * k);
  std::memcpy(Mm.d.data(), M, sizeof(double) * k * I);
  int rc = nes_update(ns, 0, Sm, Mm, k, Hm);
  if (rc == 0) std::memcpy(H, Hm.d.data(), sizeof(double) * I * k);
  return rc;
}

// Wall time of the iteration loop of the last oracle_nmf / oracle_mu_f32 call.
double oracle_last_loop_seconds() { return g_last_loop_seconds; }

// Loads a host BLAS for the CPU baseline: `path` is a shared library that
// exports cblas_dgemm (and optionally cblas_sgemm); nthreads > 0 is passed to
// MKL_Set_Num_Threads / openblas_set_num_threads when the library has one.
// Returns 1 with cblas_sgemm, 0 with cblas_dgemm only, -1 when the library or
// cblas_dgemm is missing (the blocked GEMM stays in use).
int oracle_use_blas(const char *path, int nthreads) {
  if (!path) {
    g_dgemm = nullptr;
    g_sgemm = nullptr;
    return -1;
  }
  void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!h) return -1;
  cblas_dgemm_fn d = (cblas_dgemm_fn)dlsym(h, "cblas_dgemm");
  if (!d) return -1;
  if (nthreads > 0) {
    typedef void (*set_fn)(int);
    set_fn f = (set_fn)dlsym(h, "MKL_Set_Num_Threads");
    if (!f) f = (set_fn)dlsym(h, "openblas_set_num_threads");
    if (f) f(nthreads);
  }
  g_dgemm = d;
  g_sgemm = (cblas_sgemm_fn)dlsym(h, "cblas_sgemm");
  return g_sgemm ? 1 : 0;
}

// FP32 variant of the MU loop for a dtype-matched CPU baseline (the reference
// computes in FP64 only): MUNMF::computeNMF (nmf/mu.hpp:48-91) in single
// precision through the loaded cblas_sgemm, error by computeObjectiveError's
// trace formula (common/nmf.hpp:390-424, A^T W recomputed) accumulated in
// FP64.  Returns -3 when no cblas_sgemm was loaded.
int oracle_mu_f32(const float *A, long m, long n, float *W, float *H, int k, int iters,
                  double *relerr) {
  if (!g_sgemm) return -3;
  if (!fits_int(m, n, k, m, n, m)) return -2;
  const int M = (int)m, Nn = (int)n, K = (int)k;
  std::vector<float> AtW((size_t)n * k), WtW((size_t)k * k), HWtW((size_t)n * k);
  std::vector<float> AH((size_t)m * k), HtH((size_t)k * k), WHtH((size_t)m * k);
  double normA2 = 0;
#pragma omp parallel for reduction(+ : normA2)
  for (long e = 0; e < m * n; ++e) normA2 += (double)A[e] * (double)A[e];
  const float eps = (float)EPS_1EM16;
  auto t0 = std::chrono::steady_clock::now();
  for (int it = 0; it < iters; ++it) {
    g_sgemm(102, 112, 111, Nn, K, M, 1.0f, A, M, W, M, 0.0f, AtW.data(), Nn);
    g_sgemm(102, 112, 111, K, K, M, 1.0f, W, M, W, M, 0.0f, WtW.data(), K);
    g_sgemm(102, 111, 111, Nn, K, K, 1.0f, H, Nn, WtW.data(), K, 0.0f, HWtW.data(), Nn);
#pragma omp parallel for
    for (long e = 0; e < n * k; ++e) H[e] = (H[e] * AtW[e]) / (HWtW[e] + eps);
    g_sgemm(102, 111, 111, M, K, Nn, 1.0f, A, M, H, Nn, 0.0f, AH.data(), M);
    g_sgemm(102, 112, 111, K, K, Nn, 1.0f, H, Nn, H, Nn, 0.0f, HtH.data(), K);
    g_sgemm(102, 111, 111, M, K, K, 1.0f, W, M, HtH.data(), K, 0.0f, WHtH.data(), M);
#pragma omp parallel for
    for (long e = 0; e < m * k; ++e) W[e] = (W[e] * AH[e]) / (WHtH[e] + eps);
    // computeObjectiveError: A^T W again, traces in FP64.
    g_sgemm(102, 112, 111, Nn, K, M, 1.0f, A, M, W, M, 0.0f, AtW.data(), Nn);
    g_sgemm(102, 112, 111, K, K, M, 1.0f, W, M, W, M, 0.0f, WtW.data(), K);
    g_sgemm(102, 112, 111, K, K, Nn, 1.0f, H, Nn, H, Nn, 0.0f, HtH.data(), K);
    double tr1 = 0, tr2 = 0;
    for (long e = 0; e < n * k; ++e) tr1 += (double)H[e] * (double)AtW[e];
    for (int a = 0; a < K; ++a)
      for (int b = 0; b < K; ++b) tr2 += (double)WtW[a + K * b] * (double)HtH[b + K * a];
    double fit2 = normA2 - 2 * tr1 + tr2;
    if (relerr) relerr[it] = (fit2 > 0 ? std::sqrt(fit2) : std::sqrt(normA2)) / std::sqrt(normA2);
  }
  // normalize_by_W after the loop (nmf/mu.hpp:90, common/nmf.hpp:152-161).
  for (long c = 0; c < k; ++c) {
    double nrm = 0;
    for (long i = 0; i < m; ++i) nrm += (double)W[i + m * c] * (double)W[i + m * c];
    nrm = std::sqrt(nrm);
    if (nrm > 0) {
      for (long i = 0; i < m; ++i) W[i + m * c] = (float)(W[i + m * c] / nrm);
      for (long j = 0; j < n; ++j) H[j + n * c] = (float)(H[j + n * c] * nrm);
    }
  }
  g_last_loop_seconds =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

// One NTF mode update with the Hadamard-of-Grams S (k x k) and the MTTKRP M
// (k x I_n) given: NTFHALS / NTFMU / NTFANLSBPP::update (ntf/ntfhals.hpp:
// 13-29, ntf/ntfmu.hpp:11-16, ntf/ntfanlsbpp.hpp:15-57) on the current
// factor H (I_n x k), overwritten.  Returns 0, -1 (BPP did not converge) or
// -2 (unknown algo).
int oracle_ntf_mode_update(int algo, long In, int k, double *H, const double *M, const double *S) {
  Mat Hm(In, k), Mm(k, In), Sm(k, k);
  std::memcpy(Hm.d.data(), H, sizeof(double) * In * k);
  std::memcpy(Mm.d.data(), M, sizeof(double) * In * k);
  std::memcpy(Sm.d.data(), S, sizeof(double) * k * k);
  int rc = ntf_update(algo, Hm, Mm, Sm, k);
  std::memcpy(H, Hm.d.data(), sizeof(double) * In * k);
  return rc;
}

// KRP leaving out mode n; out is (prod others) x k column-major.
void oracle_krp_leave_out_one(int N, const long *dims, const double *const *factors, int k, int n,
                              double *out) {
  std::vector<Mat> F(N);
  for (int i = 0; i < N; ++i) {
    F[i] = Mat(dims[i], k);
    std::memcpy(F[i].d.data(), factors[i], sizeof(double) * dims[i] * k);
  }
  Mat K;
  krp_leave_out_one(F, n, k, K);
  std::memcpy(out, K.d.data(), sizeof(double) * K.d.size());
}

// MTTKRP of mode n, out k x I_n.  method 0: Tensor::mttkrp with a formed KRP;
// method 1: dimension tree with split s (the full sweep 0..n is replayed so the
// projections are formed as in a real sweep).
void oracle_mttkrp(int N, const long *dims, const double *X, const double *const *factors, int k,
                   int n, int method, int split, double *out) {
  Tens T{N, std::vector<long>(dims, dims + N), X};
  std::vector<Mat> F(N);
  for (int i = 0; i < N; ++i) {
    F[i] = Mat(dims[i], k);
    std::memcpy(F[i].d.data(), factors[i], sizeof(double) * dims[i] * k);
  }
  Mat M;
  if (method == 0) {
    Mat K;
    krp_leave_out_one(F, n, k, K);
    mttkrp_direct(T, K, n, k, M);
  } else {
    DimTree dt;
    dt.s = split;
    for (int j = 0; j <= n; ++j) mttkrp_dimtree(T, F, j, k, dt, M);
  }
  std::memcpy(out, M.d.data(), sizeof(double) * M.d.size());
}

// AUNTF run: factors (I_i x k) are the raw initial factors; they are first
// normalised with NCPFactors::normalize() (as the AUNTF constructor does,
// ntf/auntf.hpp:88), then `iters` iterations of algo (0 MU, 1 HALS, 2 BPP)
// run.  Factors and lambda are returned; relerr[it] is the relative error
// after iteration it (computeObjectiveError).  dimtree: 0 direct MTTKRP,
// 1 dimension tree with split N/2.
int oracle_ntf(int algo, int N, const long *dims, const double *X, double *const *factors, int k,
               int iters, int dimtree, double *lambda, double *relerr) {
  NTFState s;
  s.X = Tens{N, std::vector<long>(dims, dims + N), X};
  s.k = k;
  s.F.resize(N);
  for (int i = 0; i < N; ++i) {
    s.F[i] = Mat(dims[i], k);
    std::memcpy(s.F[i].d.data(), factors[i], sizeof(double) * dims[i] * k);
  }
  s.lambda.assign(k, 1.0);
  double nx = 0;
  long ne = s.X.numel();
#pragma omp parallel for reduction(+ : nx)
  for (long e = 0; e < ne; ++e) nx += X[e] * X[e];
  s.normX2 = nx;
  s.dimtree = dimtree != 0 && N >= 3;
  s.dt.s = N / 2;
  ncp_normalize_all(s);
  int rc = ntf_loop(s, algo, iters, relerr);
  for (int i = 0; i < N; ++i) std::memcpy(factors[i], s.F[i].d.data(), sizeof(double) * dims[i] * k);
  for (int c = 0; c < k; ++c) lambda[c] = s.lambda[c];
  return rc;
}

// CP loss 0.5 ||X - [[A_0..A_{N-1}]]||^2 (no lambda) and, when grads is
// non-null, its gradient with respect to every factor:
// dL/dA_n = A_n * (Hadamard of the other Grams) - MTTKRP_n^T.  This is the
// transposed NTFNES gradient S*Y - M of ntf/ntfnes.hpp:138-147.
double oracle_cp_loss(int N, const long *dims, const double *X, const double *const *factors,
                      int k, double *const *grads) {
  Tens T{N, std::vector<long>(dims, dims + N), X};
  std::vector<Mat> F(N);
  for (int i = 0; i < N; ++i) {
    F[i] = Mat(dims[i], k);
    std::memcpy(F[i].d.data(), factors[i], sizeof(double) * dims[i] * k);
  }
  Mat K;
  krp_leave_out_one(F, 0, k, K);
  long I0 = dims[0], rest = K.r;
  std::vector<double> low((size_t)I0 * rest);
  gemm('N', 'T', I0, rest, k, 1.0, F[0].d.data(), I0, K.d.data(), rest, 0.0, low.data(), I0);
  double err = 0;
  long ne = T.numel();
#pragma omp parallel for reduction(+ : err)
  for (long e = 0; e < ne; ++e) {
    double d = X[e] - low[e];
    err += d * d;
  }
  if (grads) {
    for (int n = 0; n < N; ++n) {
      Mat S(k, k);
      std::fill(S.d.begin(), S.d.end(), 1.0);
      for (int i = 0; i < N; ++i) {
        if (i == n) continue;
        Mat G = mul(F[i], 'T', F[i], 'N');
        for (size_t e = 0; e < S.d.size(); ++e) S.d[e] *= G.d[e];
      }
      Mat Kn, M;
      krp_leave_out_one(F, n, k, Kn);
      mttkrp_direct(T, Kn, n, k, M);
      Mat AS = mul(F[n], 'N', S, 'N');
      for (long r = 0; r < dims[n]; ++r)
        for (long c = 0; c < k; ++c) grads[n][r + dims[n] * c] = AS(r, c) - M(c, r);
    }
  }
  return 0.5 * err;
}

// DistR2::update, distnmf/distr2.hpp:35-61, for one side: left is the 2x2
// Gram (column-major), right the 2 x n right-hand side; out receives the 2 x n
// solution.
void oracle_r2_update(const double *left, const double *right, long n, double *out) {
  double a = left[0], c = left[1], b = left[2], d = left[3];  // column-major [[a b],[c d]]
  // arma::solve(fast) on a 2x2 system = LU with partial pivoting (dgesv).
  bool sw = std::fabs(c) > std::fabs(a);
  double u11 = sw ? c : a, u12 = sw ? d : b;
  double l21 = (sw ? a : c) / u11;
  double u22 = (sw ? b : d) - l21 * u12;
  double b0 = std::sqrt(a), b1 = std::sqrt(d);
  for (long j = 0; j < n; ++j) {
    double r0 = right[2 * j], r1 = right[2 * j + 1];
    double z0 = sw ? r1 : r0, z1 = (sw ? r0 : r1) - l21 * z0;
    double g1 = z1 / u22;
    double g0 = (z0 - u12 * g1) / u11;
    if (g0 < 0 || g1 < 0) {
      double u = r0 / a, v = r1 / d;
      if (b0 * u >= b1 * v)
        v = 0;
      else
        u = 0;
      g0 = u;
      g1 = v;
    }
    out[2 * j] = g0;
    out[2 * j + 1] = g1;
  }
}

// Rank-2 NMF of one hierarchical node (Node::split, hiernmf/node.hpp:104-188):
// DistAUNMF::computeNMF at p = 1 (distnmf/aunmf.hpp:495-657), H given W then
// W given H with DistR2's update, on A_node = A0(:, cols) (all m rows kept, as
// Node::allocate does).  A0 is CSR (m x n); W (m x 2) and H (ncols x 2) are
// column-major and overwritten.
void oracle_r2_nmf_csr(long m, const long *indptr, const int *indices, const double *data,
                       const long *cols, long ncols, long n, double *W, double *H, int iters) {
  std::vector<long> colmap(n, -1);
  for (long j = 0; j < ncols; ++j) colmap[cols[j]] = j;
  std::vector<double> WtA(2 * ncols), AH(2 * m), G(4), out;
  for (int it = 0; it < iters; ++it) {
    double g00 = 0, g01 = 0, g11 = 0;
    for (long i = 0; i < m; ++i) {
      double a = W[i], b = W[i + m];
      g00 += a * a;
      g01 += a * b;
      g11 += b * b;
    }
    G = {g00, g01, g01, g11};
    std::fill(WtA.begin(), WtA.end(), 0.0);
    for (long i = 0; i < m; ++i)
      for (long e = indptr[i]; e < indptr[i + 1]; ++e) {
        long j = colmap[indices[e]];
        if (j < 0) continue;
        WtA[2 * j] += W[i] * data[e];
        WtA[2 * j + 1] += W[i + m] * data[e];
      }
    out.assign(2 * ncols, 0.0);
    oracle_r2_update(G.data(), WtA.data(), ncols, out.data());
    for (long j = 0; j < ncols; ++j) {
      H[j] = out[2 * j];
      H[j + ncols] = out[2 * j + 1];
    }
    g00 = g01 = g11 = 0;
    for (long j = 0; j < ncols; ++j) {
      double a = H[j], b = H[j + ncols];
      g00 += a * a;
      g01 += a * b;
      g11 += b * b;
    }
    G = {g00, g01, g01, g11};
    std::fill(AH.begin(), AH.end(), 0.0);
    for (long i = 0; i < m; ++i)
      for (long e = indptr[i]; e < indptr[i + 1]; ++e) {
        long j = colmap[indices[e]];
        if (j < 0) continue;
        AH[2 * i] += data[e] * H[j];
        AH[2 * i + 1] += data[e] * H[j + ncols];
      }
    out.assign(2 * m, 0.0);
    oracle_r2_update(G.data(), AH.data(), m, out.data());
    for (long i = 0; i < m; ++i) {
      W[i] = out[2 * i];
      W[i + m] = out[2 * i + 1];
    }
  }
}

// powIter (hiernmf/matutils.hpp:37-120) at p = 1 on A0(:, cols) from q0:
// q = q0/||q0||; repeat: z = A q, q = A^T z, sigma = ||q||, q /= sigma.
double oracle_powiter_csr(long m, const long *indptr, const int *indices, const double *data,
                          const long *cols, long ncols, long n, const double *q0, int iters) {
  std::vector<long> colmap(n, -1);
  for (long j = 0; j < ncols; ++j) colmap[cols[j]] = j;
  std::vector<double> q(q0, q0 + ncols), z(m);
  double nq = 0;
  for (double v : q) nq += v * v;
  nq = std::sqrt(nq);
  for (double &v : q) v /= nq;
  double sigma = 1;
  for (int it = 0; it < iters; ++it) {
    for (long i = 0; i < m; ++i) {
      double s = 0;
      for (long e = indptr[i]; e < indptr[i + 1]; ++e) {
        long j = colmap[indices[e]];
        if (j >= 0) s += data[e] * q[j];
      }
      z[i] = s;
    }
    std::fill(q.begin(), q.end(), 0.0);
    for (long i = 0; i < m; ++i)
      for (long e = indptr[i]; e < indptr[i + 1]; ++e) {
        long j = colmap[indices[e]];
        if (j >= 0) q[j] += data[e] * z[i];
      }
    double s2 = 0;
    for (double v : q) s2 += v * v;
    sigma = std::sqrt(s2);
    for (double &v : q) v /= sigma;
  }
  return sigma;
}

}  // extern "C"
