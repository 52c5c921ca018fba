// NMF<T> and its shared-memory update algorithms, with the reference's API
// (common/nmf.hpp:24-520; nmf/mu.hpp, nmf/hals.hpp, nmf/bppnmf.hpp).
// computeNMF() runs the whole loop on the MI355X through planc_nmf_run:
// A, W and H are uploaded once, the loop is device resident, W and H come
// back at the end.  T is the dense matrix type (planc::Mat).
#ifndef PLANC_AMD_HOST_NMF_HPP_
#define PLANC_AMD_HOST_NMF_HPP_

#include "core.hpp"

namespace planc {

template <class T>
class NMF {
 protected:
  const T &A;     /// input matrix of size m x n (caller owns it)
  MAT W, H;       /// left and right low-rank factors, m x k and n x k
  MAT Winit, Hinit;
  unsigned int m, n, k;
  double objective_err = 1e12, fit_err_sq = 0;
  double normA = 0, normW = 0, normH = 0, l1normW = 0, l1normH = 0;
  double m_symm_reg = -1;
  unsigned int m_num_iterations = 20;
  double m_tolerance = -1;
  algotype m_updalgo = MU;
  FVEC m_regW{0.0f, 0.0f}, m_regH{0.0f, 0.0f};
  int m_precision = default_precision();
  std::vector<double> m_relerr;  /// relative error after every iteration

  /// normalize_by_W (common/nmf.hpp:152-161) on the host copies.
  void normalize_by_W() {
    for (unsigned c = 0; c < k; ++c) {
      double s = 0;
      for (unsigned i = 0; i < m; ++i) s += W(i, c) * W(i, c);
      s = std::sqrt(s);
      if (s > 0) {
        for (unsigned i = 0; i < m; ++i) W(i, c) /= s;
        for (unsigned j = 0; j < n; ++j) H(j, c) *= s;
      }
    }
  }

  /// One device run of `algo` for num_iterations() iterations.
  void run_on_device(int algo) {
    planc_nmf_desc d{};
    d.algo = algo;
    d.dtype = m_precision;
    d.m = m;
    d.n = n;
    d.k = (int)k;
    d.iters = (int)m_num_iterations;
    d.regW[0] = m_regW[0];
    d.regW[1] = m_regW[1];
    d.regH[0] = m_regH[0];
    d.regH[1] = m_regH[1];
    d.compute_error = 1;  // always on in the shared-memory driver (nmf/nmf.cpp:205)
    d.symm_reg = m_symm_reg;
    m_relerr.assign(m_num_iterations, 0.0);
    device_nmf(A, d, W, H, m_relerr);
    normA = A.norm_fro();
    for (unsigned it = 0; it < m_num_iterations; ++it) printObjective(it);
  }

  // The device loop for a dense A (planc_nmf_run) ...
  static void device_nmf(const Mat &A, const planc_nmf_desc &d, MAT &W, MAT &H,
                         std::vector<double> &relerr) {
    Device &dev = Device::get();
    const size_t m = d.m, n = d.n, k = d.k;
    DeviceBuffer<float> dA(m * n), dW(m * k), dH(n * k);
    upload_f32(A.memptr(), m * n, dA);
    upload_f32(W.memptr(), m * k, dW);
    upload_f32(H.memptr(), n * k, dH);
    dev.check(planc_nmf_run(dev.ctx(), &d, dA.get(), (long)m, nullptr, nullptr, dW.get(), dH.get(),
                            relerr.data()),
              "planc_nmf_run");
    download_f32(dW, W.memptr(), m * k);
    download_f32(dH, H.memptr(), n * k);
  }
  // ... and for a sparse A (planc_nmf_sparse_run, the reference's BUILD_SPARSE build).
  static void device_nmf(const SpMat &A, const planc_nmf_desc &d, MAT &W, MAT &H,
                         std::vector<double> &relerr) {
    Device &dev = Device::get();
    const size_t m = d.m, n = d.n, k = d.k, nnz = A.n_nonzero();
    DeviceBuffer<long> dp(m + 1);
    DeviceBuffer<int> di(nnz ? nnz : 1);
    DeviceBuffer<float> dv(nnz ? nnz : 1), dW(m * k), dH(n * k);
    dp.upload(A.row_ptr.data());
    if (nnz) {
      di.upload(A.col_idx.data());
      upload_f32(A.values.data(), nnz, dv);
    }
    upload_f32(W.memptr(), m * k, dW);
    upload_f32(H.memptr(), n * k, dH);
    dev.check(planc_nmf_sparse_run(dev.ctx(), &d, dp.get(), di.get(), dv.get(), dW.get(), dH.get(),
                                   relerr.data()),
              "planc_nmf_sparse_run");
    download_f32(dW, W.memptr(), m * k);
    download_f32(dH, H.memptr(), n * k);
  }

 public:
  /// Random initialisation (seeds 89 / 73 as the reference, through the
  /// Philox stream since Armadillo's RNG is not available), then
  /// normalize_by_W (common/nmf.hpp:179-199).
  NMF(const T &input, const unsigned int rank) : A(input) {
    m = A.n_rows;
    n = A.n_cols;
    k = rank;
    W.zeros(m, k);
    H.zeros(n, k);
    W.randu(89, 0);
    H.randu(73, 0);
    normalize_by_W();
    normA = A.norm_fro();
  }
  /// Given initial factors, for comparing algorithms from the same start
  /// (common/nmf.hpp:206-222).
  NMF(const T &input, const MAT &leftlowrankfactor, const MAT &rightlowrankfactor) : A(input) {
    if (leftlowrankfactor.n_cols != rightlowrankfactor.n_cols)
      throw Error("NMF: W and H must have the same number of columns", PLANC_EINVAL);
    W = leftlowrankfactor;
    H = rightlowrankfactor;
    Winit = W;
    Hinit = H;
    m = A.n_rows;
    n = A.n_cols;
    k = W.n_cols;
    normA = A.norm_fro();
  }
  virtual ~NMF() {}

  virtual void computeNMF() = 0;

  MAT getLeftLowRankFactor() { return W; }
  MAT getRightLowRankFactor() { return H; }

  /// Objective and fit of the current factors (common/nmf.hpp:390-424),
  /// evaluated on the host in FP64 (a diagnostic; the loop evaluates the same
  /// quantity on the device every iteration).
  void computeObjectiveError() {
    double tr1 = 0, tr2 = 0;
    std::vector<double> WtW((size_t)k * k, 0.0), HtH((size_t)k * k, 0.0);
    for (unsigned a = 0; a < k; ++a)
      for (unsigned b = 0; b < k; ++b) {
        double s = 0, t = 0;
        for (unsigned i = 0; i < m; ++i) s += W(i, a) * W(i, b);
        for (unsigned j = 0; j < n; ++j) t += H(j, a) * H(j, b);
        WtW[a + (size_t)k * b] = s;
        HtH[a + (size_t)k * b] = t;
      }
    tr1 = trace_hatw(A);
    for (unsigned a = 0; a < k; ++a)
      for (unsigned b = 0; b < k; ++b) tr2 += WtW[a + (size_t)k * b] * HtH[b + (size_t)k * a];
    double froW = 0, froH = 0;
    for (unsigned c = 0; c < k; ++c) {
      froW += WtW[c + (size_t)k * c];
      froH += HtH[c + (size_t)k * c];
    }
    normW = std::sqrt(froW);
    normH = std::sqrt(froH);
    fit_err_sq = normA * normA - 2 * tr1 + tr2;
    objective_err = fit_err_sq + m_regW[0] * froW + m_regH[0] * froH;
  }

  /// tr(H^T A^T W) for a dense or a sparse A.
  double trace_hatw(const Mat &A) const {
    double t = 0;
    for (unsigned j = 0; j < n; ++j)
      for (unsigned i = 0; i < m; ++i) {
        double wh = 0;
        for (unsigned c = 0; c < k; ++c) wh += W(i, c) * H(j, c);
        t += A(i, j) * wh;
      }
    return t;
  }
  double trace_hatw(const SpMat &A) const {
    double t = 0;
    for (unsigned i = 0; i < m; ++i)
      for (long e = A.row_ptr[i]; e < A.row_ptr[i + 1]; ++e) {
        double wh = 0;
        for (unsigned c = 0; c < k; ++c) wh += W(i, c) * H(A.col_idx[e], c);
        t += A.values[e] * wh;
      }
    return t;
  }

  /// printObjective (common/nmf.hpp:461-478) for iteration itr of the last
  /// computeNMF (the relative error recorded on the device).
  void printObjective(const int itr) {
    double rel = (itr >= 0 && (size_t)itr < m_relerr.size()) ? m_relerr[itr] : 0.0;
    INFO << "Completed it = " << itr << "::algo::" << m_updalgo << "::k::" << k << std::endl;
    INFO << "error::" << rel * normA << "::relative error::" << rel << std::endl;
  }

  void num_iterations(const int it) { m_num_iterations = it; }
  const unsigned int num_iterations() const { return m_num_iterations; }
  void tolerance(const double tol) { m_tolerance = tol; }
  double tolerance() { return m_tolerance; }
  void regW(const FVEC &r) { m_regW = r.size() >= 2 ? r : FVEC{0.0f, 0.0f}; }
  void regH(const FVEC &r) { m_regH = r.size() >= 2 ? r : FVEC{0.0f, 0.0f}; }
  FVEC regW() { return m_regW; }
  FVEC regH() { return m_regH; }
  void symm_reg(const double &s) { m_symm_reg = s; }
  double symm_reg() { return m_symm_reg; }
  void updalgo(algotype a) { m_updalgo = a; }
  /// Bulk-product precision: PLANC_F32 (default), PLANC_BF16 or PLANC_F16.
  void precision(int p) { m_precision = p; }
  /// Relative error after each iteration of the last computeNMF.
  const std::vector<double> &relative_errors() const { return m_relerr; }
  double objective() const { return objective_err; }
};

/// MUNMF (nmf/mu.hpp:48-91): H then W multiplicative updates, error every
/// iteration, normalize_by_W after the loop.
template <class T>
class MUNMF : public NMF<T> {
 public:
  MUNMF(const T &A, int lowrank) : NMF<T>(A, lowrank) { this->m_updalgo = MU; }
  MUNMF(const T &A, const MAT &llf, const MAT &rlf) : NMF<T>(A, llf, rlf) { this->m_updalgo = MU; }
  void computeNMF() { this->run_on_device(PLANC_MU); }
};

/// HALSNMF (nmf/hals.hpp:50-116): normalize_by_W in the constructor, HALS
/// sweeps over H's and W's columns, normalize_by_W every iteration.
template <class T>
class HALSNMF : public NMF<T> {
 public:
  HALSNMF(const T &A, int lowrank) : NMF<T>(A, lowrank) {
    this->m_updalgo = HALS;
    this->normalize_by_W();  // nmf/hals.hpp:41
  }
  HALSNMF(const T &A, const MAT &llf, const MAT &rlf) : NMF<T>(A, llf, rlf) {
    this->m_updalgo = HALS;
    this->normalize_by_W();  // nmf/hals.hpp:46
  }
  void computeNMF() { this->run_on_device(PLANC_HALS); }
};

/// BPPNMF (nmf/bppnmf.hpp:187-241): W given H, then H given W, each by
/// block principal pivoting NNLS in FP64; normalize_by_W after the loop.
template <class T>
class BPPNMF : public NMF<T> {
 public:
  BPPNMF(const T &A, int lowrank) : NMF<T>(A, lowrank) { this->m_updalgo = ANLSBPP; }
  BPPNMF(const T &A, const MAT &llf, const MAT &rlf) : NMF<T>(A, llf, rlf) {
    this->m_updalgo = ANLSBPP;
  }
  void computeNMF() { this->run_on_device(PLANC_ANLSBPP); }
};

/// AOADMMNMF (nmf/aoadmm.hpp:11-195): normalize_by_W in the constructor,
/// per iteration an ADMM solve for H then for W (duals carried over, zero at
/// the start of computeNMF), no final normalisation.
template <class T>
class AOADMMNMF : public NMF<T> {
 public:
  AOADMMNMF(const T &A, int lowrank) : NMF<T>(A, lowrank) {
    this->m_updalgo = AOADMM;
    this->normalize_by_W();  // nmf/aoadmm.hpp:79
  }
  AOADMMNMF(const T &A, const MAT &llf, const MAT &rlf) : NMF<T>(A, llf, rlf) {
    this->m_updalgo = AOADMM;
    this->normalize_by_W();  // nmf/aoadmm.hpp:83
  }
  void computeNMF() { this->run_on_device(PLANC_AOADMM); }
};

/// GNSYMNMF (nmf/gnsym.hpp:11-262): symmetric NMF A ~ H H^T of a square A by
/// Gauss-Newton steps with conjugate gradients; W is cleared (empty).
template <class T>
class GNSYMNMF : public NMF<T> {
 public:
  GNSYMNMF(const T &A, int lowrank) : NMF<T>(A, lowrank) { this->m_updalgo = GNSYM; }
  GNSYMNMF(const T &A, const MAT &llf, const MAT &rlf) : NMF<T>(A, llf, rlf) {
    this->m_updalgo = GNSYM;
  }
  void computeNMF() {
    if (this->m != this->n) throw Error("GNSYMNMF: A must be square", PLANC_EINVAL);
    this->run_on_device(PLANC_GNSYM);
    this->W.zeros(0, 0);  // this->W.clear(), nmf/gnsym.hpp:73,98
  }
};

}  // namespace planc

#endif  // PLANC_AMD_HOST_NMF_HPP_
