// NMFDriver (nmf/nmf.cpp:16-321) for a dense (MAT) or sparse (SP_MAT, the
// reference's BUILD_SPARSE build) input: A is generated (rand_uniform,
// rand_normal, rand_lowrank) from the Philox streams or loaded (raw-ascii
// text for a dense A, coord_ascii triples for a sparse A), optionally
// normalised, and factored by the algorithm of -a on one MI355X.
#ifndef PLANC_AMD_HOST_NMF_DRIVER_HPP_
#define PLANC_AMD_HOST_NMF_DRIVER_HPP_

#include <algorithm>
#include <cmath>

#include "nmf.hpp"
#include "parsecommandline.hpp"

namespace planc {

// The input stage shared by the drivers: the options, and A generated or
// loaded as nmf/nmf.cpp:73-168 does.
template <class T>
class DriverInput {
 protected:
  ParseCommandLine pc;
  static constexpr uint64_t kSeed = 1210873;  // kW_seed_idx of nmf/nmf.cpp:70
  bool m_symm_flag = false;
  double m_symm_reg = -1;

  DriverInput(int argc, char *argv[]) : pc(argc, argv) {}

  // ---- dense inputs (nmf/nmf.cpp:131-168) --------------------------------
  void generate(MAT &A, const std::string &type, unsigned long long m, unsigned long long n,
                unsigned long long k) {
    A.zeros(m, n);
    if (type == "uniform") {
      A.randu(kSeed, 0);
    } else if (type == "normal") {
      for (unsigned long long e = 0; e < m * n; ++e) {
        double z = planc_rng::normal(kSeed, 0, e);
        A.memptr()[e] = z;
      }
    } else {
      MAT Wt(m, k), Ht(k, n);
      Wt.randu(kSeed, 1);
      Ht.randu(kSeed, 2);
      for (unsigned long long j = 0; j < n; ++j)
        for (unsigned long long i = 0; i < m; ++i) {
          double s = 0;
          for (unsigned long long c = 0; c < k; ++c)
            s += Wt(i, c) * (m_symm_flag ? Wt(j, c) : Ht(c, j));  // symmetric: Wtrue Wtrue^T
          A(i, j) = s;
        }
    }
    if (m_symm_flag && type != "lowrank") {
      MAT B = A;
      for (unsigned long long j = 0; j < n; ++j)
        for (unsigned long long i = 0; i < m; ++i) A(i, j) = 0.5 * (B(i, j) + B(j, i));
    }
    if (type == "normal")  // negatives zeroed (nmf/nmf.cpp:146)
      for (unsigned long long e = 0; e < m * n; ++e) A.memptr()[e] = std::max(A.memptr()[e], 0.0);
    if (pc.adj_rand())  // kalpha = 1, kbeta = 0 of the dense build (nmf/nmf.cpp:39-44,164-167)
      for (unsigned long long e = 0; e < m * n; ++e) A.memptr()[e] = std::ceil(A.memptr()[e]);
  }
  bool load(MAT &A, const std::string &name) { return A.load_raw_ascii(name); }
  void normalise(MAT &A) {
    const unsigned long long m = A.n_rows, n = A.n_cols;
    if (pc.input_normalization() == L2NORM) {  // arma::normalise: unit columns, zero columns kept
      for (unsigned long long j = 0; j < n; ++j) {
        double s = 0;
        for (unsigned long long i = 0; i < m; ++i) s += A(i, j) * A(i, j);
        s = std::sqrt(s);
        if (s > 0)
          for (unsigned long long i = 0; i < m; ++i) A(i, j) /= s;
      }
    } else if (pc.input_normalization() == MAXNORM) {
      double mx = A.memptr()[0];
      for (unsigned long long e = 0; e < m * n; ++e) mx = std::max(mx, A.memptr()[e]);
      const double inv = 1 / mx;
      for (unsigned long long e = 0; e < m * n; ++e) A.memptr()[e] *= inv;
    }
  }
  static double mean(const MAT &A) {
    double s = 0;
    for (unsigned long long e = 0; e < A.n_elem(); ++e) s += A.memptr()[e];
    return s / (double)A.n_elem();
  }
  static double maxval(const MAT &A) {
    double mx = A.memptr()[0];
    for (unsigned long long e = 0; e < A.n_elem(); ++e) mx = std::max(mx, A.memptr()[e]);
    return mx;
  }

  // ---- sparse inputs (nmf/nmf.cpp:73-130) --------------------------------
  // Positions: entry (i, j) is present when its Philox uniform of stream 3
  // is below the sparsity (-s); values U[0,1) (uniform), N(0,1) (normal) or
  // (W H)(i, j) (lowrank) from streams 4 / 1-2.  Symmetric inputs average A
  // and A^T (0.5 (A + A^T)) over the union of positions, as the reference
  // does with sparsity halved.
  void generate(SpMat &A, const std::string &type, unsigned long long m, unsigned long long n,
                unsigned long long k) {
    const double dens = m_symm_flag ? 0.5 * pc.sparsity() : pc.sparsity();
    std::vector<std::pair<std::pair<long, long>, double>> t;
    MAT Wt, Ht;
    if (type == "lowrank") {
      Wt.zeros(m, k);
      Ht.zeros(k, n);
      Wt.randu(kSeed, 1);
      Ht.randu(kSeed, 2);
    }
    for (unsigned long long j = 0; j < n; ++j)
      for (unsigned long long i = 0; i < m; ++i) {
        const uint64_t e = i + m * j;
        if (!(planc_rng::u01(planc_rng::word(kSeed, 3, e)) < dens)) continue;
        double v;
        if (type == "uniform") {
          v = planc_rng::u01(planc_rng::word(kSeed, 4, e));
        } else if (type == "normal") {
          v = planc_rng::normal(kSeed, 4, e);
        } else {
          v = 0;
          for (unsigned long long c = 0; c < k; ++c)
            v += Wt(i, c) * (m_symm_flag ? Wt(j, c) : Ht(c, j));
        }
        t.push_back({{(long)i, (long)j}, v});
        if (m_symm_flag) t.push_back({{(long)j, (long)i}, v});
      }
    if (m_symm_flag)
      for (auto &x : t) x.second *= 0.5;  // 0.5 (A + A^T)
    A.from_triplets(m, n, t);
    // Adjust and project the non-zeros (nmf/nmf.cpp:118-129; kbeta = 0).
    for (double &v : A.values) {
      if (pc.adj_rand()) v = std::ceil(v);
      if (v < 0) v = 0;
    }
  }
  bool load(SpMat &A, const std::string &name) { return A.load_coord_ascii(name); }
  void normalise(SpMat &A) {
    if (pc.input_normalization() == L2NORM) {
      std::vector<double> cn(A.n_cols, 0.0);
      for (size_t e = 0; e < A.values.size(); ++e) cn[A.col_idx[e]] += A.values[e] * A.values[e];
      for (double &c : cn) c = std::sqrt(c);
      for (size_t e = 0; e < A.values.size(); ++e)
        if (cn[A.col_idx[e]] > 0) A.values[e] /= cn[A.col_idx[e]];
    } else if (pc.input_normalization() == MAXNORM) {
      double mx = A.values.empty() ? 0.0 : A.values[0];
      for (double v : A.values) mx = std::max(mx, v);
      if (A.values.size() < A.n_rows * A.n_cols) mx = std::max(mx, 0.0);  // implicit zeros
      const double inv = 1 / mx;
      for (double &v : A.values) v *= inv;
    }
  }
  static double mean(const SpMat &A) {
    double s = 0;
    for (double v : A.values) s += v;
    return s / ((double)A.n_rows * (double)A.n_cols);
  }
  static double maxval(const SpMat &A) {
    double mx = A.values.size() < A.n_rows * A.n_cols ? 0.0 : A.values[0];
    for (double v : A.values) mx = std::max(mx, v);
    return mx;
  }

  // A from -i: "rand_uniform" / "rand_normal" / "rand_lowrank" generated at
  // -d m n with rank k, otherwise the file loaded (m, n set from it).
  T read_input(unsigned long long &m, unsigned long long &n, unsigned long long k) {
    T A;
    std::string name = pc.input_file_name();
    const std::string pre = "rand_";
    if (name.compare(0, pre.size(), pre) == 0) {
      std::string type = name.substr(pre.size());
      if (type != "uniform" && type != "normal" && type != "lowrank")
        throw Error("unknown random input type " + type, PLANC_EINVAL);
      generate(A, type, m, n, k);
    } else {
      if (!load(A, name)) throw Error("cannot load " + name, PLANC_EINVAL);
      m = A.n_rows;
      n = A.n_cols;
    }
    return A;
  }
};

template <class T>
class NMFDriver : DriverInput<T> {
  using DriverInput<T>::pc;
  using DriverInput<T>::m_symm_flag;
  using DriverInput<T>::m_symm_reg;

  template <class NMFTYPE>
  void callNMF() {
    unsigned long long m = pc.globalm(), n = pc.globaln(), k = pc.lowrankk();
    T A = this->read_input(m, n, k);
    this->normalise(A);
    MAT W(m, k), H(n, k);
    W.randu((uint64_t)pc.initseed(), 0);
    H.randu((uint64_t)pc.initseed(), 1);
    if (m_symm_flag) {
      // nmf/nmf.cpp:191-199: H scaled by 2 sqrt(mean(A)/k), W = H; a zero
      // --symm value becomes max(A)^2.
      const double s = 2 * std::sqrt(this->mean(A) / (double)k);
      for (unsigned long long e = 0; e < H.n_elem(); ++e) H.memptr()[e] *= s;
      W = H;
      if (m_symm_reg == 0.0) {
        const double mx = this->maxval(A);
        m_symm_reg = mx * mx;
      }
    }
    NMFTYPE nmf(A, W, H);
    nmf.num_iterations(pc.iterations());
    nmf.symm_reg(m_symm_reg);
    nmf.updalgo(pc.lucalgo());
    nmf.regW(pc.regW());
    nmf.regH(pc.regH());
    nmf.precision(pc.precision());
    INFO << "completed constructor::A::" << m << "x" << n << std::endl;
    tic();
    nmf.computeNMF();
    INFO << "time taken:" << toc() << std::endl;
    if (!pc.output_file_name().empty()) {
      nmf.getLeftLowRankFactor().save_raw_ascii(pc.output_file_name() + "_W");
      nmf.getRightLowRankFactor().save_raw_ascii(pc.output_file_name() + "_H");
    }
  }

 public:
  NMFDriver(int argc, char *argv[]) : DriverInput<T>(argc, argv) {
    pc.parseplancopts();
    m_symm_reg = pc.symm_reg();
    // Symmetric NMF (nmf/nmf.cpp:256-268): a square A and ANLS-BPP or GNSYM.
    if (m_symm_reg != -1) {
      m_symm_flag = true;
      if (pc.globalm() != pc.globaln())
        throw Error("Symmetric Regularization enabled and input matrix is not square", PLANC_EINVAL);
      if (pc.lucalgo() != ANLSBPP && pc.lucalgo() != GNSYM)
        throw Error("Symmetric Regularization enabled is only enabled for ANLSBPP and GNSYM",
                    PLANC_EINVAL);
    }
    pc.printConfig();
    if (pc.dry_run()) return;
    switch (pc.lucalgo()) {
      case MU: callNMF<MUNMF<T>>(); break;
      case HALS: callNMF<HALSNMF<T>>(); break;
      case ANLSBPP: callNMF<BPPNMF<T>>(); break;
      case AOADMM: callNMF<AOADMMNMF<T>>(); break;
      case GNSYM: callNMF<GNSYMNMF<T>>(); break;
      default:
        throw Error("Unsupported algorithm " + std::to_string(pc.lucalgo()), PLANC_EUNSUPPORTED);
    }
  }
};

}  // namespace planc

#endif  // PLANC_AMD_HOST_NMF_DRIVER_HPP_
