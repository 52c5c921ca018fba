// dense_nmf: drop-in for PLANC's shared-memory NMF driver (nmf/nmf.cpp:46-333)
// on one MI355X.  Same flags; A is generated (rand_uniform, rand_normal,
// rand_lowrank) from the Philox streams or loaded from a raw-ascii text file;
// W and H start from U[0,1) drawn with --seed (nmf/nmf.cpp:188-190).
#include <algorithm>
#include <cmath>

#include "planc/nmf.hpp"
#include "planc/parsecommandline.hpp"

namespace planc {

class NMFDriver {
  ParseCommandLine pc;

  template <class NMFTYPE>
  void callNMF() {
    unsigned long long m = pc.globalm(), n = pc.globaln(), k = pc.lowrankk();
    MAT A;
    std::string name = pc.input_file_name();
    const std::string pre = "rand_";
    if (name.compare(0, pre.size(), pre) == 0) {
      std::string type = name.substr(pre.size());
      if (type != "uniform" && type != "normal" && type != "lowrank")
        throw Error("unknown random input type " + type, PLANC_EINVAL);
      A.zeros(m, n);
      const uint64_t seed = 1210873;  // kW_seed_idx of nmf/nmf.cpp:70
      if (type == "uniform") {
        A.randu(seed, 0);
      } else if (type == "normal") {
        for (unsigned long long e = 0; e < m * n; ++e) {
          double z = planc_rng::normal(seed, 0, e);
          A.memptr()[e] = z < 0 ? 0.0 : z;  // negatives zeroed (nmf/nmf.cpp:146)
        }
      } else {
        MAT Wt(m, k), Ht(k, n);
        Wt.randu(seed, 1);
        Ht.randu(seed, 2);
        for (unsigned long long j = 0; j < n; ++j)
          for (unsigned long long i = 0; i < m; ++i) {
            double s = 0;
            for (unsigned long long c = 0; c < k; ++c) s += Wt(i, c) * Ht(c, j);
            A(i, j) = s;
          }
      }
      if (pc.adj_rand()) {
        // A = ceil(kalpha * A + kbeta) with the dense build's kalpha = 1,
        // kbeta = 0 (nmf/nmf.cpp:39-44,164-167); generated inputs only.
        for (unsigned long long e = 0; e < m * n; ++e) A.memptr()[e] = std::ceil(A.memptr()[e]);
      }
    } else {
      if (!A.load_raw_ascii(name)) throw Error("cannot load " + name, PLANC_EINVAL);
      m = A.n_rows;
      n = A.n_cols;
    }
    // Input normalisation (nmf/nmf.cpp:174-185): l2 makes every column of
    // A unit-norm (arma::normalise leaves zero columns zero), max scales A by
    // 1 / max(A).
    if (pc.input_normalization() == L2NORM) {
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
    MAT W(m, k), H(n, k);
    W.randu((uint64_t)pc.initseed(), 0);
    H.randu((uint64_t)pc.initseed(), 1);
    NMFTYPE nmf(A, W, H);
    nmf.num_iterations(pc.iterations());
    nmf.symm_reg(pc.symm_reg());
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
  NMFDriver(int argc, char *argv[]) : pc(argc, argv) {
    pc.parseplancopts();
    pc.printConfig();
    if (pc.dry_run()) return;
    switch (pc.lucalgo()) {
      case MU: callNMF<MUNMF<MAT>>(); break;
      case HALS: callNMF<HALSNMF<MAT>>(); break;
      case ANLSBPP: callNMF<BPPNMF<MAT>>(); break;
      default:
        throw Error("Unsupported algorithm " + std::to_string(pc.lucalgo()), PLANC_EUNSUPPORTED);
    }
  }
};

}  // namespace planc

int main(int argc, char *argv[]) {
  try {
    planc::NMFDriver d(argc, argv);
    fflush(stdout);
  } catch (const std::exception &e) {
    ERR << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
