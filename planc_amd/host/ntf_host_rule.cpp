// ntf_host_rule: the AUNTF plugin contract (ntf/auntf.hpp:45) exercised from
// the host.  A user rule derives from planc::AUNTF and implements
// update(mode); here a host FP64 NTF-MU (the rule of ntf/ntfmu.hpp:11-16)
// computed from gram_without_one and ncp_mttkrp_t[mode], while the tensor
// stays on the device and the MTTKRPs run there.  Also runs the built-in
// NTFHALS through the same host loop (use_device_loop(false)) and through
// the device loop, so both paths can be compared.
//   ntf_host_rule <tensor file with .info> <k> <iters> <dimtree 0|1>
#include <cstdio>

#include "planc/ntf.hpp"

namespace {

class HostNTFMU : public planc::AUNTF {
 protected:
  planc::MAT update(const int mode) {
    const planc::MAT &H = m_ncp_factors.factor(mode);  // I x k
    const planc::MAT &M = ncp_mttkrp_t[mode];           // k x I
    const planc::MAT &S = gram_without_one;             // k x k
    const unsigned long long I = H.n_rows, k = H.n_cols;
    planc::MAT out(k, I);
    for (unsigned long long r = 0; r < I; ++r)
      for (unsigned long long c = 0; c < k; ++c) {
        double hs = 0;
        for (unsigned long long t = 0; t < k; ++t) hs += H(r, t) * S(t, c);
        out(c, r) = H(r, c) * M(c, r) / (hs + 0.000001);  // EPSILON, common/utils.h:49
      }
    return out;
  }

 public:
  HostNTFMU(const planc::Tensor &t, int k) : AUNTF(t, k, MU) {}
};

void print(const char *tag, const std::vector<double> &r) {
  std::printf("%s", tag);
  for (double v : r) std::printf(" %.12g", v);
  std::printf("\n");
}

}  // namespace

int main(int argc, char *argv[]) {
  if (argc != 5) {
    ERR << "usage: ntf_host_rule <tensor> <k> <iters> <dimtree>" << std::endl;
    return 2;
  }
  try {
    planc::Tensor X;
    X.read(argv[1]);
    const int k = std::atoi(argv[2]), iters = std::atoi(argv[3]);
    const bool dt = std::atoi(argv[4]) != 0;
    HostNTFMU mu(X, k);
    mu.num_it(iters);
    mu.compute_error(true);
    mu.dim_tree(dt);
    mu.computeNTF();
    print("host_mu", mu.relative_errors());
    planc::NTFHALS h1(X, k, HALS), h2(X, k, HALS);
    for (planc::NTFHALS *h : {&h1, &h2}) {
      h->num_it(iters);
      h->compute_error(true);
      h->dim_tree(dt);
    }
    h1.use_device_loop(false);
    h1.computeNTF();
    h2.computeNTF();
    print("hals_host_loop", h1.relative_errors());
    print("hals_device_loop", h2.relative_errors());
  } catch (const std::exception &e) {
    ERR << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
