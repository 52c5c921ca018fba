// dense_ntf: drop-in for PLANC's shared-memory NTF driver (ntf/ntf.cpp:22-71)
// on one MI355X.  `-i rand_*` builds a uniform random tensor of `-d` (as the
// reference does), any other name reads `<name>` with its `.info` file.
#include "planc/ntf.hpp"
#include "planc/parsecommandline.hpp"

namespace planc {

class NTFDriver {
 public:
  template <class NTFTYPE>
  void callNTF(ParseCommandLine &pc) {
    Tensor t(pc.dimensions());
    std::string fn = pc.input_file_name();
    INFO << "Input filename = " << fn << std::endl;
    const bool npy = fn.size() > 4 && fn.compare(fn.size() - 4, 4, ".npy") == 0;
    if (npy) {
      // A NumPy array (the reader of common/npyio.hpp, which the reference's
      // distributed driver uses, distntf/distntfio.hpp:141).
      NumPyArray a;
      a.load(fn);
      a.printInfo();
      t = *a.m_input_tensor;
    } else if (!fn.empty() && fn.compare(0, 5, "rand_") != 0) {
      t.read(fn);
    }
    NTFTYPE solver(t, (int)pc.lowrankk(), pc.lucalgo());
    solver.num_it(pc.iterations());
    solver.compute_error(pc.compute_error());
    solver.precision(pc.precision());
    if (pc.dim_tree()) solver.dim_tree(true);
    tic();
    solver.computeNTF();
    INFO << "time taken:" << toc() << std::endl;
  }
};

}  // namespace planc

int main(int argc, char *argv[]) {
  try {
    planc::ParseCommandLine pc(argc, argv);
    pc.parseplancopts();
    pc.printConfig();
    if (pc.dry_run()) return 0;
    planc::NTFDriver d;
    switch (pc.lucalgo()) {
      case MU: d.callNTF<planc::NTFMU>(pc); break;
      case HALS: d.callNTF<planc::NTFHALS>(pc); break;
      case ANLSBPP: d.callNTF<planc::NTFANLSBPP>(pc); break;
      case AOADMM: d.callNTF<planc::NTFAOADMM>(pc); break;
      default:
        ERR << "Wrong algorithm choice. Quitting.." << pc.lucalgo() << std::endl;
        return 1;
    }
  } catch (const std::exception &e) {
    ERR << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
