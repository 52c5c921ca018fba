// dense_nmf: drop-in for PLANC's shared-memory NMF driver (nmf/nmf.cpp:46-333)
// on one MI355X.  Same flags; A is generated (rand_uniform, rand_normal,
// rand_lowrank) from the Philox streams or loaded from a raw-ascii text file;
// W and H start from U[0,1) drawn with --seed (nmf/nmf.cpp:188-190).
#include "planc/nmf_driver.hpp"

int main(int argc, char *argv[]) {
  try {
    planc::NMFDriver<planc::MAT> d(argc, argv);
    fflush(stdout);
  } catch (const std::exception &e) {
    ERR << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
