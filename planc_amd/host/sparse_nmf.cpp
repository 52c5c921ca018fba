// sparse_nmf: drop-in for PLANC's shared-memory NMF driver built with
// BUILD_SPARSE (nmf/nmf.cpp:46-333 with SP_MAT A) on one MI355X.  A is
// generated with density -s (rand_uniform, rand_normal, rand_lowrank) or
// loaded from coord_ascii triples (common/utils.hpp:326-328).
#include "planc/nmf_driver.hpp"

int main(int argc, char *argv[]) {
  try {
    planc::NMFDriver<planc::SP_MAT> d(argc, argv);
    fflush(stdout);
  } catch (const std::exception &e) {
    ERR << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
