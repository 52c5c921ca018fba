// hiernmf: drop-in for PLANC's hierarchical rank-2 NMF driver
// (hiernmf/hiernmf.cpp:196-202, the BUILD_SPARSE build) on one MI355X.
#include "planc/hiernmf_driver.hpp"

int main(int argc, char *argv[]) {
  try {
    planc::HierNMFDriver d(argc, argv);
    fflush(stdout);
  } catch (const std::exception &e) {
    ERR << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
