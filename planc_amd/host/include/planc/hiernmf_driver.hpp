// HierNMFDriver (hiernmf/hiernmf.cpp:17-194) on one MI355X: a sparse A is
// generated or loaded as the NMF drivers do (no normalisation: the reference
// reads it with NONE, hiernmf/hiernmf.cpp:78-84), the best-first tree with
// the --nodes budget is built by planc_hiernmf_run, the node queue is printed
// and, with -o, every node but the root is written as the reference writes
// it (hiernmf/hiernmf.cpp:159-180): <prefix>_<index>_W and _H hold the
// parent's factor columns on the node's side as raw FP64 (DistIO::writeOutput
// for one process, distnmf/distio.hpp:640-673), <prefix>_<index>_cols the
// node's columns as raw ASCII, one per line.
#ifndef PLANC_AMD_HOST_HIERNMF_DRIVER_HPP_
#define PLANC_AMD_HOST_HIERNMF_DRIVER_HPP_

#include <cstdio>
#include <unordered_map>

#include "nmf_driver.hpp"

namespace planc {

class HierNMFDriver : DriverInput<SpMat> {
  static void write_raw(const std::string &fn, const double *p, size_t cnt) {
    FILE *f = std::fopen(fn.c_str(), "wb");
    if (!f) throw Error("cannot write " + fn, PLANC_EINVAL);
    const size_t w = std::fwrite(p, sizeof(double), cnt, f);
    std::fclose(f);
    if (w != cnt) throw Error("short write to " + fn, PLANC_EINVAL);
  }
  static void write_cols(const std::string &fn, const long *c, long cnt) {
    std::ofstream out(fn);
    if (!out) throw Error("cannot write " + fn, PLANC_EINVAL);
    for (long j = 0; j < cnt; ++j) out << c[j] << "\n";
  }

  void build_tree() {
    unsigned long long m = pc.globalm(), n = pc.globaln();
    SpMat A = read_input(m, n, pc.lowrankk());
    // The loop of hiernmf/hiernmf.cpp:138-147 accepts at most nodes - 1
    // nodes, so the reported nodes lie on levels below nodes: a level cap of
    // nodes + 1 (the library's limit is 62) never changes the tree.
    const int budget = pc.nodes() < 1 ? 1 : pc.nodes();
    planc_hier_desc d{};
    d.m = (long)m;
    d.n = (long)n;
    d.depth = std::min(62, std::max(2, budget + 1));
    d.iters = pc.iterations();
    d.initseed = pc.initseed();
    d.nodes = budget;
    std::vector<float> vals(A.values.begin(), A.values.end());
    const size_t cap = budget > 1 ? 2 * (size_t)budget - 1 : 3;
    std::vector<planc_hier_node> table(cap);
    std::vector<long> cols((size_t)n * d.depth);
    const bool out = !pc.output_file_name().empty();
    std::vector<double> H(out ? 2 * cols.size() : 0), W(out ? 2 * (size_t)m * cap : 0);
    int count = 0;
    Device &dev = Device::get();
    tic();
    dev.check(planc_hiernmf_run(dev.ctx(), &d, A.row_ptr.data(), A.col_idx.data(), vals.data(),
                                table.data(), &count, cols.data(), out ? H.data() : nullptr,
                                out ? W.data() : nullptr),
              "planc_hiernmf_run");
    std::printf("H2NMF took %.3lf secs.\n", toc());
    std::printf("idx level m n sigma score\n");
    std::unordered_map<long, int> slot;
    for (int t = 0; t < count; ++t) {
      const planc_hier_node &x = table[t];
      slot[x.index] = t;
      std::printf("%ld %d %llu %ld %f %f\n", x.index, x.level, m, x.ncols, x.sigma, x.score);
    }
    if (!out) return;
    for (int t = 0; t < count; ++t) {
      const planc_hier_node &x = table[t];
      if (x.index == 0) continue;
      const planc_hier_node &p = table[slot.at((x.index - 1) / 2)];
      const int s = x.parent_side;
      const std::string pre = pc.output_file_name() + "_" + std::to_string(x.index);
      write_raw(pre + "_W", W.data() + (size_t)slot.at(p.index) * 2 * m + (size_t)s * m, m);
      std::vector<double> h(p.ncols);
      for (long j = 0; j < p.ncols; ++j) h[j] = H[2 * (p.col_offset + j) + s];
      write_raw(pre + "_H", h.data(), h.size());
      write_cols(pre + "_cols", cols.data() + x.col_offset, x.ncols);
    }
  }

 public:
  HierNMFDriver(int argc, char *argv[]) : DriverInput<SpMat>(argc, argv) {
    pc.parseplancopts();
    pc.printConfig();
    if (pc.dry_run()) return;
    build_tree();
  }
};

}  // namespace planc

#endif  // PLANC_AMD_HOST_HIERNMF_DRIVER_HPP_
