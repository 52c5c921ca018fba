// format_check: reads and writes the on-disk formats of the host API (no GPU
// work): Tensor .info + raw doubles (common/tensor.hpp:430-503), NumPy .npy
// (common/npyio.hpp), coord_ascii sparse matrices (common/utils.hpp:326-328)
// and raw_ascii dense matrices.  Prints what it read and whether a write /
// read round trip reproduces it exactly.
//   format_check tensor <file>    format_check npy <file>
//   format_check coord <file>     format_check raw <file>
#include <cstdio>

#include "planc/ntf.hpp"

using namespace planc;

static double sum(const std::vector<double> &v) {
  double s = 0;
  for (double x : v) s += x;
  return s;
}

int main(int argc, char *argv[]) {
  if (argc != 3) {
    ERR << "usage: format_check {tensor|npy|coord|raw} <file>" << std::endl;
    return 2;
  }
  const std::string what = argv[1], fn = argv[2];
  try {
    std::printf("%s", "");
    if (what == "tensor") {
      Tensor t;
      t.read(fn);
      std::printf("modes %d numel %llu sum %.17g first %.17g last %.17g\n", t.modes(), t.numel(),
                  sum(t.m_data), t.m_data.front(), t.m_data.back());
      const std::string copy = fn + ".copy.bin";
      t.write(copy);
      Tensor u;
      u.read(copy);
      std::printf("roundtrip %s\n", (u.m_data == t.m_data && u.dimensions() == t.dimensions()) ? "ok" : "FAIL");
    } else if (what == "npy") {
      NumPyArray a;
      a.load(fn);
      std::printf("modes %lld fortran %d word %lld dims", (long long)a.modes(), (int)a.fortran_order(),
                  (long long)a.word_size());
      for (auto d : a.dims()) std::printf(" %llu", d);
      const Tensor &t = *a.m_input_tensor;
      std::printf("\nnumel %llu sum %.17g first %.17g last %.17g\n", t.numel(), sum(t.m_data), t.m_data.front(),
                  t.m_data.back());
    } else if (what == "coord") {
      SpMat A;
      if (!A.load_coord_ascii(fn)) throw Error("cannot load " + fn, PLANC_EINVAL);
      std::printf("rows %llu cols %llu nnz %llu sum %.17g\n", A.n_rows, A.n_cols, A.n_nonzero(), sum(A.values));
      const std::string copy = fn + ".copy.txt";
      A.save_coord_ascii(copy);
      SpMat B;
      B.load_coord_ascii(copy);
      std::printf("roundtrip %s\n", (B.row_ptr == A.row_ptr && B.col_idx == A.col_idx && B.values == A.values)
                                        ? "ok" : "FAIL");
    } else if (what == "raw") {
      MAT A;
      if (!A.load_raw_ascii(fn)) throw Error("cannot load " + fn, PLANC_EINVAL);
      double s = 0;
      for (unsigned long long e = 0; e < A.n_elem(); ++e) s += A.memptr()[e];
      std::printf("rows %llu cols %llu sum %.17g\n", A.n_rows, A.n_cols, s);
    } else {
      ERR << "unknown format " << what << std::endl;
      return 2;
    }
  } catch (const std::exception &e) {
    ERR << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
