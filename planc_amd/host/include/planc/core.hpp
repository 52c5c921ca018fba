// Host-side types of the planc_amd C++ API: the column-major matrix that
// stands in for arma::mat (PLANC's MAT, common/utils.h:58), the algotype
// enum (common/utils.h:10-11), timing helpers and the device context.  All
// arithmetic of the hot path runs in libplanc_hip.so through the C ABI of
// include/planc_hip.h; this header only moves data and reports errors.
#ifndef PLANC_AMD_HOST_CORE_HPP_
#define PLANC_AMD_HOST_CORE_HPP_

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stack>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../../include/planc_hip.h"
#include "../../../csrc/philox.h"

#ifndef INFO
#define INFO std::cout
#endif
#ifndef ERR
#define ERR std::cerr
#endif

// Same numbering as the reference (common/utils.h:10-11).
enum algotype { MU, HALS, ANLSBPP, NAIVEANLSBPP, AOADMM, NESTEROV, CPALS, GNSYM, R2, PGD, PGNCG };

namespace planc {

// ---------------------------------------------------------------------------
// Column-major FP64 matrix with the subset of arma::mat's interface the NMF /
// NTF API uses (n_rows, n_cols, memptr, operator(), t, zeros, randu).
class Mat {
 public:
  unsigned long long n_rows = 0, n_cols = 0;
  Mat() {}
  Mat(unsigned long long r, unsigned long long c) : n_rows(r), n_cols(c), d_(r * c, 0.0) {}
  double &operator()(unsigned long long i, unsigned long long j) { return d_[i + n_rows * j]; }
  double operator()(unsigned long long i, unsigned long long j) const { return d_[i + n_rows * j]; }
  double *memptr() { return d_.data(); }
  const double *memptr() const { return d_.data(); }
  unsigned long long n_elem() const { return n_rows * n_cols; }
  void zeros(unsigned long long r, unsigned long long c) {
    n_rows = r;
    n_cols = c;
    d_.assign(r * c, 0.0);
  }
  void zeros() { std::fill(d_.begin(), d_.end(), 0.0); }
  Mat t() const {
    Mat o(n_cols, n_rows);
    for (unsigned long long j = 0; j < n_cols; ++j)
      for (unsigned long long i = 0; i < n_rows; ++i) o(j, i) = (*this)(i, j);
    return o;
  }
  // Fills with U[0,1) from the Philox stream (seed, stream), column-major
  // element order (planc_amd/datagen.py uniform()).
  void randu(uint64_t seed, uint64_t stream) {
    for (unsigned long long e = 0; e < d_.size(); ++e)
      d_[e] = planc_rng::u01(planc_rng::word(seed, stream, e));
  }
  double norm_fro() const {
    double s = 0;
    for (double v : d_) s += v * v;
    return std::sqrt(s);
  }
  // Whitespace-separated text, one row per line (arma::raw_ascii).
  void save_raw_ascii(const std::string &fn) const {
    std::ofstream o(fn);
    o.precision(17);
    for (unsigned long long i = 0; i < n_rows; ++i) {
      for (unsigned long long j = 0; j < n_cols; ++j) o << (j ? " " : "") << (*this)(i, j);
      o << "\n";
    }
  }
  bool load_raw_ascii(const std::string &fn) {
    std::ifstream in(fn);
    if (!in) return false;
    std::vector<std::vector<double>> rows;
    std::string line;
    while (std::getline(in, line)) {
      std::istringstream ss(line);
      std::vector<double> r;
      double v;
      while (ss >> v) r.push_back(v);
      if (!r.empty()) rows.push_back(r);
    }
    if (rows.empty()) return false;
    zeros(rows.size(), rows[0].size());
    for (size_t i = 0; i < rows.size(); ++i) {
      if (rows[i].size() != n_cols) return false;
      for (size_t j = 0; j < n_cols; ++j) (*this)(i, j) = rows[i][j];
    }
    return true;
  }

 private:
  std::vector<double> d_;
};

// ---------------------------------------------------------------------------
// Sparse matrix (stands in for arma::sp_mat, PLANC's SP_MAT): held as CSR
// (row pointers, sorted column ids, FP64 values), the layout the device
// products take.  coord_ascii text is one "row col value" triple per line,
// 0-based (Armadillo's coord_ascii, read by common/utils.hpp:326-328);
// entries given twice are summed.
class SpMat {
 public:
  unsigned long long n_rows = 0, n_cols = 0;
  std::vector<long> row_ptr{0};
  std::vector<int> col_idx;
  std::vector<double> values;
  SpMat() {}
  SpMat(unsigned long long r, unsigned long long c) : n_rows(r), n_cols(c), row_ptr(r + 1, 0) {}
  unsigned long long n_nonzero() const { return values.size(); }
  double norm_fro() const {
    double s = 0;
    for (double v : values) s += v * v;
    return std::sqrt(s);
  }
  /// Builds the CSR from triplets (sorted by row, then column; duplicates summed).
  void from_triplets(unsigned long long r, unsigned long long c,
                     std::vector<std::pair<std::pair<long, long>, double>> &t) {
    std::sort(t.begin(), t.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    n_rows = r;
    n_cols = c;
    row_ptr.assign(r + 1, 0);
    col_idx.clear();
    values.clear();
    for (size_t e = 0; e < t.size(); ++e) {
      if (!col_idx.empty() && e > 0 && t[e].first == t[e - 1].first) {
        values.back() += t[e].second;
        continue;
      }
      row_ptr[t[e].first.first + 1]++;
      col_idx.push_back((int)t[e].first.second);
      values.push_back(t[e].second);
    }
    for (unsigned long long i = 0; i < r; ++i) row_ptr[i + 1] += row_ptr[i];
  }
  bool load_coord_ascii(const std::string &fn) {
    std::ifstream in(fn);
    if (!in) return false;
    std::vector<std::pair<std::pair<long, long>, double>> t;
    long r, c, mr = -1, mc = -1;
    double v;
    std::string line;
    while (std::getline(in, line)) {
      if (line.find_first_not_of(" \t\r") == std::string::npos) continue;  // blank line
      std::istringstream ss(line);
      if (!(ss >> r >> c >> v)) return false;  // malformed triple
      if (r < 0 || c < 0) return false;
      t.push_back({{r, c}, v});
      mr = std::max(mr, r);
      mc = std::max(mc, c);
    }
    from_triplets((unsigned long long)(mr + 1), (unsigned long long)(mc + 1), t);
    return true;
  }
  void save_coord_ascii(const std::string &fn) const {
    std::ofstream o(fn);
    o.precision(17);
    for (unsigned long long i = 0; i < n_rows; ++i)
      for (long e = row_ptr[i]; e < row_ptr[i + 1]; ++e) o << i << " " << col_idx[e] << " " << values[e] << "\n";
  }
};

typedef Mat MAT;
typedef SpMat SP_MAT;
typedef std::vector<double> VEC;
typedef std::vector<float> FVEC;
typedef std::vector<unsigned long long> UVEC;

// tic/toc as in common/utils.hpp:44-57.
inline std::stack<std::chrono::steady_clock::time_point> &tictoc_stack() {
  static thread_local std::stack<std::chrono::steady_clock::time_point> s;
  return s;
}
inline void tic() { tictoc_stack().push(std::chrono::steady_clock::now()); }
inline double toc() {
  auto d = std::chrono::steady_clock::now() - tictoc_stack().top();
  tictoc_stack().pop();
  return std::chrono::duration<double>(d).count();
}

// ---------------------------------------------------------------------------
// Errors of the C ABI become exceptions on the C++ side (the reference's
// drivers catch std::exception, nmf/nmf.cpp:326-333).
class Error : public std::runtime_error {
 public:
  int status;
  Error(const std::string &m, int s) : std::runtime_error(m), status(s) {}
};

// One device context per host thread (the ABI's threading rule).
class Device {
 public:
  static Device &get() {
    static thread_local Device d;
    return d;
  }
  planc_ctx *ctx() {
    if (!ctx_) {
      int rc = planc_create(device_, &ctx_);
      if (rc) throw Error("planc_create: no usable MI355X device", rc);
    }
    return ctx_;
  }
  void check(int rc, const char *what) {
    if (rc) throw Error(std::string(what) + ": " + planc_last_error(ctx_), rc);
  }
  ~Device() {
    if (ctx_) planc_destroy(ctx_);
  }

 private:
  Device() {}
  int device_ = 0;
  planc_ctx *ctx_ = nullptr;
};

// Owning device buffer of n elements of T.
template <class T>
class DeviceBuffer {
 public:
  explicit DeviceBuffer(size_t n) : n_(n) {
    Device &d = Device::get();
    void *p = nullptr;
    d.check(planc_malloc(d.ctx(), &p, sizeof(T) * (n ? n : 1)), "planc_malloc");
    p_ = (T *)p;
  }
  ~DeviceBuffer() {
    if (p_) planc_free(Device::get().ctx(), p_);
  }
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  T *get() { return p_; }
  void upload(const T *h) {
    Device &d = Device::get();
    d.check(planc_memcpy_h2d(d.ctx(), p_, h, sizeof(T) * n_), "planc_memcpy_h2d");
  }
  void download(T *h) {
    Device &d = Device::get();
    d.check(planc_memcpy_d2h(d.ctx(), h, p_, sizeof(T) * n_), "planc_memcpy_d2h");
  }

 private:
  T *p_ = nullptr;
  size_t n_;
};

// FP64 host matrix <-> FP32 device copy.
inline void upload_f32(const double *h, size_t n, DeviceBuffer<float> &d) {
  std::vector<float> t(n);
  for (size_t i = 0; i < n; ++i) t[i] = (float)h[i];
  d.upload(t.data());
}
inline void download_f32(DeviceBuffer<float> &d, double *h, size_t n) {
  std::vector<float> t(n);
  d.download(t.data());
  for (size_t i = 0; i < n; ++i) h[i] = t[i];
}

// Precision of the bulk products (PLANC_F32 / PLANC_BF16 / PLANC_F16).
inline int &default_precision() {
  static int p = PLANC_F32;
  return p;
}

}  // namespace planc

#endif  // PLANC_AMD_HOST_CORE_HPP_
