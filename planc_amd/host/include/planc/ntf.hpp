// Tensor, NCPFactors and AUNTF with the reference's API (common/tensor.hpp,
// common/ncpfactors.hpp, ntf/auntf.hpp, ntf/ntfhals.hpp, ntf/ntfmu.hpp,
// ntf/ntfanlsbpp.hpp).  computeNTF() of the built-in update rules runs the
// whole alternating loop on the MI355X (planc_ntf_run); a user subclass that
// overrides update(mode) keeps the reference's plugin contract: the device
// computes each MTTKRP (KRP fused) and the host calls update(mode).
#ifndef PLANC_AMD_HOST_NTF_HPP_
#define PLANC_AMD_HOST_NTF_HPP_

#include <memory>

#include "core.hpp"

namespace planc {

/// Dense N-way tensor, mode 0 fastest (common/tensor.hpp:23-28, 36-521).
class Tensor {
 public:
  std::vector<double> m_data;

  Tensor() {}
  explicit Tensor(const UVEC &dims) : m_dims(dims) {
    m_numel = 1;
    for (auto d : dims) m_numel *= d;
    m_data.assign(m_numel, 0.0);
    randu(103);
  }
  Tensor(const UVEC &dims, const double *data) : m_dims(dims) {
    m_numel = 1;
    for (auto d : dims) m_numel *= d;
    m_data.assign(data, data + m_numel);
  }
  int modes() const { return (int)m_dims.size(); }
  UVEC dimensions() const { return m_dims; }
  int dimension(int i) const { return (int)m_dims[i]; }
  unsigned long long numel() const { return m_numel; }
  /// Uniform [0,1) entries from the Philox stream of `seed` (deterministic,
  /// unlike the reference's shared mt19937 under OpenMP, tensor.hpp:219-235).
  void randu(const int seed = 103) {
    for (unsigned long long e = 0; e < m_numel; ++e)
      m_data[e] = planc_rng::u01(planc_rng::word((uint64_t)seed, 0, e));
  }
  /// Squared Frobenius norm, as Tensor::norm() (common/tensor.hpp:350-356).
  double norm() const {
    double s = 0;
    for (double v : m_data) s += v * v;
    return s;
  }
  /// Squared difference (common/tensor.hpp:361-369).
  double err(const Tensor &b) const {
    double s = 0;
    for (unsigned long long e = 0; e < m_numel; ++e) {
      double d = m_data[e] - b.m_data[e];
      s += d * d;
    }
    return s;
  }
  /// `<name>.info` (modes, dims) + raw little-endian doubles
  /// (common/tensor.hpp:430-503).
  void write(const std::string &filename) const {
    std::string base = filename.substr(0, filename.find_last_of('.'));
    std::ofstream info(base + ".info");
    info << m_dims.size() << "\n";
    for (auto d : m_dims) info << d << "\n";
    info << "\n";
    FILE *fp = fopen(filename.c_str(), "wb");
    if (!fp) throw Error("Tensor::write: cannot open " + filename, PLANC_EINVAL);
    size_t nw = fwrite(m_data.data(), sizeof(double), m_numel, fp);
    fclose(fp);
    if (nw != m_numel) throw Error("Tensor::write: short write", PLANC_EINVAL);
  }
  void read(const std::string &filename) {
    std::string base = filename.substr(0, filename.find_last_of('.'));
    std::ifstream info(base + ".info");
    if (!info) throw Error("Tensor::read: missing " + base + ".info", PLANC_EINVAL);
    int modes = 0;
    info >> modes;
    m_dims.assign(modes, 0);
    for (int i = 0; i < modes; ++i) info >> m_dims[i];
    m_numel = 1;
    for (auto d : m_dims) m_numel *= d;
    m_data.assign(m_numel, 0.0);
    FILE *fp = fopen(filename.c_str(), "rb");
    if (!fp) throw Error("Tensor::read: cannot open " + filename, PLANC_EINVAL);
    size_t nr = fread(m_data.data(), sizeof(double), m_numel, fp);
    fclose(fp);
    if (nr != m_numel) throw Error("Tensor::read: short read", PLANC_EINVAL);
  }

 private:
  UVEC m_dims;
  unsigned long long m_numel = 0;
};

/// CP factors I_i x k plus lambda (common/ncpfactors.hpp:22-446).
class NCPFactors {
 public:
  NCPFactors(const UVEC &dims, const int &k, bool trans) : m_dims(dims), m_k(k) {
    m_factors.resize(dims.size());
    for (size_t i = 0; i < dims.size(); ++i) {
      unsigned long long r = dims[i] > 0 ? dims[i] : 1;
      if (trans)
        m_factors[i].zeros(k, r);
      else
        m_factors[i].zeros(r, k);
      m_factors[i].randu(103, i);
    }
    m_lambda.assign(k, 1.0);
  }
  int rank() const { return m_k; }
  UVEC dimensions() const { return m_dims; }
  MAT &factor(const int i) { return m_factors[i]; }
  const MAT &factor(const int i) const { return m_factors[i]; }
  int modes() const { return (int)m_factors.size(); }
  VEC lambda() const { return m_lambda; }
  void set(const int i, const MAT &f) {
    if (f.n_elem() != m_factors[i].n_elem()) throw Error("NCPFactors::set: size mismatch", PLANC_EINVAL);
    m_factors[i] = f;
  }
  void set_lambda(const VEC &l) { m_lambda = l; }
  /// Hadamard of the Grams except mode n (gram_leave_out_one, :139-148).
  void gram_leave_out_one(const unsigned int n, MAT *out) const {
    out->zeros(m_k, m_k);
    for (int a = 0; a < m_k; ++a)
      for (int b = 0; b < m_k; ++b) (*out)(a, b) = 1.0;
    for (size_t i = 0; i < m_factors.size(); ++i) {
      if (i == n) continue;
      const MAT &F = m_factors[i];
      for (int a = 0; a < m_k; ++a)
        for (int b = 0; b < m_k; ++b) {
          double s = 0;
          for (unsigned long long r = 0; r < F.n_rows; ++r) s += F(r, a) * F(r, b);
          (*out)(a, b) *= s;
        }
    }
  }
  /// normalize() (:329-339): lambda = product of the column norms.
  void normalize() {
    std::fill(m_lambda.begin(), m_lambda.end(), 1.0);
    for (auto &F : m_factors)
      for (int c = 0; c < m_k; ++c) {
        double s = 0;
        for (unsigned long long r = 0; r < F.n_rows; ++r) s += F(r, c) * F(r, c);
        s = std::sqrt(s);
        if (s > 0)
          for (unsigned long long r = 0; r < F.n_rows; ++r) F(r, c) /= s;
        m_lambda[c] *= s;
      }
  }
  /// normalize(mode) (:346-351): lambda replaced by that mode's norms.
  void normalize(int mode) {
    MAT &F = m_factors[mode];
    for (int c = 0; c < m_k; ++c) {
      double s = 0;
      for (unsigned long long r = 0; r < F.n_rows; ++r) s += F(r, c) * F(r, c);
      s = std::sqrt(s);
      m_lambda[c] = s;
      if (s > 0)
        for (unsigned long long r = 0; r < F.n_rows; ++r) F(r, c) /= s;
    }
  }

 private:
  UVEC m_dims;
  int m_k;
  std::vector<MAT> m_factors;
  VEC m_lambda;
};

/// MTTKRP of mode n (k x I_n) on the device with the KRP fused
/// (common/ntf_utils.hpp:49-53 + common/tensor.hpp:246-327).
inline void mttkrp(const int n, const Tensor &X, const NCPFactors &F, MAT *out) {
  Device &dev = Device::get();
  const int N = X.modes(), k = F.rank();
  DeviceBuffer<float> dX(X.numel());
  upload_f32(X.m_data.data(), X.numel(), dX);
  std::vector<std::unique_ptr<DeviceBuffer<float>>> df;
  std::vector<const float *> ptrs;
  std::vector<long> dims;
  for (int i = 0; i < N; ++i) {
    dims.push_back((long)X.dimension(i));
    df.emplace_back(new DeviceBuffer<float>((size_t)X.dimension(i) * k));
    upload_f32(F.factor(i).memptr(), (size_t)X.dimension(i) * k, *df.back());
    ptrs.push_back(df.back()->get());
  }
  DeviceBuffer<float> dout((size_t)k * X.dimension(n));
  dev.check(planc_mttkrp(dev.ctx(), PLANC_F32, dX.get(), N, dims.data(), ptrs.data(), k, n, dout.get(), k),
            "planc_mttkrp");
  out->zeros(k, X.dimension(n));
  download_f32(dout, out->memptr(), (size_t)k * X.dimension(n));
}

/// Alternating-update NTF (ntf/auntf.hpp:40-251).
class AUNTF {
 protected:
  NCPFactors m_ncp_factors;
  std::vector<MAT> ncp_mttkrp_t;  /// k x I_n MTTKRP of each mode
  MAT gram_without_one;
  /// Update rule of one mode: returns the new factor transposed (k x I_n).
  /// Built-in rules return an empty matrix and run on the device instead.
  virtual MAT update(const int mode) { return MAT(); }
  /// Device algorithm of a built-in rule, or -1 for a host update(mode).
  virtual int device_algo() const { return -1; }

 private:
  const Tensor &m_input_tensor;
  int m_num_it = 20;
  int m_current_it = 0;
  bool m_compute_error = false;
  bool m_enable_dim_tree = false;
  bool m_accelerated = false;
  const int m_low_rank_k;
  const algotype m_updalgo;
  double m_rel_error = 0;
  std::vector<double> m_relerr;
  int m_precision = default_precision();

  void run_device_loop() {
    Device &dev = Device::get();
    const int N = m_input_tensor.modes(), k = m_low_rank_k;
    DeviceBuffer<float> dX(m_input_tensor.numel());
    upload_f32(m_input_tensor.m_data.data(), m_input_tensor.numel(), dX);
    std::vector<std::unique_ptr<DeviceBuffer<float>>> df;
    std::vector<float *> ptrs;
    planc_ntf_desc d{};
    d.algo = device_algo();
    d.dtype = m_precision;
    d.nmodes = N;
    for (int i = 0; i < N; ++i) {
      d.dims[i] = m_input_tensor.dimension(i);
      size_t sz = (size_t)d.dims[i] * k;
      df.emplace_back(new DeviceBuffer<float>(sz));
      // The device normalises like the constructor; pass the unnormalised
      // factors (lambda folded into the last mode).
      MAT F = m_ncp_factors.factor(i);
      if (i == N - 1) {
        VEC l = m_ncp_factors.lambda();
        for (int c = 0; c < k; ++c)
          for (unsigned long long r = 0; r < F.n_rows; ++r) F(r, c) *= l[c];
      }
      upload_f32(F.memptr(), sz, *df.back());
      ptrs.push_back(df.back()->get());
    }
    d.k = k;
    d.iters = m_num_it;
    d.dimtree = m_enable_dim_tree ? 1 : 0;
    d.compute_error = m_compute_error ? 1 : 0;
    DeviceBuffer<double> dl(k);
    m_relerr.assign(m_num_it, 0.0);
    dev.check(planc_ntf_run(dev.ctx(), &d, dX.get(), nullptr, ptrs.data(), dl.get(), m_relerr.data()),
              "planc_ntf_run");
    for (int i = 0; i < N; ++i) {
      MAT F(d.dims[i], k);
      download_f32(*df[i], F.memptr(), (size_t)d.dims[i] * k);
      m_ncp_factors.set(i, F);
    }
    VEC l(k);
    dl.download(l.data());
    m_ncp_factors.set_lambda(l);
    for (m_current_it = 0; m_current_it < m_num_it; ++m_current_it) {
      INFO << "iter::" << m_current_it << std::endl;
      if (m_compute_error) {
        m_rel_error = m_relerr[m_current_it];
        INFO << "relative_error at it::" << m_current_it << "::" << m_rel_error << std::endl;
      }
    }
  }

  // Host update rules: device MTTKRP per mode, host update(mode).
  void run_host_update_loop() {
    const int N = m_input_tensor.modes();
    for (m_current_it = 0; m_current_it < m_num_it; ++m_current_it) {
      INFO << "iter::" << m_current_it << std::endl;
      for (int j = 0; j < N; ++j) {
        m_ncp_factors.gram_leave_out_one(j, &gram_without_one);
        mttkrp(j, m_input_tensor, m_ncp_factors, &ncp_mttkrp_t[j]);
        MAT factor = update(j);
        m_ncp_factors.set(j, factor.t());
        m_ncp_factors.normalize(j);
      }
    }
  }

 public:
  AUNTF(const Tensor &i_tensor, const int i_k, algotype i_algo)
      : m_ncp_factors(i_tensor.dimensions(), i_k, false),
        m_input_tensor(i_tensor),
        m_low_rank_k(i_k),
        m_updalgo(i_algo) {
    m_ncp_factors.normalize();
    gram_without_one.zeros(i_k, i_k);
    ncp_mttkrp_t.resize(i_tensor.modes());
  }
  virtual ~AUNTF() {}
  NCPFactors &ncp_factors() { return m_ncp_factors; }
  void dim_tree(bool i_dim_tree) { m_enable_dim_tree = i_dim_tree; }
  double current_error() const { return m_rel_error; }
  void num_it(const int i_n) { m_num_it = i_n; }
  void compute_error(bool e) { m_compute_error = e; }
  void accelerated(const bool &a) {
    m_accelerated = a;
    m_compute_error = true;
  }
  int current_it() const { return m_current_it; }
  /// Bulk MTTKRP precision: PLANC_F32 (default), PLANC_BF16 or PLANC_F16.
  void precision(int p) { m_precision = p; }
  const std::vector<double> &relative_errors() const { return m_relerr; }
  /// reset (ntf/auntf.hpp:186-198): each mode set and normalised, then
  /// lambda taken from new_factors.
  void reset(const NCPFactors &new_factors, bool trans = false) {
    for (int i = 0; i < m_input_tensor.modes(); ++i) {
      m_ncp_factors.set(i, trans ? new_factors.factor(i).t() : new_factors.factor(i));
      m_ncp_factors.normalize(i);
    }
    m_ncp_factors.set_lambda(new_factors.lambda());
  }
  void computeNTF() {
    if (device_algo() >= 0)
      run_device_loop();
    else
      run_host_update_loop();
  }
};

/// NTFHALS::update (ntf/ntfhals.hpp:13-29) on the device.
class NTFHALS : public AUNTF {
 protected:
  int device_algo() const { return PLANC_HALS; }

 public:
  NTFHALS(const Tensor &t, const int k, algotype a) : AUNTF(t, k, a) {}
};

/// NTFMU::update (ntf/ntfmu.hpp:11-16) on the device.
class NTFMU : public AUNTF {
 protected:
  int device_algo() const { return PLANC_MU; }

 public:
  NTFMU(const Tensor &t, const int k, algotype a) : AUNTF(t, k, a) {}
};

/// NTFANLSBPP::update (ntf/ntfanlsbpp.hpp:15-57) on the device.
class NTFANLSBPP : public AUNTF {
 protected:
  int device_algo() const { return PLANC_ANLSBPP; }

 public:
  NTFANLSBPP(const Tensor &t, const int k, algotype a) : AUNTF(t, k, a) {}
};

}  // namespace planc

#endif  // PLANC_AMD_HOST_NTF_HPP_
