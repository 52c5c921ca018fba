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

/// Reader of a NumPy .npy file into a Tensor (NumPyArray, common/npyio.hpp:
/// 12-107): the header's descr, fortran_order and shape are parsed (format
/// versions 1-3), and numel little-endian doubles are read in file order into
/// the mode-0-fastest Tensor.  As in the reference the fortran_order flag is
/// recorded but not applied, so a C-order array comes in with its dimensions
/// as written and its elements in C order (a warning is printed).  Only 8-byte
/// floats are accepted ('<f8'; the reference reads doubles whatever the word
/// size).
class NumPyArray {
 public:
  Tensor *m_input_tensor = nullptr;
  NumPyArray() {}
  ~NumPyArray() { delete m_input_tensor; }
  NumPyArray(const NumPyArray &) = delete;
  NumPyArray &operator=(const NumPyArray &) = delete;
  bool fortran_order() const { return m_fortran_order; }
  int64_t word_size() const { return m_word_size; }
  int64_t modes() const { return m_modes; }
  const UVEC &dims() const { return m_dims; }

  void load(const std::string &fname) {
    FILE *fp = fopen(fname.c_str(), "rb");
    if (!fp) throw Error("NumPyArray: could not load the file " + fname, PLANC_EINVAL);
    std::string header;
    try {
      header = read_header(fp);
      parse_header(header);
    } catch (...) {
      fclose(fp);
      throw;
    }
    if (!m_fortran_order)
      ERR << "NumPyArray: " << fname << " is in C order; read in file order as mode-0-fastest" << std::endl;
    delete m_input_tensor;
    std::vector<double> buf;
    unsigned long long numel = 1;
    for (auto d : m_dims) numel *= d;
    buf.resize(numel);
    size_t nread = fread(buf.data(), sizeof(double), numel, fp);
    fclose(fp);
    if (nread != numel)
      throw Error("NumPyArray: read " + std::to_string(nread) + " of " + std::to_string(numel) + " values",
                  PLANC_EINVAL);
    m_input_tensor = new Tensor(m_dims, buf.data());
  }
  void printInfo() const {
    INFO << "modes::" << m_modes << "::dims::";
    for (auto d : m_dims) INFO << d << " ";
    INFO << "::fortran_order::" << m_fortran_order << "::word_size::" << m_word_size << std::endl;
  }

 private:
  int64_t m_word_size = 0;
  bool m_fortran_order = false;
  int64_t m_modes = 0;
  UVEC m_dims;

  static std::string read_header(FILE *fp) {
    unsigned char pre[8];
    if (fread(pre, 1, 8, fp) != 8 || pre[0] != 0x93 || std::memcmp(pre + 1, "NUMPY", 5) != 0)
      throw Error("NumPyArray: not an .npy file", PLANC_EINVAL);
    const int major = pre[6];
    size_t hlen = 0;
    if (major == 1) {
      unsigned char l[2];
      if (fread(l, 1, 2, fp) != 2) throw Error("NumPyArray: short header", PLANC_EINVAL);
      hlen = l[0] | (l[1] << 8);
    } else if (major == 2 || major == 3) {
      unsigned char l[4];
      if (fread(l, 1, 4, fp) != 4) throw Error("NumPyArray: short header", PLANC_EINVAL);
      hlen = (size_t)l[0] | ((size_t)l[1] << 8) | ((size_t)l[2] << 16) | ((size_t)l[3] << 24);
    } else {
      throw Error("NumPyArray: unknown format version", PLANC_EINVAL);
    }
    std::string h(hlen, '\0');
    if (fread(&h[0], 1, hlen, fp) != hlen) throw Error("NumPyArray: short header", PLANC_EINVAL);
    return h;
  }
  void parse_header(const std::string &h) {
    auto value_after = [&](const std::string &key) -> size_t {
      size_t p = h.find("'" + key + "'");
      if (p == std::string::npos) throw Error("NumPyArray: header lacks " + key, PLANC_EINVAL);
      p = h.find(':', p);
      if (p == std::string::npos) throw Error("NumPyArray: bad header", PLANC_EINVAL);
      return h.find_first_not_of(" ", p + 1);
    };
    size_t p = value_after("descr");
    const std::string descr = h.substr(p + 1, h.find('\'', p + 1) - p - 1);
    if (descr.size() < 3 || (descr[0] != '<' && descr[0] != '|'))
      throw Error("NumPyArray: only little-endian data (descr " + descr + ")", PLANC_EINVAL);
    m_word_size = std::atoi(descr.c_str() + 2);
    if (descr[1] != 'f' || m_word_size != 8)
      throw Error("NumPyArray: only 8-byte floats are read (descr " + descr + ")", PLANC_EINVAL);
    p = value_after("fortran_order");
    m_fortran_order = h.compare(p, 4, "True") == 0;
    p = value_after("shape");
    const size_t e = h.find(')', p);
    std::string shp = h.substr(p + 1, e - p - 1);
    m_dims.clear();
    std::stringstream ss(shp);
    std::string tok;
    while (std::getline(ss, tok, ',')) {
      size_t b = tok.find_first_not_of(" ");
      if (b == std::string::npos) continue;
      m_dims.push_back(std::stoull(tok.substr(b)));
    }
    m_modes = (int64_t)m_dims.size();
    if (m_modes == 0) throw Error("NumPyArray: scalar arrays are not tensors", PLANC_EINVAL);
  }
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

/// Alternating-update NTF (ntf/auntf.hpp:40-251).  update(mode) is the
/// plugin point, pure virtual as in the reference: it returns the new
/// factor of `mode` transposed (k x I_n) from gram_without_one (the Hadamard
/// product of the other modes' Grams) and ncp_mttkrp_t[mode].
///
/// computeNTF() of a class derived from AUNTF runs the reference's loop with
/// the tensor resident on the device: per mode the Grams on the host, the
/// MTTKRP on the device (dimension tree with the reference's reuse when
/// dim_tree(true), else the direct form), update(mode) on the host, then
/// normalize(mode); the relative error after each iteration when
/// compute_error(true) (the fused CP-loss kernel), accelerate() when
/// accelerated(true).  The built-in rules (NTFHALS, NTFMU, NTFANLSBPP,
/// NTFAOADMM) implement update(mode) with single-mode device kernels and, by
/// default, run the whole alternating loop on the device (planc_ntf_run);
/// use_device_loop(false) sends them through the host loop as well.
class AUNTF {
 protected:
  NCPFactors m_ncp_factors;
  std::vector<MAT> ncp_mttkrp_t;  /// k x I_n MTTKRP of each mode
  MAT gram_without_one;
  virtual MAT update(const int mode) = 0;
  /// Device algorithm of a built-in rule, or -1 (host update(mode) only).
  virtual int device_algo() const { return -1; }
  /// Called after each iteration when accelerated(true) (ntf/auntf.hpp:171).
  virtual void accelerate() {}

 private:
  const Tensor &m_input_tensor;
  int m_num_it = 20;
  int m_current_it = 0;
  bool m_compute_error = false;
  bool m_enable_dim_tree = false;
  bool m_accelerated = false;
  bool m_device_loop = true;
  const int m_low_rank_k;
  const algotype m_updalgo;
  double m_rel_error = 0;
  double m_normX2 = 0;
  std::vector<double> m_relerr;
  int m_precision = default_precision();
  // Device copy of the tensor for the host-update loop (uploaded once per
  // computeNTF) and per-mode factor buffers.
  std::unique_ptr<DeviceBuffer<float>> m_dX;
  std::vector<std::unique_ptr<DeviceBuffer<float>>> m_dF;

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
    d.normX2 = m_normX2;
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

  // ---- host-update loop ------------------------------------------------
  void upload_factor(int i) {
    const MAT &F = m_ncp_factors.factor(i);
    upload_f32(F.memptr(), F.n_elem(), *m_dF[i]);
  }
  // MTTKRP of mode j into ncp_mttkrp_t[j] (k x I_j).  With the dimension
  // tree (split s = N/2 - as the shared-memory AUNTF, ntf/auntf.hpp:123 -
  // or N-2 for N = 2): the right contraction at j = 0 and the left one at
  // j = s + 1, reused by the multi-TTVs of the modes of their block
  // (dimtree/ddt.hpp:134-243).
  void device_mttkrp(int j, DeviceBuffer<float> *Pbuf, DeviceBuffer<float> *Qbuf) {
    Device &dev = Device::get();
    const int N = m_input_tensor.modes(), k = m_low_rank_k;
    std::vector<long> dims(N);
    std::vector<const float *> ptrs(N);
    for (int i = 0; i < N; ++i) {
      dims[i] = m_input_tensor.dimension(i);
      ptrs[i] = m_dF[i]->get();
    }
    DeviceBuffer<float> dout((size_t)k * dims[j]);
    if (m_enable_dim_tree && N >= 3) {
      const int s = N / 2 < N - 1 ? N / 2 : N - 2;
      long L = 1, R = 1;
      for (int i = 0; i <= s; ++i) L *= dims[i];
      for (int i = s + 1; i < N; ++i) R *= dims[i];
      if (j == 0)
        dev.check(planc_mttkrp_partial(dev.ctx(), PLANC_F32, m_dX->get(), N, dims.data(), ptrs.data(), k, s, 0,
                                       Pbuf->get(), L),
                  "planc_mttkrp_partial");
      if (j == s + 1)
        dev.check(planc_mttkrp_partial(dev.ctx(), PLANC_F32, m_dX->get(), N, dims.data(), ptrs.data(), k, s, 1,
                                       Qbuf->get(), R),
                  "planc_mttkrp_partial");
      if (j <= s)
        dev.check(planc_multi_ttv(dev.ctx(), Pbuf->get(), L, s + 1, dims.data(), ptrs.data(), j, k,
                                  dout.get(), k),
                  "planc_multi_ttv");
      else
        dev.check(planc_multi_ttv(dev.ctx(), Qbuf->get(), R, N - 1 - s, dims.data() + s + 1,
                                  ptrs.data() + s + 1, j - s - 1, k, dout.get(), k),
                  "planc_multi_ttv");
    } else {
      dev.check(planc_mttkrp(dev.ctx(), PLANC_F32, m_dX->get(), N, dims.data(), ptrs.data(), k, j,
                             dout.get(), k),
                "planc_mttkrp");
    }
    ncp_mttkrp_t[j].zeros(k, dims[j]);
    download_f32(dout, ncp_mttkrp_t[j].memptr(), (size_t)k * dims[j]);
  }

  void run_host_update_loop() {
    const int N = m_input_tensor.modes(), k = m_low_rank_k;
    m_dX.reset(new DeviceBuffer<float>(m_input_tensor.numel()));
    upload_f32(m_input_tensor.m_data.data(), m_input_tensor.numel(), *m_dX);
    m_dF.clear();
    long L = 1, R = 1;
    const int s = N / 2 < N - 1 ? N / 2 : N - 2;
    for (int i = 0; i < N; ++i) {
      m_dF.emplace_back(new DeviceBuffer<float>((size_t)m_input_tensor.dimension(i) * k));
      upload_factor(i);
      if (i <= s)
        L *= m_input_tensor.dimension(i);
      else
        R *= m_input_tensor.dimension(i);
    }
    std::unique_ptr<DeviceBuffer<float>> P, Q;
    if (m_enable_dim_tree && N >= 3) {
      P.reset(new DeviceBuffer<float>((size_t)L * k));
      Q.reset(new DeviceBuffer<float>((size_t)R * k));
    }
    const double normX2 = m_input_tensor.norm();
    m_relerr.assign(m_num_it, 0.0);
    for (m_current_it = 0; m_current_it < m_num_it; ++m_current_it) {
      INFO << "iter::" << m_current_it << std::endl;
      for (int j = 0; j < N; ++j) {
        m_ncp_factors.gram_leave_out_one(j, &gram_without_one);
        device_mttkrp(j, P.get(), Q.get());
        MAT factor = update(j);
        m_ncp_factors.set(j, factor.t());
        m_ncp_factors.normalize(j);
        upload_factor(j);
      }
      if (m_compute_error) {
        m_rel_error = computeObjectiveError(normX2);
        m_relerr[m_current_it] = m_rel_error;
        INFO << "relative_error at it::" << m_current_it << "::" << m_rel_error << std::endl;
      }
      if (m_accelerated) accelerate();
    }
    m_dX.reset();
    m_dF.clear();
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
    // ||X||^2 once, as the reference's constructor takes m_normA (ntf/auntf.hpp:101).
    for (unsigned long long e = 0; e < i_tensor.numel(); ++e) m_normX2 += i_tensor.m_data[e] * i_tensor.m_data[e];
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
  /// Built-in rules only: false runs them through the host-update loop.
  void use_device_loop(bool d) { m_device_loop = d; }
  const std::vector<double> &relative_errors() const { return m_relerr; }
  /// reset (ntf/auntf.hpp:186-198): each mode set and normalised, then
  /// lambda taken from new_factors.
  void reset(const NCPFactors &new_factors, bool trans = false) {
    for (int i = 0; i < m_input_tensor.modes(); ++i) {
      m_ncp_factors.set(i, trans ? new_factors.factor(i).t() : new_factors.factor(i));
      m_ncp_factors.normalize(i);
      if (!m_dF.empty()) upload_factor(i);
    }
    m_ncp_factors.set_lambda(new_factors.lambda());
  }
  /// computeObjectiveError (ntf/auntf.hpp:200-246): sqrt(||X - [[lambda; F]]||^2
  /// / ||X||^2), the residual from the fused CP-loss kernel (lambda folded
  /// into mode 0).
  double computeObjectiveError(double normX2 = -1) {
    Device &dev = Device::get();
    const int N = m_input_tensor.modes(), k = m_low_rank_k;
    if (normX2 < 0) normX2 = m_input_tensor.norm();
    std::unique_ptr<DeviceBuffer<float>> own;
    const float *dX;
    if (m_dX) {
      dX = m_dX->get();
    } else {
      own.reset(new DeviceBuffer<float>(m_input_tensor.numel()));
      upload_f32(m_input_tensor.m_data.data(), m_input_tensor.numel(), *own);
      dX = own->get();
    }
    std::vector<std::unique_ptr<DeviceBuffer<float>>> df;
    std::vector<const float *> ptrs;
    std::vector<long> dims;
    VEC l = m_ncp_factors.lambda();
    for (int i = 0; i < N; ++i) {
      MAT F = m_ncp_factors.factor(i);
      if (i == 0)
        for (int c = 0; c < k; ++c)
          for (unsigned long long r = 0; r < F.n_rows; ++r) F(r, c) *= l[c];
      dims.push_back((long)F.n_rows);
      df.emplace_back(new DeviceBuffer<float>(F.n_elem()));
      upload_f32(F.memptr(), F.n_elem(), *df.back());
      ptrs.push_back(df.back()->get());
    }
    double loss = 0;
    dev.check(planc_cp_loss_fwd(dev.ctx(), dX, N, dims.data(), ptrs.data(), k, &loss), "planc_cp_loss_fwd");
    return std::sqrt(2.0 * loss / normX2);
  }
  void computeNTF() {
    if (device_algo() >= 0 && m_device_loop)
      run_device_loop();
    else
      run_host_update_loop();
  }
};

namespace detail {
// Uploads a factor H (I x k), M (k x I) and S (k x k) for a single-mode
// device update; returns the factor transposed (k x I) after `call`.
template <class F>
inline MAT single_mode(const MAT &H, const MAT &M, const MAT &S, F call) {
  const size_t I = H.n_rows, k = H.n_cols;
  DeviceBuffer<float> dH(I * k), dM(I * k);
  DeviceBuffer<double> dS(k * k);
  upload_f32(H.memptr(), I * k, dH);
  upload_f32(M.memptr(), I * k, dM);
  dS.upload(S.memptr());
  call(dH.get(), dM.get(), dS.get(), (long)I, (int)k);
  MAT out(I, k);
  download_f32(dH, out.memptr(), I * k);
  return out.t();
}
}  // namespace detail

/// NTFHALS::update (ntf/ntfhals.hpp:13-29).
class NTFHALS : public AUNTF {
 protected:
  int device_algo() const { return PLANC_HALS; }
  MAT update(const int mode) {
    Device &dev = Device::get();
    return detail::single_mode(m_ncp_factors.factor(mode), ncp_mttkrp_t[mode], gram_without_one,
                               [&](float *H, const float *M, const double *S, long I, int k) {
                                 dev.check(planc_hals_update_rows(dev.ctx(), H, I, k, I, M, k, 1, S),
                                           "planc_hals_update_rows");
                               });
  }

 public:
  NTFHALS(const Tensor &t, const int k, algotype a) : AUNTF(t, k, a) {}
};

/// NTFMU::update (ntf/ntfmu.hpp:11-16), EPSILON = 1e-6.
class NTFMU : public AUNTF {
 protected:
  int device_algo() const { return PLANC_MU; }
  MAT update(const int mode) {
    Device &dev = Device::get();
    return detail::single_mode(m_ncp_factors.factor(mode), ncp_mttkrp_t[mode], gram_without_one,
                               [&](float *H, const float *M, const double *S, long I, int k) {
                                 dev.check(planc_mu_update(dev.ctx(), H, I, k, I, M, k, 1, S, 1e-6),
                                           "planc_mu_update");
                               });
  }

 public:
  NTFMU(const Tensor &t, const int k, algotype a) : AUNTF(t, k, a) {}
};

/// NTFANLSBPP::update (ntf/ntfanlsbpp.hpp:15-57): BPP on every row.
class NTFANLSBPP : public AUNTF {
 protected:
  int device_algo() const { return PLANC_ANLSBPP; }
  MAT update(const int mode) {
    Device &dev = Device::get();
    return detail::single_mode(m_ncp_factors.factor(mode), ncp_mttkrp_t[mode], gram_without_one,
                               [&](float *H, const float *M, const double *S, long I, int k) {
                                 dev.check(planc_bpp(dev.ctx(), S, k, M, k, 1, I, H, 1, I), "planc_bpp");
                               });
  }

 public:
  NTFANLSBPP(const Tensor &t, const int k, algotype a) : AUNTF(t, k, a) {}
};

/// NTFAOADMM::update (ntf/ntfaoadmm.hpp:23-77): ADMM solves with per-mode
/// dual factors (zero at construction), the dual's columns normalised after
/// each solve.
class NTFAOADMM : public AUNTF {
  std::vector<MAT> m_aux;  // m_ncp_aux, I_n x k per mode

 protected:
  int device_algo() const { return PLANC_AOADMM; }
  MAT update(const int mode) {
    Device &dev = Device::get();
    MAT &D = m_aux[mode];
    const size_t I = D.n_rows, k = D.n_cols;
    DeviceBuffer<float> dD(I * k);
    upload_f32(D.memptr(), I * k, dD);
    MAT out = detail::single_mode(m_ncp_factors.factor(mode), ncp_mttkrp_t[mode], gram_without_one,
                                  [&](float *H, const float *M, const double *S, long In, int kk) {
                                    dev.check(planc_admm_update(dev.ctx(), S, kk, M, kk, 1, H, In, dD.get(),
                                                                In, In),
                                              "planc_admm_update");
                                  });
    download_f32(dD, D.memptr(), I * k);
    for (size_t c = 0; c < k; ++c) {  // m_ncp_aux.normalize(mode)
      double nrm = 0;
      for (size_t r = 0; r < I; ++r) nrm += D(r, c) * D(r, c);
      nrm = std::sqrt(nrm);
      if (nrm > 0)
        for (size_t r = 0; r < I; ++r) D(r, c) /= nrm;
    }
    return out;
  }

 public:
  NTFAOADMM(const Tensor &t, const int k, algotype a) : AUNTF(t, k, a) {
    for (int i = 0; i < t.modes(); ++i) m_aux.emplace_back(t.dimension(i), k);
  }
};

}  // namespace planc

#endif  // PLANC_AMD_HOST_NTF_HPP_
