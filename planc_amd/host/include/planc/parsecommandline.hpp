// ParseCommandLine with PLANC's flags, defaults and accessors
// (common/parsecommandline.h:81-108, common/parsecommandline.hpp:126-300),
// for the dense_nmf / dense_ntf drivers.  Added: --precision {fp32,bf16,fp16}
// selects the arithmetic of the bulk products.
#ifndef PLANC_AMD_HOST_PARSECOMMANDLINE_HPP_
#define PLANC_AMD_HOST_PARSECOMMANDLINE_HPP_

#include <getopt.h>

#include "core.hpp"

namespace planc {

// Input normalisation of dense_nmf (normtype, common/utils.h).
enum normtype { NONE, L2NORM, MAXNORM };

enum { NUMKBLOCKS = 1000, NORMALIZATION, DIMTREE, SYMMETRICREG, ADJRAND, NUMLUCITERS, INITSEED,
       PRECISION, DRYRUN };

class ParseCommandLine {
  int m_argc;
  char **m_argv;
  algotype m_lucalgo = ANLSBPP;
  bool m_compute_error = false;
  int m_num_it = 20;
  int m_num_k_blocks = 1;
  bool m_dim_tree = true;
  bool m_adj_rand = false;
  std::string m_Afile_name, m_outputfile_name;
  unsigned long long m_k = 20, m_globalm = 0, m_globaln = 0;
  int m_initseed = 193957;  // reference default (parsecommandline.hpp:153)
  FVEC m_regW{0.0f, 0.0f}, m_regH{0.0f, 0.0f};
  float m_sparsity = 0.01f;
  int m_pr = 1, m_pc = 1;
  double m_symm_reg = -1, m_tolerance = -1;
  int m_num_modes = 0;
  UVEC m_dimensions;
  UVEC m_proc_grids;
  FVEC m_regularizers;
  int m_max_luciters = -1;
  int m_num_nodes = 0;
  int m_precision = PLANC_F32;
  bool m_dry_run = false;
  normtype m_input_normalization = NONE;

  void parse_array(char opt, const char *input) {
    std::vector<std::string> toks;
    std::stringstream ss(input);
    std::string s;
    while (std::getline(ss, s, ' '))
      if (!s.empty()) toks.push_back(s);
    if (m_num_modes == 0) {
      m_num_modes = (int)toks.size();
      m_dimensions.assign(m_num_modes, 0);
      m_regularizers.assign(2 * m_num_modes, 0.0f);
      m_proc_grids.assign(m_num_modes, 1);
    }
    for (size_t i = 0; i < toks.size(); ++i) {
      if (opt == 'd' && i < m_dimensions.size()) m_dimensions[i] = std::stoull(toks[i]);
      if (opt == 'r' && i < m_regularizers.size()) m_regularizers[i] = std::stof(toks[i]);
      if (opt == 'p' && i < m_proc_grids.size()) m_proc_grids[i] = std::stoull(toks[i]);
    }
  }

 public:
  ParseCommandLine(int argc, char **argv) : m_argc(argc), m_argv(argv) {}

  static void print_usage() {
    INFO << "usage: -a algo (0 MU, 1 HALS, 2 ANLSBPP) -d \"dims\" -k rank -t iterations\n"
            "       -i input (rand_uniform | rand_normal | rand_lowrank | file) -o output prefix\n"
            "       -e 1 (report error) -r \"regs\" --dimtree 0/1 --seed N\n"
            "       --precision fp32|bf16|fp16" << std::endl;
  }

  void parseplancopts() {
    static struct option opts[] = {{"algo", required_argument, 0, 'a'},
                                   {"dimensions", required_argument, 0, 'd'},
                                   {"error", required_argument, 0, 'e'},
                                   {"input", required_argument, 0, 'i'},
                                   {"lowrank", required_argument, 0, 'k'},
                                   {"output", required_argument, 0, 'o'},
                                   {"processors", required_argument, 0, 'p'},
                                   {"regularizer", required_argument, 0, 'r'},
                                   {"sparsity", required_argument, 0, 's'},
                                   {"iter", required_argument, 0, 't'},
                                   {"nodes", required_argument, 0, 'n'},
                                   {"tolerance", required_argument, 0, 'l'},
                                   {"numkblocks", required_argument, 0, NUMKBLOCKS},
                                   {"normalization", required_argument, 0, NORMALIZATION},
                                   {"dimtree", required_argument, 0, DIMTREE},
                                   {"symm", required_argument, 0, SYMMETRICREG},
                                   {"adjrand", no_argument, 0, ADJRAND},
                                   {"luciters", required_argument, 0, NUMLUCITERS},
                                   {"seed", required_argument, 0, INITSEED},
                                   {"precision", required_argument, 0, PRECISION},
                                   {"dry-run", no_argument, 0, DRYRUN},
                                   {0, 0, 0, 0}};
    int opt, idx;
    optind = 1;
    while ((opt = getopt_long(m_argc, m_argv, "a:d:e:i:c:k:o:p:r:s:t:h:l:n:", opts, &idx)) != -1) {
      switch (opt) {
        case 'a': m_lucalgo = static_cast<algotype>(atoi(optarg)); break;
        case 'e': m_compute_error = atoi(optarg); break;
        case 'l': m_tolerance = atof(optarg); break;
        case 'i': m_Afile_name = optarg; break;
        case 'k': m_k = atoi(optarg); break;
        case 'o': m_outputfile_name = optarg; break;
        case 'd':
        case 'r':
        case 'p': parse_array((char)opt, optarg); break;
        case 's': m_sparsity = atof(optarg); break;
        case 't': m_num_it = atoi(optarg); break;
        case 'n': m_num_nodes = atoi(optarg); break;
        case NUMKBLOCKS: m_num_k_blocks = atoi(optarg); break;
        case NORMALIZATION: {  // parsecommandline.hpp:210-217
          std::string t = optarg;
          if (t == "l2")
            m_input_normalization = L2NORM;
          else if (t == "max")
            m_input_normalization = MAXNORM;
          break;
        }
        case DIMTREE: m_dim_tree = atoi(optarg); break;
        case SYMMETRICREG: m_symm_reg = atof(optarg); break;
        case ADJRAND: m_adj_rand = true; break;
        case NUMLUCITERS: m_max_luciters = atoi(optarg); break;
        case INITSEED: m_initseed = atoi(optarg); break;
        case PRECISION: {
          std::string p = optarg;
          if (p == "fp32") m_precision = PLANC_F32;
          else if (p == "bf16") m_precision = PLANC_BF16;
          else if (p == "fp16") m_precision = PLANC_F16;
          else throw Error("--precision must be fp32, bf16 or fp16", PLANC_EINVAL);
          break;
        }
        case DRYRUN: m_dry_run = true; break;
        case 'h': print_usage(); exit(0);
        default: print_usage(); throw Error("failed while processing arguments", PLANC_EINVAL);
      }
    }
    if (m_Afile_name.empty()) {
      print_usage();
      throw Error("Input not given.", PLANC_EINVAL);
    }
    if (m_num_modes == 2) {  // parsecommandline.hpp:288-299
      m_globalm = m_dimensions[0];
      m_globaln = m_dimensions[1];
      m_regW = {m_regularizers[0], m_regularizers[1]};
      m_regH = {m_regularizers[2], m_regularizers[3]};
      m_pr = (int)m_proc_grids[0];
      m_pc = (int)m_proc_grids[1];
    }
  }

  void printConfig() const {
    INFO << "a::" << m_lucalgo << "::i::" << m_Afile_name << "::k::" << m_k << "::m::" << m_globalm
         << "::n::" << m_globaln << "::t::" << m_num_it << "::error::" << m_compute_error
         << "::dimtree::" << m_dim_tree << "::input normalization::" << m_input_normalization
         << "::adjrand::" << m_adj_rand << "::initseed::" << m_initseed << "::precision::"
         << (m_precision == PLANC_F32 ? "fp32" : m_precision == PLANC_BF16 ? "bf16" : "fp16")
         << "::regW::" << m_regW[0] << "," << m_regW[1] << "::regH::" << m_regH[0] << ","
         << m_regH[1] << std::endl;
  }

  algotype lucalgo() const { return m_lucalgo; }
  bool compute_error() const { return m_compute_error; }
  int iterations() const { return m_num_it; }
  unsigned long long lowrankk() const { return m_k; }
  unsigned long long globalm() const { return m_globalm; }
  unsigned long long globaln() const { return m_globaln; }
  std::string input_file_name() const { return m_Afile_name; }
  std::string output_file_name() const { return m_outputfile_name; }
  FVEC regW() const { return m_regW; }
  FVEC regH() const { return m_regH; }
  float sparsity() const { return m_sparsity; }
  bool dim_tree() const { return m_dim_tree; }
  double symm_reg() const { return m_symm_reg; }
  double tolerance() const { return m_tolerance; }
  int initseed() const { return m_initseed; }
  int num_modes() const { return m_num_modes; }
  UVEC dimensions() const { return m_dimensions; }
  int num_k_blocks() const { return m_num_k_blocks; }
  int nodes() const { return m_num_nodes; }
  int max_luciters() const { return m_max_luciters; }
  bool adj_rand() const { return m_adj_rand; }
  normtype input_normalization() const { return m_input_normalization; }
  int precision() const { return m_precision; }
  bool dry_run() const { return m_dry_run; }
};

}  // namespace planc

#endif  // PLANC_AMD_HOST_PARSECOMMANDLINE_HPP_
