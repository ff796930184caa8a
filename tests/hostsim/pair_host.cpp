// pair_host.cpp -- test-only host build of the library's per-pair device code
// (bia_3d_gpu_amd/csrc/bia_pair.hpp), so its control flow can be checked against the oracle on
// the CPU.  Never loaded by the package; built by __graft_entry__.build() into tests/hostsim/.
#include "../../bia_3d_gpu_amd/csrc/bia_pair.hpp"

extern "C" void hs_pairs(int method, int64_t n, double eps, double dt, int itr, const double* rel, double* out,
                         int* branch, int* iters, unsigned* flags, int mode, const float* table, int tn) {
  double hist[bia::kMaxItr];
  const bia::LearnTable T{table, tn};
  for (int64_t k = 0; k < n; k++) {
    bia::Hist h{hist, 1};
    int br = 0, it = 0;
    unsigned fl = 0;
    if (method == 0)
      bia::pair_plain<0>(eps, dt, itr, rel + 6 * k, out + 6 * k, &br, &it, &fl, h);
    else if (mode == 1)
      bia::pair_eval<1, bia::kLearn>(eps, dt, itr, rel + 6 * k, out + 6 * k, &br, &it, &fl, h, T);
    else if (mode == 2)
      bia::pair_eval<1, bia::kGuess>(eps, dt, itr, rel + 6 * k, out + 6 * k, &br, &it, &fl, h, T);
    else if (mode == 3)
      bia::pair_eval<1, bia::kPlainLearn>(eps, dt, itr, rel + 6 * k, out + 6 * k, &br, &it, &fl, h, T);
    else
      bia::pair_plain<1>(eps, dt, itr, rel + 6 * k, out + 6 * k, &br, &it, &fl, h);
    branch[k] = br;
    iters[k] = it;
    flags[k] = fl;
  }
}

// the plain mode's range-specific sin/cos (|x| <= pi) and log (positive normal x)
extern "C" void hs_sincos_pi(int64_t n, const double* x, double* s, double* c) {
  for (int64_t k = 0; k < n; k++) bia::sincos_pi(x[k], s + k, c + k);
}
extern "C" void hs_log_pos(int64_t n, const double* x, double* y) {
  for (int64_t k = 0; k < n; k++) y[k] = bia::log_pos(x[k]);
}
