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
extern "C" void hs_atan2_upper(int64_t n, const double* y, const double* x, double* out) {
  for (int64_t k = 0; k < n; k++) out[k] = bia::atan2_upper(y[k], x[k]);
}
extern "C" void hs_acos_unit(int64_t n, const double* q, double* out) {
  for (int64_t k = 0; k < n; k++) out[k] = bia::acos_unit(q[k]);
}
extern "C" void hs_log_pos(int64_t n, const double* x, double* y) {
  for (int64_t k = 0; k < n; k++) y[k] = bia::log_pos(x[k]);
}

// The reference's best-iterate loop (bia/find_dXYUV_kernel.cu:186-217) as it stands, every
// iteration evaluated, beside the library's hyp_loop (which takes the exact-repeat exit without
// the evaluation when the anomaly repeats one of the last two evaluated): both from the same
// set-up and guess selection of each hyperbolic pair.  th_rep[k] = 1 where the reference loop's
// exact-repeat exit came on an anomaly that repeats one of the last two.
static double ref_hyp_loop(const bia::HypState& S, double th_end, double err, int itr, int* ir_out, unsigned* fl,
                           double* resid, int* th_rep) {
  double hist[bia::kMaxItr], d;
  if (itr > bia::kMaxItr) itr = bia::kMaxItr;
  int ir = 0;
  double err_stock = err, th_stock = 0, p1 = NAN, p2 = NAN;
  bool updated = false, cycled = false;
  *th_rep = 0;
  while (ir < itr) {
    const double th_now = th_end;
    ir++;
    const bool rep = th_now == p1 || th_now == p2;
    p2 = p1;
    p1 = th_now;
    th_end = bia::hyp_newton(th_now, S.TE, S.P, &err, &d);
    for (int num = ir - 1; num >= 1; num--)
      if (hist[num - 1] == err) {
        cycled = true;
        break;
      }
    if (cycled) {
      *th_rep = rep;
      break;
    }
    hist[ir - 1] = err;
    if (fabs(err) < fabs(err_stock)) {
      err_stock = err;
      th_stock = th_end;
      updated = true;
    }
    if (fabs(err_stock) < bia::kConvH) break;
  }
  if (cycled) *fl |= bia::kCycle;
  else if (!(fabs(err_stock) < bia::kConvH)) *fl |= bia::kItrCap;
  if (!updated) *fl |= bia::kStockDefault;
  *ir_out = ir;
  *resid = err_stock;
  return th_stock;
}

// out[k] = {theta, residual, iterations, flags} of the library loop (lib) and the reference loop
// (ref); hyp[k] = 0 for pairs that are not hyperbolic (left unset otherwise)
extern "C" void hs_hyp_loops(int64_t n, double eps, double dt, int itr, const double* rel, double* lib, double* ref,
                             int* hyp, int* th_rep) {
  double hbuf[bia::kMaxItr];
  for (int64_t k = 0; k < n; k++) {
    const double* r = rel + 6 * k;
    const double R[3] = {r[0], r[1], r[2]}, G[3] = {r[3], r[4], r[5]};
    double exi[3], eeta[3], yz[4];
    bia::frame(R, G, exi, eeta, yz);
    const double rr = sqrt(yz[0] * yz[0] + yz[1] * yz[1]);
    hyp[k] = rr * (yz[2] * yz[2] + yz[3] * yz[3]) + 2 * eps > 0;
    th_rep[k] = 0;
    if (!hyp[k]) continue;
    bia::HypState S;
    unsigned f0 = 0;
    bia::hyp_setup(eps, dt, yz, S, &f0);
    double e_ini, e_gss, d;
    const double a = bia::hyp_newton(S.th_ini, S.TE, S.P, &e_ini, &d);
    const double b = bia::hyp_newton(S.th_gss, S.TE, S.P, &e_gss, &d);
    const bool ini = fabs(e_gss) > 10 * fabs(e_ini);
    int ir1 = 0, ir2 = 0;
    unsigned f1 = 0, f2 = 0;
    double res1 = 0, res2 = 0;
    const double t1 = bia::hyp_loop(S, ini ? a : b, ini ? e_ini : e_gss, itr, bia::Hist{hbuf, 1}, &ir1, &f1, &res1);
    const double t2 = ref_hyp_loop(S, ini ? a : b, ini ? e_ini : e_gss, itr, &ir2, &f2, &res2, th_rep + k);
    double* o1 = lib + 4 * k;
    double* o2 = ref + 4 * k;
    o1[0] = t1, o1[1] = res1, o1[2] = ir1, o1[3] = f1;
    o2[0] = t2, o2[1] = res2, o2[2] = ir2, o2[3] = f2;
  }
}
