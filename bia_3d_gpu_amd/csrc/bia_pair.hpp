// bia_pair.hpp -- per-pair two-body propagators of the Binary Interaction Approximation.
//
// One pair (i, j) is propagated over dt as an isolated two-body problem
//   r'' = eps * r / |r|^3        (eps = -2Gm for gravity, +/-2kq^2/m for Coulomb)
// in its own orbital plane, then mapped back to 3D with the equal-mass half split.
//
// Branches (reference: bia/find_dXYUV_kernel.cu):
//   hyperbola  Newton in the true anomaly on the hyperbolic time equation  (:94-238)
//   ellipse    Newton on Kepler's equation in the eccentric anomaly        (:242-364)
//   parabola   Barker's equation, Cardano guess + Newton polish           (:367-465)
//   RKF45      adaptive Fehlberg 4(5) with step doubling on (xi,eta,g)    (bia/bia-rkf.cu:1-103)
//
// The reference iteration control (best iterate, exact-repeat exit, ITR cap, tolerance
// 2.22e-16 on the dimensionless time residual) is kept in kPlain mode.  The hyperbola's
// Newton residual is evaluated from one sincos of the half angle:
//   sin((t0 +- t)/2) = sin(t0/2)cos(t/2) +- cos(t0/2)sin(t/2), cos t = (c-s)(c+s), sin t = 2sc
// which is the reference's residual up to rounding, at a third of the transcendental cost.
//
// Everything here is __host__ __device__ so the same source can be exercised on the CPU.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define BIA_HD __host__ __device__ inline
#define BIA_HDF __host__ __device__ __forceinline__
#define BIA_HDN __host__ __device__ inline __attribute__((noinline))  // rare branches, kept out of line
#else
#define BIA_HD inline
#define BIA_HDF inline
#define BIA_HDN inline
#endif

namespace bia {

constexpr double kPi = 3.14159265358979323846264338327950288419716939937510582097494;
constexpr double kConvH = 2.22e-16;  // find_dXYUV_kernel.cu:18
constexpr double kConvE = 2e-16;     // find_dXYUV_kernel.cu:284
constexpr int kMaxItr = 20;          // size of the reference's residual history (:192)

enum Branch : int { kHyperbola = 0, kParabola = 1, kEllipse = 2, kRkf = 3 };

// per-pair status bits
enum Flag : unsigned {
  kStockDefault = 1u,  // best-iterate slot never updated: the reference returns its initial value
  kCycle = 2u,         // exact repeat of an earlier residual ended the loop
  kItrCap = 4u,        // iteration cap reached without |residual| < tolerance
  kAcosClamp = 8u,     // |cos(theta_ini)| > 1 clamp taken (hyperbola)
  kRkfFail = 16u,      // per-pair RKF ended before reaching dt
  kNonFinite = 32u,    // non-finite output
  kLearned = 64u,      // learned initial guess accepted
  kFallback = 128u     // learned path fell back to the reference guess
};

// kPlain: the reference's guess selection and iteration control.
// kLearn: learned guess (network tabulated over its input features) + safeguarded Newton.
// kGuess: the reference's closed-form guess + the same safeguarded Newton (no network).
enum Mode : int { kPlain = 0, kLearn = 1, kGuess = 2 };

// Learned-initialiser table: the network's output t(x1, v) for the two hyperbolic branches on a
// regular grid, x1 in [-1, 1] (nx points), v in [0, 1] (nv points), bilinear interpolation.
struct LearnTable {
  const float* t;  // [2][nv][nx]
  int nx, nv;
};

// Residual history used by the reference's exact-repeat exit.  On the device it lives in
// LDS (one column per lane); on the host it is a plain array.
struct Hist {
  double* p;
  int stride;
  BIA_HDF double get(int k) const { return p[k * stride]; }
  BIA_HDF void set(int k, double v) const { p[k * stride] = v; }
};

BIA_HDF void sincos_(double x, double* s, double* c) { ::sincos(x, s, c); }

// ------------------------------------------------------------------------------------
// orbital-plane frame, find_dXYUV_kernel.cu:467-488:
//   e_eta = g/|g|, e_zeta = (r x g)/|r x g|, e_xi = e_eta x e_zeta,
//   yz = (r.e_xi, r.e_eta, 0, |g|)
// ------------------------------------------------------------------------------------
BIA_HDF void frame(const double R[3], const double G[3], double exi[3], double eeta[3], double yz[4]) {
  double gm = sqrt(G[0] * G[0] + G[1] * G[1] + G[2] * G[2]);
  eeta[0] = G[0] / gm;
  eeta[1] = G[1] / gm;
  eeta[2] = G[2] / gm;
  double L0 = R[1] * G[2] - R[2] * G[1];
  double L1 = R[2] * G[0] - R[0] * G[2];
  double L2 = R[0] * G[1] - R[1] * G[0];
  double Lm = sqrt(L0 * L0 + L1 * L1 + L2 * L2);
  double z0 = L0 / Lm, z1 = L1 / Lm, z2 = L2 / Lm;
  exi[0] = eeta[1] * z2 - eeta[2] * z1;
  exi[1] = eeta[2] * z0 - eeta[0] * z2;
  exi[2] = eeta[0] * z1 - eeta[1] * z0;
  yz[0] = R[0] * exi[0] + R[1] * exi[1] + R[2] * exi[2];
  yz[1] = R[0] * eeta[0] + R[1] * eeta[1] + R[2] * eeta[2];
  yz[2] = 0;
  yz[3] = gm;
}

// ------------------------------------------------------------------------------------
// hyperbola
// ------------------------------------------------------------------------------------
struct HypP {
  double th0, sn0, cs0, tn0, s0h, c0h;  // theta_0, sin/cos/tan theta_0 as the reference forms them,
                                        // sin/cos(theta_0 / 2)
};

// residual of the dimensionless time equation and one Newton update (find_theta_end_cu :57-66)
//   f(t) = tn0 sin t / (cos t - cs0) + ln( sin((t0+t)/2) / sin((t0-t)/2) ) - TE
BIA_HDF double hyp_newton(double th, double TE, const HypP& P, double* err, double* dfd) {
  double sh, ch;
  sincos_(0.5 * th, &sh, &ch);
  double snp = P.s0h * ch + P.c0h * sh;
  double snm = P.s0h * ch - P.c0h * sh;
  double cs = (ch - sh) * (ch + sh);
  double sn = 2 * sh * ch;
  double Dinv = 1 / (cs - P.cs0);
  double d = P.tn0 * (1 - P.cs0 * cs) * Dinv * Dinv + P.sn0 / (2 * snp * snm);
  double e = P.tn0 * sn * Dinv - TE + log(snp / snm);
  *err = e;
  *dfd = d;
  return th - e / d;
}

struct HypState {
  HypP P;
  double TE;        // dimensionless target time
  double th_ini;    // initial true anomaly
  double b, g0_sn0, u, v, cs_a, sn_a;
  double th_gss;    // the reference's closed-form guess (:172)
};

// set-up shared by both modes: find_dXYUV_kernel.cu:98-172
BIA_HDF void hyp_setup(double eps, double dt, const double yz[4], HypState& S, unsigned* fl) {
  double X = yz[0], Y = yz[1], U = yz[2], V = yz[3];
  double r = sqrt(X * X + Y * Y);
  double W = U * U + V * V + 2 * eps / r;
  double r_g = X * U + Y * V;
  double L = X * V - Y * U;
  double g0 = sqrt(W);
  double b0 = eps / W;
  double aL = fabs(L);
  S.b = aL / g0;
  double bg = b0 * g0;
  S.P.th0 = atan2(aL, bg);
  double den = sqrt(bg * bg + L * L);
  S.P.sn0 = aL / den;
  S.P.cs0 = bg / den;
  S.P.tn0 = S.P.sn0 / S.P.cs0;
  sincos_(0.5 * S.P.th0, &S.P.s0h, &S.P.c0h);
  double q = S.P.cs0 + S.b * S.P.sn0 / r;
  double th_ini, cs;
  if (fabs(q) > 1) {  // :119-126
    th_ini = 1e-30;
    cs = cos(th_ini);
    *fl |= kAcosClamp;
  } else {
    th_ini = acos(q);
    cs = q;
  }
  if (L * r_g < 0) th_ini = -th_ini;
  double sh, ch;
  sincos_(0.5 * th_ini, &sh, &ch);
  double sn = 2 * sh * ch;
  double x = r * cs, y = r * sn;
  S.g0_sn0 = (L > 0) ? g0 / S.P.sn0 : -g0 / S.P.sn0;
  S.u = S.g0_sn0 * (S.P.cs0 * sn);
  S.v = S.g0_sn0 * (1 - S.P.cs0 * cs);
  double nrm = sqrt((X * X + Y * Y) * (x * x + y * y));
  S.cs_a = (x * X + y * Y) / nrm;
  S.sn_a = (x * Y - y * X) / nrm;
  double snp = S.P.s0h * ch + S.P.c0h * sh;
  double snm = S.P.s0h * ch - S.P.c0h * sh;
  double time_ini = (S.P.tn0 * sn / (cs - S.P.cs0) + log(snp / snm)) * b0 / g0;  // :153-155
  double time_end = (L > 0) ? time_ini + dt : time_ini - dt;
  S.TE = time_end * g0 / b0;
  S.th_ini = th_ini;
  S.th_gss = 2 / kPi * S.P.th0 * atan(S.TE / S.P.tn0);
}

// state after dt, rotated back to the (xi, eta) frame: :220-237
BIA_HDF void hyp_finish(const HypState& S, double th_end, const double yz[4], double yn[4]) {
  double sn_end, cs_end;
  sincos_(th_end, &sn_end, &cs_end);
  double u_end = S.g0_sn0 * S.P.cs0 * sn_end;
  double v_end = S.g0_sn0 * (1 - S.P.cs0 * cs_end);
  double du = u_end - S.u, dv = v_end - S.v;
  double r_end = S.b * S.P.sn0 / (cs_end - S.P.cs0);
  double x_end = r_end * cs_end, y_end = r_end * sn_end;
  yn[0] = x_end * S.cs_a - y_end * S.sn_a;
  yn[1] = x_end * S.sn_a + y_end * S.cs_a;
  yn[2] = yz[2] + (du * S.cs_a - dv * S.sn_a);
  yn[3] = yz[3] + (du * S.sn_a + dv * S.cs_a);
}

// the reference's guess selection and best-iterate loop: :169-217
BIA_HDF double hyp_solve_plain(const HypState& S, int itr, Hist hist, int* iters, unsigned* fl) {
  double e_ini, e_gss, d;
  double th_ini_end = hyp_newton(S.th_ini, S.TE, S.P, &e_ini, &d);
  double th_gss_end = hyp_newton(S.th_gss, S.TE, S.P, &e_gss, &d);
  double th_end, err;
  if (fabs(e_gss) > 10 * fabs(e_ini)) {
    th_end = th_ini_end;
    err = e_ini;
  } else {
    th_end = th_gss_end;
    err = e_gss;
  }
  if (itr > kMaxItr) itr = kMaxItr;
  int ir = 0;
  double err_stock = err, th_stock = 0;
  bool updated = false, cycled = false;
  while (ir < itr) {
    double th_now = th_end;
    ir++;
    th_end = hyp_newton(th_now, S.TE, S.P, &err, &d);
    for (int num = ir - 1; num >= 1; num--) {
      if (hist.get(num - 1) == err) {
        cycled = true;
        break;
      }
    }
    if (cycled) break;
    hist.set(ir - 1, err);
    if (fabs(err) < fabs(err_stock)) {
      err_stock = err;
      th_stock = th_end;
      updated = true;
    }
    if (fabs(err_stock) < kConvH) break;
  }
  if (cycled) *fl |= kCycle;
  else if (!(fabs(err_stock) < kConvH)) *fl |= kItrCap;
  if (!updated) *fl |= kStockDefault;
  *iters = ir + 2;  // two set-up Newton evaluations + the loop
  return th_stock;
}

// ------------------------------------------------------------------------------------
// learned initial guess and safeguarded Newton (no reference counterpart)
//
// Features of a hyperbolic pair: x1 = theta_gss / theta_0 (the reference's closed-form guess
// (:172), a function of the dimensionless mean anomaly TE and e), v = e^(-1/4) = |cs0|^(1/4),
// branch b = 0 attractive (cs0 < 0) / 1 repulsive.  Target t = (theta/theta_0 - x1) * e, the
// guess's scaled residual, O(1) and smooth; theta_L = theta_0 (x1 + t / e).
// ------------------------------------------------------------------------------------
BIA_HDF void hyp_features(const HypState& S, double* x1, double* v, int* b) {
  *x1 = S.th_gss / S.P.th0;
  *v = sqrt(sqrt(fabs(S.P.cs0)));
  *b = (S.P.cs0 < 0) ? 0 : 1;
}

BIA_HDF double table_lookup(const LearnTable& T, int b, double x1, double v) {
  double fx = (x1 + 1) * 0.5 * (T.nx - 1);
  double fv = v * (T.nv - 1);
  fx = fmin(fmax(fx, 0.0), T.nx - 1.000001);
  fv = fmin(fmax(fv, 0.0), T.nv - 1.000001);
  const int ix = (int)fx, iv = (int)fv;
  const double ax = fx - ix, av = fv - iv;
  const float* p = T.t + ((size_t)b * T.nv + iv) * T.nx + ix;
  const double t00 = p[0], t01 = p[1], t10 = p[T.nx], t11 = p[T.nx + 1];
  return (1 - av) * ((1 - ax) * t00 + ax * t01) + av * ((1 - ax) * t10 + ax * t11);
}

// Newton on the reference's residual, kept inside a bracket of (-theta_0, theta_0): the residual
// is monotone there, so each evaluation moves one end of the bracket, and a step leaving the
// bracket (or a non-finite residual near a pole) is replaced by bisection.  Stops after a Newton
// step |d theta| <= 1e-8 theta_0: quadratic convergence leaves ~1e-16 theta_0 after applying it.
BIA_HDF double hyp_solve_safe(const HypState& S, double th, int kmax, int* evals, bool* ok) {
  double lo = -S.P.th0, hi = S.P.th0;
  if (!(th > lo && th < hi)) th = 0.5 * (lo + hi) + 0.5 * (S.th_gss - 0.5 * (lo + hi));
  *ok = false;
  int k = 0;
  double prev = 0;  // size of the last Newton step taken (0 after a bisection)
  for (; k < kmax; k++) {
    double err, d;
    double th_new = hyp_newton(th, S.TE, S.P, &err, &d);
    if (err == 0) {
      *ok = true;
      *evals = k + 1;
      return th;
    }
    const bool newton_ok = isfinite(err) && isfinite(d) && d != 0 && isfinite(th_new);
    // converged when the error left after applying this step is below 1e-16 theta_0.  Newton's
    // error contracts as e' = C e^2 and each step measures the error it removes, so after two
    // Newton steps of sizes p then s the remaining error is about C s^2 = s^3 / p^2.  The first
    // step has no such estimate and must itself be below 1e-10 theta_0.  (Tested before the
    // bracket update, whose end may sit exactly at th when the step rounds to zero.)
    const double step = fabs(th_new - th);
    if (newton_ok && (k == 0 || prev == 0 ? step <= 1e-10 * S.P.th0
                                          : step * step * step <= 1e-16 * S.P.th0 * prev * prev)) {
      *ok = true;
      *evals = k + 1;
      return th_new;
    }
    if (!newton_ok) {
      if (th > 0) hi = th;  // the only poles sit at the ends of the interval
      else lo = th;
    } else if ((err > 0) == (d > 0)) {
      hi = th;  // residual increasing through zero above th
    } else {
      lo = th;
    }
    const bool take = newton_ok && th_new > lo && th_new < hi;
    th = take ? th_new : 0.5 * (lo + hi);
    prev = take ? step : 0;
  }
  *evals = k;
  return th;
}

// ------------------------------------------------------------------------------------
// ellipse: find_dXYUV_kernel.cu:242-364
// ------------------------------------------------------------------------------------
BIA_HDN void ellipse(double eps, double dt, int itr, const double yz[4], double yn[4], int* iters,
                    unsigned* fl) {
  double X = yz[0], Y = yz[1], U = yz[2], V = yz[3];
  double g2 = U * U + V * V;
  double r_ini = sqrt(X * X + Y * Y);
  double W = g2 + 2 * eps / r_ini;
  double L = X * V - Y * U;
  double r_g = X * U + Y * V;
  double g_a = (L >= 0) ? sqrt(-W) : -sqrt(-W);
  double a = eps / W;
  double b = L / g_a;
  double omega_a = g_a / a;
  double xi_ecc = asin(L / (a * g_a));
  double ecc = cos(xi_ecc);
  double xi_ini;
  if (r_g == 0) {  // apsides, :315-330
    if (r_ini < a) {
      xi_ini = 0;
    } else if (r_ini == a) {
      xi_ini = atan2(Y, X);
      ecc = 0;
    } else {
      xi_ini = (L >= 0) ? kPi : -kPi;
    }
  } else {
    xi_ini = atan2(r_g / g_a, a - r_ini);
  }
  double s_ini, c_ini;
  sincos_(xi_ini, &s_ini, &c_ini);
  double alpha = atan2(Y, X) - atan2(b * s_ini, a * (c_ini - ecc));
  double tau_ini = xi_ini - ecc * s_ini;
  double time_ini = tau_ini / omega_a;
  double time_end = time_ini + dt;
  double tau_end = omega_a * time_end;
  if (dt == 0) {  // the reference leaves YN unwritten here (:340-344); return the unchanged state
    yn[0] = yz[0]; yn[1] = yz[1]; yn[2] = yz[2]; yn[3] = yz[3];
    *iters = 0;
    return;
  }
  // newton_ellipse :242-288
  double xi_end;
  int ir = 0;
  double error_ini = tau_end - xi_ini + ecc * s_ini;
  double delta_ini = error_ini * (1 - ecc * c_ini);  // the reference multiplies here
  if (error_ini == 0) {
    xi_end = xi_ini;
  } else {
    double st, ct;
    sincos_(tau_end, &st, &ct);
    double xi_2 = (ct < 0.9) ? tau_end : ((st >= 0) ? kPi / 2 : -kPi / 2);
    double s2, c2;
    sincos_(xi_2, &s2, &c2);
    double error_tm2 = tau_end - xi_2 + ecc * s2;
    double delta_xi2 = error_tm2 / (1 - ecc * c2);
    double xi = (fabs(error_ini) < fabs(error_tm2)) ? xi_2 + delta_ini : xi_2 + delta_xi2;
    double err_stock = 100, th_stock = 0;
    bool conv = false;
    while (ir <= itr) {
      ir++;
      double xi_1 = xi, s1, c1;
      sincos_(xi_1, &s1, &c1);
      double e = tau_end - xi_1 + ecc * s1;
      xi = xi_1 + e / (1 - ecc * c1);
      if (fabs(e) < fabs(err_stock)) {
        err_stock = e;
        th_stock = xi;
      }
      if (fabs(err_stock) < kConvE) {
        conv = true;
        break;
      }
    }
    if (!conv) *fl |= kItrCap;
    xi_end = th_stock;
  }
  double sn_end, cs_end;
  sincos_(xi_end, &sn_end, &cs_end);
  double xx = a * (cs_end - ecc);
  double yy = b * sn_end;
  double rr = a * (1 - ecc * cs_end);
  double uu = -a * g_a * sn_end / rr;
  double vv = b * g_a * cs_end / rr;
  double sa, ca;
  sincos_(alpha, &sa, &ca);
  yn[0] = xx * ca - yy * sa;
  yn[1] = xx * sa + yy * ca;
  yn[2] = uu * ca - vv * sa;
  yn[3] = uu * sa + vv * ca;
  *iters = ir;
}

// ------------------------------------------------------------------------------------
// parabola (W == 0 exactly): find_dXYUV_kernel.cu:367-465
// ------------------------------------------------------------------------------------
BIA_HDF double par_time(double th, double rmg) {
  double sn, cs;
  sincos_(th, &sn, &cs);
  return (2 + cs) * sn / (1 + cs) / (1 + cs) * 4 * rmg / 3;
}
BIA_HDF double par_dtdth(double th, double rmg) {
  double t = 1 / cos(th / 2);
  t = t * t;
  t = t * t;
  return rmg * t;
}

BIA_HDN void parabola(double eps, double dt, const double yz[4], double yn[4], int* iters) {
  double r_ini = sqrt(yz[0] * yz[0] + yz[1] * yz[1]);
  double r_dot_g = yz[0] * yz[2] + yz[1] * yz[3];
  double r_crs_g = yz[0] * yz[3] - yz[1] * yz[2];
  double g_max = -2 * eps / r_crs_g;
  double r_min = r_crs_g / g_max;
  double rmg = r_min / g_max;
  double x_ini = 2 * r_min - r_ini;
  double tmp = 2 * r_min * sqrt(r_ini / r_min - 1);
  double y_ini = (r_dot_g > 0) ? tmp : -tmp;
  double time_end = par_time(atan2(y_ini, x_ini), rmg) + dt;
  double te = 3 * g_max / (4 * r_min) * time_end;
  double Fp = cbrt(sqrt(te * te + 1) + te);
  double Fm = cbrt(sqrt(te * te + 1) - te);
  // Newton polish with the reference's bookkeeping (:411-440), history kept in registers:
  // the loop only ever looks back at the current and best iterates plus exact repeats,
  // which are tracked by a running list of at most 20 residuals.
  double th[20], et[20];
  th[0] = 2 * atan(Fp - Fm);
  et[0] = par_time(th[0], rmg) - time_end;
  int i_min = 0, i;
  double et_min = fabs(et[0]);
  for (i = 1; i < 20; i++) {
    if (et[i - 1] == 0) {
      i_min = i - 1;
      break;
    }
    th[i] = th[i - 1] - et[i - 1] / par_dtdth(th[i - 1], rmg);
    et[i] = par_time(th[i], rmg) - time_end;
    if (th[i] == th[i - 1]) {
      i_min = i;
      break;
    }
    if (et_min > fabs(et[i])) {
      i_min = i;
      et_min = fabs(et[i]);
    }
    for (int j = 0; j < i; j++) {
      if (et[i] == et[j]) {
        et_min = fabs(et[i_min]);
        i_min = j;
        break;
      }
    }
    if (i != i_min && et[i] == et[i_min]) break;
  }
  double th_end = th[i_min];
  double sb, cb, st, ct;
  sincos_(th_end / 2, &sb, &cb);
  double r_end = r_min / (cb * cb);
  double g_end = g_max * cb;
  sincos_(th_end, &st, &ct);
  double x_end = r_end * ct, y_end = r_end * st;
  double u_end = -g_end * sb, v_end = g_end * cb;
  double gamma = atan2(x_ini * yz[1] - y_ini * yz[0], x_ini * yz[0] + y_ini * yz[1]);
  double sg, cg;
  sincos_(gamma, &sg, &cg);
  yn[0] = x_end * cg - y_end * sg;
  yn[1] = x_end * sg + y_end * cg;
  yn[2] = u_end * cg - v_end * sg;
  yn[3] = u_end * sg + v_end * cg;
  *iters = (i < 20) ? i : 19;
}

// ------------------------------------------------------------------------------------
// per-pair RKF45: bia/bia-rkf.cu:1-103 (RELERR 0, ABSERR 1e-16, ITEMAX 1e5)
// ------------------------------------------------------------------------------------
namespace fehl {
constexpr double B21 = 1e0 / 4, B31 = 3e0 / 32, B32 = 9e0 / 32, B41 = 1932e0 / 2197, B42 = -7200e0 / 2197,
                 B43 = 7296e0 / 2197, B51 = 439e0 / 216, B52 = -8e0, B53 = 3680e0 / 513, B54 = -845e0 / 4104,
                 B61 = -8e0 / 27, B62 = 2e0, B63 = -3544e0 / 2565, B64 = 1859e0 / 4104, B65 = -11e0 / 40,
                 GA1 = 16e0 / 135, GA3 = 6656e0 / 12825, GA4 = 28561e0 / 56430, GA5 = -9e0 / 50,
                 GA6 = 2e0 / 55, DA1 = 1e0 / 360, DA3 = -128e0 / 4275, DA4 = -2197e0 / 75240,
                 DA5 = 1e0 / 50, DA6 = 2e0 / 55;
}

BIA_HDF void kep_f(double eps, const double* y, double* f) {
  double rsq = y[0] * y[0] + y[1] * y[1];
  double aaa = eps / (rsq * sqrt(rsq));
  f[0] = y[2];
  f[1] = y[3];
  f[2] = y[0] * aaa;
  f[3] = y[1] * aaa;
}

template <bool ERR>
BIA_HDF void kep_fehl(double eps, double h, const double* y0, double* yn, double* er) {
  using namespace fehl;
  double k1[4], k2[4], k3[4], k4[4], k5[4], k6[4], w[4];
  kep_f(eps, y0, k1);
  for (int i = 0; i < 4; i++) w[i] = y0[i] + h * (B21 * k1[i]);
  kep_f(eps, w, k2);
  for (int i = 0; i < 4; i++) w[i] = y0[i] + h * (B31 * k1[i] + B32 * k2[i]);
  kep_f(eps, w, k3);
  for (int i = 0; i < 4; i++) w[i] = y0[i] + h * (B41 * k1[i] + B42 * k2[i] + B43 * k3[i]);
  kep_f(eps, w, k4);
  for (int i = 0; i < 4; i++) w[i] = y0[i] + h * (B51 * k1[i] + B52 * k2[i] + B53 * k3[i] + B54 * k4[i]);
  kep_f(eps, w, k5);
  for (int i = 0; i < 4; i++)
    w[i] = y0[i] + h * (B61 * k1[i] + B62 * k2[i] + B63 * k3[i] + B64 * k4[i] + B65 * k5[i]);
  kep_f(eps, w, k6);
  for (int i = 0; i < 4; i++)
    yn[i] = y0[i] + h * (GA1 * k1[i] + GA3 * k3[i] + GA4 * k4[i] + GA5 * k5[i] + GA6 * k6[i]);
  if (ERR)
    for (int i = 0; i < 4; i++) er[i] = h * (DA1 * k1[i] + DA3 * k3[i] + DA4 * k4[i] + DA5 * k5[i] + DA6 * k6[i]);
}

BIA_HD void rkf_pair(double eps, double dt, const double yz[4], double yn[4], int* steps, unsigned* fl,
                     double relerr = 0, double abserr = 1e-16) {
  const double MINLIMIT = 1e-300, EPSMIN = 1e-15;
  const int ITEMAX = 100000;
  double yh0[4], yl0[4], yln[4], er[4], f[4], yhn[4];
  for (int i = 0; i < 4; i++) yh0[i] = yl0[i] = yhn[i] = yz[i];
  double x = 0;
  kep_f(eps, yl0, f);
  double toldy = 1e+75;
  bool changed = false;
  for (int i = 0; i < 4; i++) {
    if (fabs(f[i]) > MINLIMIT) {
      double tol = relerr * fabs(yl0[i]) + abserr;
      toldy = fmin(toldy, fabs(tol / f[i]));
      changed = true;
    }
  }
  double h = changed ? fmin(dt / 100, pow(toldy, 0.2)) : EPSMIN * dt;
  int acc = 0;
  bool done = false;
  for (int iter = 0; iter < ITEMAX; iter++) {
    kep_fehl<true>(eps, h, yl0, yln, er);
    double erret = 0;
    for (int i = 0; i < 4; i++) {
      double et = relerr * (fabs(yl0[i]) + fabs(yln[i])) / 2 + abserr;
      if (et < MINLIMIT) iter = ITEMAX;
      else erret = fmax(erret, fabs(er[i]) / et);
    }
    if (erret > 1) {
      h *= (erret >= 59049) ? 0.1 : 0.9 / pow(erret, 0.2);
      if (h <= EPSMIN) break;
    } else if (x + h < dt) {
      kep_fehl<false>(eps, h / 2, yh0, yhn, er);
      for (int i = 0; i < 4; i++) yh0[i] = yhn[i];
      kep_fehl<false>(eps, h / 2, yh0, yhn, er);
      x += h;
      acc++;
      for (int i = 0; i < 4; i++) {
        yl0[i] = yln[i];
        yh0[i] = yhn[i];
      }
      h *= (erret <= 1.889568e-4) ? 5.0 : 0.9 / pow(erret, 0.2);
    } else {
      h = dt - x;
      kep_fehl<true>(eps, h, yl0, yln, er);
      kep_fehl<false>(eps, h / 2, yh0, yhn, er);
      for (int i = 0; i < 4; i++) yh0[i] = yhn[i];
      kep_fehl<false>(eps, h / 2, yh0, yhn, er);
      acc++;
      done = true;
      break;
    }
  }
  for (int i = 0; i < 4; i++) yn[i] = yhn[i];
  *steps = acc;
  if (!done) *fl |= kRkfFail;
}

// ------------------------------------------------------------------------------------
// one pair: frame, classifier (:503-506), 2D solve, half-split epilogue (:507-515)
// rel = (r_i - r_j, g_i - g_j); d = the six half-deltas credited to i (j gets -d)
// ------------------------------------------------------------------------------------
template <int METHOD, int MODE>  // METHOD: 0 = RKF45, 1 = ALG (reference METHOD codes)
BIA_HDF void pair_eval(double eps, double dt, int itr, const double rel[6], double d[6], int* branch, int* iters,
                       unsigned* fl, Hist hist, const LearnTable& T) {
  double exi[3], eeta[3], yz[4], yn[4];
  frame(rel, rel + 3, exi, eeta, yz);
  int it = 0;
  if (METHOD == 0) {
    *branch = kRkf;
    rkf_pair(eps, dt, yz, yn, &it, fl);
  } else {
    double W = sqrt(yz[0] * yz[0] + yz[1] * yz[1]) * (yz[2] * yz[2] + yz[3] * yz[3]) + 2 * eps;
    if (W > 0) {
      *branch = kHyperbola;
      HypState S;
      hyp_setup(eps, dt, yz, S, fl);
      double th;
      if (MODE == kPlain) {
        th = hyp_solve_plain(S, itr, hist, &it, fl);
      } else {
        double th0 = S.th_gss;
        if (MODE == kLearn) {
          double x1, v;
          int b;
          hyp_features(S, &x1, &v, &b);
          const double tp = table_lookup(T, b, x1, v);
          const double thl = S.P.th0 * (x1 + fabs(S.P.cs0) * tp);
          // the reference guess judges the learned one: a non-finite or out-of-range proposal
          // is replaced by it
          if (isfinite(thl) && fabs(thl) < S.P.th0) {
            th0 = thl;
            *fl |= kLearned;
          }
        }
        bool ok;
        th = hyp_solve_safe(S, th0, 24, &it, &ok);
        if (!ok) {  // fallback: the reference's own guess and iteration control
          int it2 = 0;
          th = hyp_solve_plain(S, itr, hist, &it2, fl);
          it += it2;
          *fl |= kFallback;
        }
      }
      hyp_finish(S, th, yz, yn);
    } else if (W == 0) {
      *branch = kParabola;
      parabola(eps, dt, yz, yn, &it);
    } else {
      *branch = kEllipse;
      ellipse(eps, dt, itr, yz, yn, &it, fl);
    }
  }
  *iters = it;
  double drx = yn[0] - yz[0], dre = yn[1] - yz[1];
  double dgx = yn[2] - yz[2], dge = yn[3] - yz[3];
  for (int c = 0; c < 3; c++) {
    d[c] = (drx * exi[c] + dre * eeta[c] - rel[3 + c] * dt) / 2;
    d[3 + c] = (dgx * exi[c] + dge * eeta[c]) / 2;
  }
  bool fin = true;
  for (int c = 0; c < 6; c++) fin = fin && isfinite(d[c]);
  if (!fin) *fl |= kNonFinite;
}

template <int METHOD>
BIA_HDF void pair_plain(double eps, double dt, int itr, const double rel[6], double d[6], int* branch, int* iters,
                        unsigned* fl, Hist hist) {
  const LearnTable none{nullptr, 0, 0};
  pair_eval<METHOD, kPlain>(eps, dt, itr, rel, d, branch, iters, fl, hist, none);
}

}  // namespace bia
