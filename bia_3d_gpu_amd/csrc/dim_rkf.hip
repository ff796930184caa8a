// dim_rkf.hip -- DIM: direct-sum N-body RKF45, the accuracy oracle of the BIA method
// (reference dim/rkf_kernel.cu:29-53 force, :75-145 Fehlberg stages, :194-495 stepper;
// driven by dim/dim_3D.f90:30).
//
// Device layout: SoA rows x,y,z,u,v,w of stride ld.  One Fehlberg stage is two kernels:
//   k_dim_force   one wave per (block of 64 particles, chunk of jc source particles): the partial
//                 direct-sum acceleration of its 64 particles over its chunk of j, in ascending j,
//                 into part[chunk][3][i].  The chunk length is chosen so the launch has thousands
//                 of waves whatever N (N = 1332: 21 x 42 waves of 32 sources each), so a small
//                 system still spreads over the whole GPU;
//   k_dim_combine one thread per particle and component: an acceleration component is the
//                 chunks' partials added in chunk order (deterministic), k_s[i] is stored and the
//                 next stage input w_s[i] = y0[i] + h sum_l b_sl k_l[i] formed in the same pass.  The last
//                 stage forms y_n and, for the error-controlled step, the per-block maximum of
//                 |err_i| / (relerr (|y0|+|yn|)/2 + abserr), so only nblocks doubles cross PCIe
//                 per trial step.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "bia_pair.hpp"
#include "bia_runtime.hpp"

namespace bia {

constexpr int kDimRows = 64;        // particles per force wave (one per lane)
constexpr int kDimMinChunk = 32;    // sources per force wave at least
constexpr int kDimTargetWaves = 4096;
constexpr int kDimCombine = 256;    // particles per combine block
constexpr int kDimPad = 256;        // row stride ld is a multiple of this

struct DimBufs {
  double* y[4];   // yl0, yln (trial step), yh0 (the high solution), yh1 (between its two half steps)
  double* k;      // 6 stages x 6 rows x ld
  double* w[2];   // ping-pong stage inputs
  double* er;     // 6 x ld
  double* part;   // nchunk x 3 x ld partial accelerations
  double* bmax;   // per-block error maxima (+1 flag slot per block)
};

// partial direct-sum acceleration (rkf_kernel.cu:29-53: a_i = sum_j eps (r_i - r_j) /
// (|r_i - r_j|^2 + 1e-99)^(3/2)) of the 64 particles of block blockIdx.x over the sources
// [blockIdx.y jc, min(np, (blockIdx.y + 1) jc)), j ascending; the sources pass through LDS 64 at a
// time (one coalesced load per lane, then broadcast reads).  CPL: the coupling of a pair is a
// per-source factor sw[j] in place of eps and a per-target factor tw[i] applied to the sums
// (gravity: sw = -G m, tw = 1; Coulomb: sw = q, tw = k q / m).
template <bool CPL>
__global__ __launch_bounds__(kDimRows) void k_dim_force(int np, int ld, double eps, int jc,
                                                        const double* __restrict__ w, const double* __restrict__ tw,
                                                        const double* __restrict__ sw, double* __restrict__ part) {
  __shared__ double src[4][kDimRows];
  const int lane = threadIdx.x;
  const int i = blockIdx.x * kDimRows + lane;
  const bool ok = i < np;
  const double xi = ok ? w[i] : 0, yi = ok ? w[ld + i] : 0, zi = ok ? w[2 * (size_t)ld + i] : 0;
  const int j0 = blockIdx.y * jc, j1 = min(np, j0 + jc);
  double sx = 0, sy = 0, sz = 0;
  for (int jb = j0; jb < j1; jb += kDimRows) {
    const int nj = min(kDimRows, j1 - jb);
    __syncthreads();  // the previous batch's readers are done
    if (lane < nj) {
      src[0][lane] = w[jb + lane];
      src[1][lane] = w[ld + jb + lane];
      src[2][lane] = w[2 * (size_t)ld + jb + lane];
      if (CPL) src[3][lane] = sw[jb + lane];
    }
    __syncthreads();
    for (int jj = 0; jj < nj; ++jj) {
      const double xij = xi - src[0][jj], yij = yi - src[1][jj], zij = zi - src[2][jj];
      const double rsq = xij * xij + yij * yij + zij * zij + 1e-99;  // rkf_kernel.cu:44
      const double e = (CPL ? src[3][jj] : eps) / (rsq * sqrt(rsq));
      sx += xij * e;
      sy += yij * e;
      sz += zij * e;
    }
  }
  if (CPL && ok) {
    const double t = tw[i];
    sx *= t;
    sy *= t;
    sz *= t;
  }
  double* p = part + (size_t)blockIdx.y * 3 * ld + i;  // i < ld: the padding columns take zeros
  p[0] = sx;
  p[ld] = sy;
  p[2 * (size_t)ld] = sz;
}

// acceleration component q of particle i: the nchunk partials in chunk order (loads issued 16
// chunks at a time, the adds in chunk order)
__device__ __forceinline__ double dim_accel(int nchunk, int ld, const double* __restrict__ part, int i, int q) {
  const size_t cs = (size_t)3 * ld;
  const double* p = part + (size_t)q * ld + i;
  double a = 0;
  int c = 0;
  for (; c + 16 <= nchunk; c += 16) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; u++) v[u] = p[(c + u) * cs];
#pragma unroll
    for (int u = 0; u < 16; u++) a += v[u];
  }
  for (; c < nchunk; ++c) a += p[c * cs];
  return a;
}

// stage S (1..6) of one Fehlberg step of size h from y0, after k_dim_force on win; one thread per
// particle and component (component blockIdx.y):
//   k_S = f(win);  S < 6: wout = y0 + h sum_{l<=S} b_{S+1,l} k_l
//                  S = 6: yn = y0 + h sum ga_l k_l; ERR: er, block max of |er|/et
template <int S, bool ERR>
__global__ __launch_bounds__(kDimCombine) void k_dim_combine(int np, int ld, int nchunk, const double* __restrict__ part,
                                                             const double* __restrict__ hp, double hs,
                                                             const double* __restrict__ y0,
                                                             const double* __restrict__ win, double* __restrict__ k,
                                                             double* __restrict__ wout, double* __restrict__ yn,
                                                             double* __restrict__ er, double* __restrict__ bmax,
                                                             double relerr, double abserr) {
  using namespace fehl;
  const int i = blockIdx.x * kDimCombine + threadIdx.x;
  const int c = blockIdx.y;
  const double h = hs * hp[0];  // hs = 1 or 1/2 (exact)
  double emax = 0, bad = 0;
  if (i < np) {
    const size_t sl = (size_t)6 * ld;  // stride between stages
    const size_t o = (size_t)c * ld + i;
    const double ks = c < 3 ? win[3 * (size_t)ld + o] : dim_accel(nchunk, ld, part, i, c - 3);
    k[(size_t)(S - 1) * sl + o] = ks;
    const double k1 = S == 1 ? ks : k[o], k2 = (S >= 2) ? (S == 2 ? ks : k[sl + o]) : 0,
                 k3 = (S >= 3) ? (S == 3 ? ks : k[2 * sl + o]) : 0, k4 = (S >= 4) ? (S == 4 ? ks : k[3 * sl + o]) : 0,
                 k5 = (S >= 5) ? (S == 5 ? ks : k[4 * sl + o]) : 0, k6 = (S >= 6) ? ks : 0;
    const double y = y0[o];
    if (S == 1) wout[o] = y + h * B21 * k1;  // rkf_kernel.cu:82
    if (S == 2) wout[o] = y + h * (B31 * k1 + B32 * k2);
    if (S == 3) wout[o] = y + h * (B41 * k1 + B42 * k2 + B43 * k3);
    if (S == 4) wout[o] = y + h * (B51 * k1 + B52 * k2 + B53 * k3 + B54 * k4);
    if (S == 5) wout[o] = y + h * (B61 * k1 + B62 * k2 + B63 * k3 + B64 * k4 + B65 * k5);
    if (S == 6) {
      const double ynv = y + h * (GA1 * k1 + GA3 * k3 + GA4 * k4 + GA5 * k5 + GA6 * k6);
      yn[o] = ynv;
      if (ERR) {
        const double ev = h * (DA1 * k1 + DA3 * k3 + DA4 * k4 + DA5 * k5 + DA6 * k6);
        er[o] = ev;
        const double et = relerr * (fabs(y) + fabs(ynv)) / 2 + abserr;
        if (et == 0) bad = 1;
        else emax = fabs(ev) / et;
      }
    }
  }
  if (S == 6 && ERR) {
    __shared__ double red[2][kDimCombine / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      emax = fmax(emax, __shfl_xor(emax, off, 64));
      bad = fmax(bad, __shfl_xor(bad, off, 64));
    }
    if (threadIdx.x % 64 == 0) {
      red[0][threadIdx.x / 64] = emax;
      red[1][threadIdx.x / 64] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int q = 1; q < kDimCombine / 64; q++) {
        emax = fmax(emax, red[0][q]);
        bad = fmax(bad, red[1][q]);
      }
      const size_t b = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
      bmax[2 * b] = emax;
      bmax[2 * b + 1] = bad;
    }
  }
}

// dst = src over 6 rows of stride ld (the accepted low solution becomes the next step's start)
__global__ __launch_bounds__(256) void k_dim_copy(int n, const double* __restrict__ src, double* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// derivative only (initial step-size estimate): f = (v, a) at y, after k_dim_force on y
__global__ __launch_bounds__(kDimCombine) void k_dim_func(int np, int ld, int nchunk, const double* __restrict__ part,
                                                          const double* __restrict__ y, double* __restrict__ f) {
  const int i = blockIdx.x * kDimCombine + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= np) return;
  f[(size_t)c * ld + i] = c < 3 ? y[(size_t)(3 + c) * ld + i] : dim_accel(nchunk, ld, part, i, c - 3);
}

// One call's launch plan.  A trial step (error-controlled Fehlberg step of yl0 into yln) and an
// accepted step (two half steps of yh0 in place through yh1, then yl0 = yln) are each captured
// once into a hipGraph and replayed; the step size is read by the kernels from device memory
// (hdev), written before each replay, so the graphs do not change from step to step.
struct DimRun {
  bia_ctx* ctx;
  int np, ld, nb;   // nb: force blocks of kDimRows particles
  int jc, nchunk;   // sources per force wave, chunks of j
  int ncb;          // combine blocks of kDimCombine particles
  double eps, relerr, abserr;
  DimBufs b;
  double* hdev;     // device: step size of the trial step [0] and of the accepted step [1]
  double* hpin;     // pinned host staging of the same two values
  int evals = 0;
  hipGraphExec_t trial = nullptr, accept = nullptr;

  const double* tw = nullptr;  // coupled force: per-target and per-source factors (device, ld), else NULL
  const double* sw = nullptr;

  void force(const double* w) {
    if (sw)
      hipLaunchKernelGGL(k_dim_force<true>, dim3(nb, nchunk), dim3(kDimRows), 0, ctx->stream, np, ld, eps, jc, w, tw,
                         sw, b.part);
    else
      hipLaunchKernelGGL(k_dim_force<false>, dim3(nb, nchunk), dim3(kDimRows), 0, ctx->stream, np, ld, eps, jc, w,
                         nullptr, nullptr, b.part);
  }

  // Fehlberg step of size hs * hdev from y0 into yn (yn may be y0: each thread reads its own
  // elements of y0 before stage 6 writes them, and the force kernels that read all of y0 run
  // before)
  void fehl(const double* y0, double* yn, const double* hp, double hs, bool err) {
    const dim3 g(ncb, 6), t(kDimCombine);
    hipStream_t s = ctx->stream;
    double *w0 = b.w[0], *w1 = b.w[1];
    force(y0);
    hipLaunchKernelGGL((k_dim_combine<1, false>), g, t, 0, s, np, ld, nchunk, b.part, hp, hs, y0, y0, b.k, w0, nullptr, nullptr, nullptr, relerr, abserr);
    force(w0);
    hipLaunchKernelGGL((k_dim_combine<2, false>), g, t, 0, s, np, ld, nchunk, b.part, hp, hs, y0, w0, b.k, w1, nullptr, nullptr, nullptr, relerr, abserr);
    force(w1);
    hipLaunchKernelGGL((k_dim_combine<3, false>), g, t, 0, s, np, ld, nchunk, b.part, hp, hs, y0, w1, b.k, w0, nullptr, nullptr, nullptr, relerr, abserr);
    force(w0);
    hipLaunchKernelGGL((k_dim_combine<4, false>), g, t, 0, s, np, ld, nchunk, b.part, hp, hs, y0, w0, b.k, w1, nullptr, nullptr, nullptr, relerr, abserr);
    force(w1);
    hipLaunchKernelGGL((k_dim_combine<5, false>), g, t, 0, s, np, ld, nchunk, b.part, hp, hs, y0, w1, b.k, w0, nullptr, nullptr, nullptr, relerr, abserr);
    force(w0);
    if (err)
      hipLaunchKernelGGL((k_dim_combine<6, true>), g, t, 0, s, np, ld, nchunk, b.part, hp, hs, y0, w0, b.k, nullptr, yn, b.er, b.bmax, relerr, abserr);
    else
      hipLaunchKernelGGL((k_dim_combine<6, false>), g, t, 0, s, np, ld, nchunk, b.part, hp, hs, y0, w0, b.k, nullptr, yn, nullptr, nullptr, relerr, abserr);
  }

  int capture(hipGraphExec_t* exec, bool accept_step) {
    hipStream_t s = ctx->stream;
    BIA_HIP_TRY(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    if (accept_step) {
      fehl(b.y[2], b.y[3], hdev + 1, 0.5, false);
      fehl(b.y[3], b.y[2], hdev + 1, 0.5, false);
      const int n6 = 6 * ld;
      hipLaunchKernelGGL(k_dim_copy, dim3((n6 + 255) / 256), dim3(256), 0, s, n6, b.y[1], b.y[0]);
    } else {
      fehl(b.y[0], b.y[1], hdev, 1.0, true);
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      BIA_HIP_TRY(ctx, e);
    }
    const hipError_t ei = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    BIA_HIP_TRY(ctx, ei);
    return BIA_OK;
  }

  // stage the step size of slot q (0 trial, 1 accepted step) and replay that step's graph.  The
  // host rewrites a slot only after the stream has been synchronised past its last copy: each
  // trial step is followed by a synchronisation before the next accepted or trial step is issued.
  int run(int q, double h, int nevals) {
    hpin[q] = h;
    BIA_HIP_TRY(ctx, hipMemcpyAsync(hdev + q, hpin + q, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    BIA_HIP_TRY(ctx, hipGraphLaunch(q == 0 ? trial : accept, ctx->stream));
    evals += nevals;
    return BIA_OK;
  }

  ~DimRun() {
    if (trial) (void)hipGraphExecDestroy(trial);
    if (accept) (void)hipGraphExecDestroy(accept);
  }
};

// graph capture is not possible on the null stream: a call on it runs on the context's private
// blocking stream, which stays ordered with null-stream work on both sides
struct CaptureStream {
  bia_ctx* ctx;
  hipStream_t user;
  explicit CaptureStream(bia_ctx* c) : ctx(c), user(c->stream) {}
  int enter() {
    if (!user) {
      if (!ctx->cap_stream) BIA_HIP_TRY(ctx, hipStreamCreate(&ctx->cap_stream));
      ctx->stream = ctx->cap_stream;
    }
    return BIA_OK;
  }
  ~CaptureStream() { ctx->stream = user; }
};

// tw, sw: host arrays (np) of the coupled force's per-target and per-source factors, or NULL for
// the reference's scalar eps
static int dim_rkf45_locked(bia_ctx* ctx, int np, double eps, const double* tw, const double* sw, double* x0,
                            double xe, double* y, int* init, double relerr, double abserr, double* hio, double* ese,
                            int* nde, int* naccept, int* nreject, int* ier, FILE* trace) {
  const double EPSMIN = 1e-15;  // rkf_kernel.cu:201-202
  const int ITEMAX = 10000;
  if (np < 1 || !x0 || !y || !hio || !ier) return BIA_ERR_ARG;
  CaptureStream cs(ctx);
  BIA_TRY(ctx, cs.enter());
  // host vectors copied from and into asynchronously, declared before the drain so that they
  // outlive it on every return path
  std::vector<double> soa, fac, f, bm, p0, outh, outl;
  StreamDrain drain{ctx};
  const int ld = ((np + kDimPad - 1) / kDimPad) * kDimPad;
  const int nb = (np + kDimRows - 1) / kDimRows;
  // chunks of j: enough force waves to fill the GPU, at least kDimMinChunk sources each
  const int nchunk = std::max(1, std::min((kDimTargetWaves + nb - 1) / nb, (np + kDimMinChunk - 1) / kDimMinChunk));
  const int jc = (np + nchunk - 1) / nchunk;
  const int ncb = (np + kDimCombine - 1) / kDimCombine;
  const size_t row = (size_t)6 * ld;
  // layout: yl0 yln yh0 yh1 | k (6 stages) | w0 w1 | er | part | bmax | h (2) | tw sw (coupled)
  const size_t npart = (size_t)nchunk * 3 * ld;
  const size_t ndbl = 4 * row + 6 * row + 2 * row + row + npart + 2 * 6 * (size_t)ncb + 2 + (sw ? 2 * (size_t)ld : 0);
  BIA_TRY(ctx, ctx->dim_work.ensure(ndbl * sizeof(double)));
  BIA_TRY(ctx, ctx->dim_h.ensure(2 * sizeof(double)));
  double* base = ctx->dim_work.as<double>();
  DimRun R{ctx, np, ld, nb, jc, nchunk, ncb, eps, relerr, abserr, {}, nullptr, ctx->dim_h.host<double>()};
  for (int q = 0; q < 4; q++) R.b.y[q] = base + q * row;
  R.b.k = base + 4 * row;
  R.b.w[0] = base + 10 * row;
  R.b.w[1] = base + 11 * row;
  R.b.er = base + 12 * row;
  R.b.part = base + 13 * row;
  R.b.bmax = base + 13 * row + npart;
  R.hdev = R.b.bmax + 2 * 6 * (size_t)ncb;
  double *yl0 = R.b.y[0], *yh0 = R.b.y[2];
  hipStream_t s = ctx->stream;
  // AoS host -> SoA device (padding rows zeroed once)
  soa.assign(row, 0.0);
  for (int i = 0; i < np; i++)
    for (int c = 0; c < 6; c++) soa[(size_t)c * ld + i] = y[(size_t)6 * i + c];
  BIA_HIP_TRY(ctx, hipMemcpyAsync(yh0, soa.data(), row * sizeof(double), hipMemcpyHostToDevice, s));
  BIA_HIP_TRY(ctx, hipMemcpyAsync(yl0, yh0, row * sizeof(double), hipMemcpyDeviceToDevice, s));
  if (sw) {
    double* d = R.hdev + 2;
    fac.assign(2 * (size_t)ld, 0.0);
    std::copy(tw, tw + np, fac.begin());
    std::copy(sw, sw + np, fac.begin() + ld);
    BIA_HIP_TRY(ctx, hipMemcpyAsync(d, fac.data(), fac.size() * sizeof(double), hipMemcpyHostToDevice, s));
    R.tw = d;
    R.sw = d + ld;
  }
  double H = *hio;
  if (init && *init == 1) {  // initial step from the derivative, :237-265
    R.force(yl0);
    hipLaunchKernelGGL(k_dim_func, dim3(ncb, 6), dim3(kDimCombine), 0, s, np, ld, nchunk, R.b.part, yl0, R.b.k);
    BIA_HIP_TRY(ctx, hipGetLastError());
    f.assign(row, 0.0);
    BIA_HIP_TRY(ctx, hipMemcpyAsync(f.data(), R.b.k, row * sizeof(double), hipMemcpyDeviceToHost, s));
    BIA_HIP_TRY(ctx, hipStreamSynchronize(s));
    double toldy = 1.7976931348623157e308;
    bool changed = false;
    for (int i = 0; i < np; i++)
      for (int c = 0; c < 6; c++) {
        const double fv = f[(size_t)c * ld + i];
        if (fabs(fv) > 0.0) {
          const double tol = relerr * fabs(y[(size_t)6 * i + c]) + abserr;
          toldy = fmin(toldy, fabs(tol / fv));
          changed = true;
        }
      }
    H = changed ? fmin((xe - *x0) / 100e0, pow(toldy, 0.2)) : EPSMIN * (xe - *x0);
    *init = 2;
    R.evals = 1;
  }
  bm.assign(2 * 6 * (size_t)ncb, 0.0);
  p0.assign(6, 0.0);
  auto trace_row = [&](double t, const double* dy) -> int {
    if (!trace) return BIA_OK;
    for (int c = 0; c < 6; c++)
      BIA_HIP_TRY(ctx, hipMemcpyAsync(&p0[c], dy + (size_t)c * ld, sizeof(double), hipMemcpyDeviceToHost, s));
    BIA_HIP_TRY(ctx, hipStreamSynchronize(s));
    fprintf(trace, "%24.16e%25.16e%25.16e%25.16e%25.16e%25.16e%25.16e\n", t, p0[0], p0[1], p0[2], p0[3], p0[4],
            p0[5]);
    return BIA_OK;
  };
  if (trace) {
    fprintf(trace, "%24s%25s%25s%25s%25s%25s%25s\n", "time", "x", "y", "z", "u", "v", "w");
    BIA_TRY(ctx, trace_row(*x0, yh0));
  }
  int acc = 0, rej = 0;
  *ier = 10000;
  double X = *x0;
  BIA_TRY(ctx, R.capture(&R.trial, false));
  BIA_TRY(ctx, R.capture(&R.accept, true));
  double* yhi = R.b.y[2];  // the high solution, advanced in place by the accepted step
  for (int iter = 0; iter < ITEMAX; iter++) {
    BIA_TRY(ctx, R.run(0, H, 6));
    BIA_HIP_TRY(ctx, hipMemcpyAsync(bm.data(), R.b.bmax, bm.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    BIA_HIP_TRY(ctx, hipStreamSynchronize(s));
    double erret = 0;
    bool bad = false;
    for (int q = 0; q < 6 * ncb; q++) {
      erret = fmax(erret, bm[2 * q]);
      bad = bad || bm[2 * q + 1] != 0;
    }
    if (bad) {
      *ier = 20000;
      break;
    }
    if (erret > 1) {
      rej++;
      H *= (erret >= 59049) ? 0.1 : 0.9 / pow(erret, 0.2);
      if (H <= EPSMIN) {
        *ier = 30000;
        break;
      }
    } else if (X + H < xe) {
      // high solution: two half steps, then the low solution moves on (:318-363)
      BIA_TRY(ctx, R.run(1, H, 12));
      X += H;
      *x0 = X;
      acc++;
      BIA_TRY(ctx, trace_row(X, yhi));
      H *= (erret <= 1.889568e-4) ? 5.0 : 0.9 / pow(erret, 0.2);
    } else {
      H = xe - X;  // last phase, :387-483
      BIA_TRY(ctx, R.run(0, H, 6));
      BIA_TRY(ctx, R.run(1, H, 12));
      acc++;
      *x0 = xe;
      // results back to AoS host arrays
      outh.assign(row, 0.0);
      outl.assign(row, 0.0);
      BIA_HIP_TRY(ctx, hipMemcpyAsync(outh.data(), yhi, row * sizeof(double), hipMemcpyDeviceToHost, s));
      BIA_HIP_TRY(ctx, hipMemcpyAsync(outl.data(), R.b.y[1], row * sizeof(double), hipMemcpyDeviceToHost, s));
      BIA_HIP_TRY(ctx, hipStreamSynchronize(s));
      for (int i = 0; i < np; i++)
        for (int c = 0; c < 6; c++) {
          const double vh = outh[(size_t)c * ld + i], vl = outl[(size_t)c * ld + i];
          y[(size_t)6 * i + c] = vh;
          if (ese) ese[(size_t)6 * i + c] = (vl - vh) / 31;
        }
      BIA_TRY(ctx, trace_row(xe, yhi));
      *ier = 0;
      break;
    }
  }
  *hio = H;
  if (nde) *nde = R.evals;
  if (naccept) *naccept = acc;
  if (nreject) *nreject = rej;
  return *ier == 0 ? BIA_OK : BIA_ERR_RKF;
}

}  // namespace bia

using namespace bia;

extern "C" {

int bia_dim_rkf45_host(bia_ctx* ctx, int np, double eps, double* x0, double xe, double* y, int* init, double relerr,
                       double abserr, double* h, double* ese, int* nde, int* naccept, int* nreject, int* ier) {
  if (!ctx) return BIA_ERR_ARG;
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->last_error[0] = 0;
  BIA_HIP_TRY(ctx, hipSetDevice(ctx->device));
  return dim_rkf45_locked(ctx, np, eps, nullptr, nullptr, x0, xe, y, init, relerr, abserr, h, ese, nde, naccept,
                          nreject, ier, nullptr);
}

int bia_dim_rkf45_coupled_host(bia_ctx* ctx, int coupling, double coup, int np, const double* mass,
                               const double* charge, double* x0, double xe, double* y, int* init, double relerr,
                               double abserr, double* h, double* ese, int* nde, int* naccept, int* nreject, int* ier) {
  if (!ctx) return BIA_ERR_ARG;
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->last_error[0] = 0;
  if (np < 1 || !(coupling == BIA_COUPLING_GRAVITY ? mass != nullptr : coupling == BIA_COUPLING_COULOMB && charge))
    return BIA_ERR_ARG;
  BIA_HIP_TRY(ctx, hipSetDevice(ctx->device));
  std::vector<double> tw(np), sw(np);
  for (int i = 0; i < np; i++) {
    if (coupling == BIA_COUPLING_GRAVITY) {  // a_i = -G sum_j m_j r_ij / r^3
      tw[i] = 1.0;
      sw[i] = -coup * mass[i];
    } else {  // a_i = (k q_i / m_i) sum_j q_j r_ij / r^3
      tw[i] = coup * charge[i] / (mass ? mass[i] : 1.0);
      sw[i] = charge[i];
    }
  }
  return dim_rkf45_locked(ctx, np, 0.0, tw.data(), sw.data(), x0, xe, y, init, relerr, abserr, h, ese, nde, naccept,
                          nreject, ier, nullptr);
}

// Reference-compatible rkf_c_.  When BIA_DIM_TRACE names a file, the accepted-step trajectory of
// particle 1 is written there in the reference's text format (its DEBUG output,
// rkf_kernel.cu:209-216,376-377,476-477).
void rkf_c_(int neq, double eps_th, double* X0, double* XE, double* YH0, int* INIT, const double* RELERR,
            const double* ABSERR, double* YHN, double* ESE, int* NDE, int* IER, double* H, double* YL0,
            double* YLN, double* ERR, double* WRK) {
  (void)ERR;
  (void)WRK;
  bia_ctx* ctx = default_ctx();
  if (neq < 6 || neq % 6 != 0) {
    snprintf(ctx->last_error, sizeof(ctx->last_error), "neq=%d", neq);
    legacy_fail("rkf_c_ arguments", BIA_ERR_ARG, ctx);
  }
  const int np = neq / 6;
  std::vector<double> y(YH0, YH0 + neq);
  FILE* tr = nullptr;
  if (const char* p = getenv("BIA_DIM_TRACE")) tr = fopen(p, "w");
  int ier = 0, nde = 0;
  int s;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->last_error[0] = 0;
    if (hipSetDevice(ctx->device) != hipSuccess) legacy_fail("rkf_c_", BIA_ERR_HIP, ctx);
    s = dim_rkf45_locked(ctx, np, eps_th, nullptr, nullptr, X0, *XE, y.data(), INIT, *RELERR, *ABSERR, H, ESE, &nde,
                         nullptr, nullptr, &ier, tr);
  }
  if (tr) fclose(tr);
  *NDE = nde;
  *IER = ier;
  if (s == BIA_ERR_RKF) {
    // the reference prints the reason and exits with status 0 (rkf_kernel.cu:298-315,491-494)
    if (ier == 20000) printf("(SUBR -RKF) TROUBLE(INAPPROPRIATE ERROR TOLERANCE)\n");
    else if (ier == 30000) printf("(SUBR -RKF) TROUBLE(TOO SMALL STEP SIZE) H %25.16e\n", *H);
    else printf("(SUBR -RKF) TROUBLE(TOO MANY ITERATIONS/ ITER =%d)\n", 10000);
    fflush(stdout);
    exit(0);
  }
  if (s != BIA_OK) legacy_fail("rkf_c_", s, ctx);
  (void)YL0;
  (void)YLN;
  memcpy(YHN, y.data(), sizeof(double) * neq);
  memcpy(YH0, y.data(), sizeof(double) * neq);
}

}  // extern "C"
