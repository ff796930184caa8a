// bia_step.hip -- whole-step pair kernels (ALG / RKF45) and the reference-compatible
// per-particle entry point find_dxyuv_kernel_.
//
// Whole step (bia_step_dev): the i<j triangle is cut into 64x64 tiles (I <= J).  One wave of
// 64 lanes owns one tile: lane l holds row particle i = 64 I + l in registers and walks the
// 64 column particles j of tile J in order (j is wave-uniform, so its six doubles are
// scalar loads).  Each lane accumulates its row sum in registers; the column sum of each j
// is a 64-lane butterfly reduction.  Results go to a partial buffer part[K][c][i] in which
// every (tile K, particle i) slot has exactly one writer:
//   tile (I,J), I<J : rows of I  -> part[J][c][i],  columns of J -> part[I][c][j]
//   tile (I,I)      : both sides -> part[I][c][i]
// so no atomics and no clearing are needed, and a second pass sums part over K in a fixed
// order (deterministic results).  The drift v*dt is added in that pass.
#include "bia_pair.hpp"
#include "bia_runtime.hpp"

namespace bia {

constexpr int kWave = 64;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}
__device__ __forceinline__ unsigned wave_sum_u(unsigned v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

struct PairStats {
  unsigned cnt[4];
  unsigned its[4];
  unsigned fl[8];
  __device__ void clear() {
    for (int k = 0; k < 4; k++) cnt[k] = its[k] = 0;
    for (int k = 0; k < 8; k++) fl[k] = 0;
  }
  __device__ void add(int br, int it, unsigned f) {
    // branch-indexed without dynamic register indexing
    for (int k = 0; k < 4; k++) {
      cnt[k] += (br == k);
      its[k] += (br == k) ? (unsigned)it : 0u;
    }
    for (int k = 0; k < 8; k++) fl[k] += (f >> k) & 1u;
  }
  __device__ void flush(unsigned long long* out) {
    unsigned v[16];
    for (int k = 0; k < 4; k++) {
      v[k] = wave_sum_u(cnt[k]);
      v[4 + k] = wave_sum_u(its[k]);
    }
    for (int k = 0; k < 8; k++) v[8 + k] = wave_sum_u(fl[k]);
    if (threadIdx.x == 0)
      for (int k = 0; k < 16; k++)
        if (v[k]) atomicAdd(out + k, (unsigned long long)v[k]);
  }
};

template <int METHOD, bool STATS>
__global__ __launch_bounds__(kWave) void k_step_tiles(int n, int lds, int ldp, double eps, double dt, int itr,
                                                      const double* __restrict__ st,
                                                      const int2* __restrict__ tiles,
                                                      double* __restrict__ part,
                                                      unsigned long long* __restrict__ stats) {
  __shared__ double hist_lds[kMaxItr * kWave];
  const int lane = threadIdx.x;
  const int2 t = tiles[blockIdx.x];
  const int I = t.x, J = t.y;
  const int i = I * kWave + lane;
  const bool row_ok = i < n;
  double ri[6];
#pragma unroll
  for (int c = 0; c < 6; c++) ri[c] = row_ok ? st[(size_t)c * lds + i] : 0.0;
  double acc[6] = {0, 0, 0, 0, 0, 0}, col[6] = {0, 0, 0, 0, 0, 0};
  PairStats ps;
  if (STATS) ps.clear();
  Hist hist{hist_lds + lane, kWave};
  const int j0 = J * kWave;
  const int jn = min(kWave, n - j0);
#pragma unroll 1
  for (int jj = 0; jj < jn; ++jj) {
    const int j = j0 + jj;
    double rj[6];
#pragma unroll
    for (int c = 0; c < 6; c++) rj[c] = st[(size_t)c * lds + j];
    double d[6] = {0, 0, 0, 0, 0, 0};
    if (row_ok && (I != J || lane < jj)) {
      double rel[6];
#pragma unroll
      for (int c = 0; c < 6; c++) rel[c] = ri[c] - rj[c];
      int br = 0, it = 0;
      unsigned fl = 0;
      pair_plain<METHOD>(eps, dt, itr, rel, d, &br, &it, &fl, hist);
      if (STATS) ps.add(br, it, fl);
    }
#pragma unroll
    for (int c = 0; c < 6; c++) {
      acc[c] += d[c];
      double s = wave_sum(d[c]);
      if (lane == jj) col[c] = -s;
    }
  }
  if (I == J) {
#pragma unroll
    for (int c = 0; c < 6; c++) part[((size_t)I * 6 + c) * ldp + i] = acc[c] + col[c];
  } else {
#pragma unroll
    for (int c = 0; c < 6; c++) {
      part[((size_t)J * 6 + c) * ldp + i] = acc[c];
      part[((size_t)I * 6 + c) * ldp + j0 + lane] = col[c];
    }
  }
  if (STATS) ps.flush(stats);
}

// delta[c][i] = sum_K part[K][c][i] (+ v*dt for positions), K ascending
__global__ __launch_bounds__(256) void k_step_reduce(int n, int lds, int ldp, int ldd, int nt, double dt,
                                                     const double* __restrict__ st,
                                                     const double* __restrict__ part,
                                                     double* __restrict__ delta) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= n) return;
  double s = 0;
  const double* p = part + (size_t)c * ldp + i;
  for (int K = 0; K < nt; ++K) s += p[(size_t)K * 6 * ldp];
  if (c < 3) s += st[(size_t)(3 + c) * lds + i] * dt;
  delta[(size_t)c * ldd + i] = s;
}

// reference per-particle form: thread per j > i on an AoS relative-state row
template <int METHOD>
__global__ __launch_bounds__(kWave) void k_legacy_row(int n, int i, double eps, double dt, int itr,
                                                      const double* __restrict__ rel_aos,
                                                      double* __restrict__ out_aos) {
  __shared__ double hist_lds[kMaxItr * kWave];
  const int j = (i + 1) + blockIdx.x * kWave + threadIdx.x;
  if (j >= n) return;
  double rel[6], d[6];
#pragma unroll
  for (int c = 0; c < 6; c++) rel[c] = rel_aos[(size_t)6 * j + c];
  int br = 0, it = 0;
  unsigned fl = 0;
  Hist hist{hist_lds + threadIdx.x, kWave};
  pair_plain<METHOD>(eps, dt, itr, rel, d, &br, &it, &fl, hist);
#pragma unroll
  for (int c = 0; c < 6; c++) out_aos[(size_t)6 * j + c] = d[c];
}

// the step on device-resident SoA state; caller holds ctx->mu
static int step_dev_locked(bia_ctx* ctx, int method, int mode, int n, double eps, double dt, int itr,
                           const double* d_state, int ld, double* d_delta, unsigned long long* d_stats) {
  if (n < 1 || ld < n || !d_state || !d_delta) return BIA_ERR_ARG;
  if (method != BIA_METHOD_ALG && method != BIA_METHOD_RKF) return BIA_ERR_ARG;
  if (mode != BIA_MODE_PLAIN) return BIA_ERR_ARG;  // learned mode dispatches elsewhere
  if (itr < 0 || itr > kMaxItr) return BIA_ERR_ARG;
  const int nt = (n + kWave - 1) / kWave;
  const int ldp = nt * kWave;  // partial rows are padded to whole tiles
  BIA_TRY(ctx, ensure_tiles(ctx, nt));
  BIA_TRY(ctx, ctx->part.ensure((size_t)nt * 6 * ldp * sizeof(double)));
  const unsigned ntiles = (unsigned)((size_t)nt * (nt + 1) / 2);
  double* part = ctx->part.as<double>();
  const int2* tiles = ctx->tiles.as<int2>();
  dim3 g(ntiles), b(kWave);
  const bool stats = d_stats != nullptr;
  BIA_TRY(ctx, prof_begin(ctx));
  if (method == BIA_METHOD_ALG) {
    if (stats)
      hipLaunchKernelGGL((k_step_tiles<1, true>), g, b, 0, ctx->stream, n, ld, ldp, eps, dt, itr, d_state, tiles,
                         part, d_stats);
    else
      hipLaunchKernelGGL((k_step_tiles<1, false>), g, b, 0, ctx->stream, n, ld, ldp, eps, dt, itr, d_state, tiles,
                         part, d_stats);
  } else {
    if (stats)
      hipLaunchKernelGGL((k_step_tiles<0, true>), g, b, 0, ctx->stream, n, ld, ldp, eps, dt, itr, d_state, tiles,
                         part, d_stats);
    else
      hipLaunchKernelGGL((k_step_tiles<0, false>), g, b, 0, ctx->stream, n, ld, ldp, eps, dt, itr, d_state, tiles,
                         part, d_stats);
  }
  BIA_HIP_TRY(ctx, hipGetLastError());
  BIA_TRY(ctx, prof_end(ctx));
  hipLaunchKernelGGL(k_step_reduce, dim3((n + 255) / 256, 6), dim3(256), 0, ctx->stream, n, ld, ldp, ld, nt, dt,
                     d_state, part, d_delta);
  BIA_HIP_TRY(ctx, hipGetLastError());
  return BIA_OK;
}

}  // namespace bia

using namespace bia;

extern "C" {

int bia_step_dev(bia_ctx* ctx, int method, int mode, int n, double eps, double dt, int itr,
                 const double* d_state, int ld, double* d_delta, unsigned long long* d_stats) {
  if (!ctx) return BIA_ERR_ARG;
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->last_error[0] = 0;
  BIA_HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (mode == BIA_MODE_LEARN) {
    if (method != BIA_METHOD_ALG) return BIA_ERR_ARG;
    return step_learn_locked(ctx, n, eps, dt, itr, d_state, ld, d_delta, d_stats);
  }
  return step_dev_locked(ctx, method, mode, n, eps, dt, itr, d_state, ld, d_delta, d_stats);
}

int bia_step_host(bia_ctx* ctx, int method, int mode, int n, double eps, double dt, int itr, const double* X,
                  const double* Y, const double* Z, const double* U, const double* V, const double* W,
                  double* dX, double* dY, double* dZ, double* dU, double* dV, double* dW,
                  unsigned long long* stats) {
  if (!ctx || n < 1) return BIA_ERR_ARG;
  const double* in[6] = {X, Y, Z, U, V, W};
  double* out[6] = {dX, dY, dZ, dU, dV, dW};
  for (int c = 0; c < 6; c++)
    if (!in[c] || !out[c]) return BIA_ERR_ARG;
  const int ld = ((n + kWave - 1) / kWave) * kWave;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->last_error[0] = 0;
    BIA_HIP_TRY(ctx, hipSetDevice(ctx->device));
    BIA_TRY(ctx, ctx->delta.ensure((size_t)6 * ld * sizeof(double) * 2 + BIA_STATS_LEN * 8));
    // host staging: [state 6*ld][delta 6*ld][stats]
    double* dst = ctx->delta.as<double>();
    double* ddl = dst + (size_t)6 * ld;
    unsigned long long* dsa = reinterpret_cast<unsigned long long*>(ddl + (size_t)6 * ld);
    BIA_HIP_TRY(ctx, hipMemsetAsync(dst, 0, (size_t)6 * ld * sizeof(double), ctx->stream));
    for (int c = 0; c < 6; c++)
      BIA_HIP_TRY(ctx, hipMemcpyAsync(dst + (size_t)c * ld, in[c], (size_t)n * sizeof(double),
                                      hipMemcpyHostToDevice, ctx->stream));
    BIA_HIP_TRY(ctx, hipMemsetAsync(dsa, 0, BIA_STATS_LEN * 8, ctx->stream));
  }
  double* dst = ctx->delta.as<double>();
  double* ddl = dst + (size_t)6 * ld;
  unsigned long long* dsa = reinterpret_cast<unsigned long long*>(ddl + (size_t)6 * ld);
  int s = bia_step_dev(ctx, method, mode, n, eps, dt, itr, dst, ld, ddl, dsa);
  if (s != BIA_OK) return s;
  std::lock_guard<std::mutex> lk(ctx->mu);
  for (int c = 0; c < 6; c++)
    BIA_HIP_TRY(ctx, hipMemcpyAsync(out[c], ddl + (size_t)c * ld, (size_t)n * sizeof(double),
                                    hipMemcpyDeviceToHost, ctx->stream));
  if (stats)
    BIA_HIP_TRY(ctx, hipMemcpyAsync(stats, dsa, BIA_STATS_LEN * 8, hipMemcpyDeviceToHost, ctx->stream));
  BIA_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return BIA_OK;
}

void find_dxyuv_kernel_(int METHOD, int N, int I, double EPS, double DT, int ITR, double* dX, double* dY,
                        double* dZ, double* dU, double* dV, double* dW, double* RXYZUVW, double* dXYZUVW) {
  bia_ctx* ctx = default_ctx();
  const int i = I - 1;  // Fortran index
  if (N < 1 || i < 0 || i >= N || ITR < 0 || ITR > kMaxItr || (METHOD != 0 && METHOD != 1)) {
    snprintf(ctx->last_error, sizeof(ctx->last_error), "METHOD=%d N=%d I=%d ITR=%d", METHOD, N, I, ITR);
    legacy_fail("find_dxyuv_kernel_ arguments", BIA_ERR_ARG, ctx);
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  const size_t bytes = (size_t)6 * N * sizeof(double);
  int s;
  if ((s = ctx->leg_in.ensure(bytes)) != BIA_OK) legacy_fail("find_dxyuv_kernel_ allocation", s, ctx);
  if ((s = ctx->leg_out.ensure(bytes)) != BIA_OK) legacy_fail("find_dxyuv_kernel_ allocation", s, ctx);
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess)
    e = hipMemcpyAsync(ctx->leg_in.p, RXYZUVW, bytes, hipMemcpyHostToDevice, ctx->stream);
  const int cnt = N - (i + 1);
  if (e == hipSuccess && cnt > 0) {
    dim3 g((cnt + kWave - 1) / kWave), b(kWave);
    if (METHOD == 1)
      hipLaunchKernelGGL(k_legacy_row<1>, g, b, 0, ctx->stream, N, i, EPS, DT, ITR, ctx->leg_in.as<double>(),
                         ctx->leg_out.as<double>());
    else
      hipLaunchKernelGGL(k_legacy_row<0>, g, b, 0, ctx->stream, N, i, EPS, DT, ITR, ctx->leg_in.as<double>(),
                         ctx->leg_out.as<double>());
    e = hipGetLastError();
    // only rows j > i are produced (the reference copies back the whole array, leaving the
    // rest unspecified); copy back exactly those
    if (e == hipSuccess)
      e = hipMemcpyAsync(dXYZUVW + (size_t)6 * (i + 1), ctx->leg_out.as<double>() + (size_t)6 * (i + 1),
                         (size_t)6 * cnt * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    snprintf(ctx->last_error, sizeof(ctx->last_error), "%s", hipGetErrorString(e));
    legacy_fail("find_dxyuv_kernel_", BIA_ERR_HIP, ctx);
  }
  // Newton's third law scatter, in the reference's order (find_dXYUV_kernel.cu:541-546)
  for (int j = i + 1; j < N; j++) {
    const double* d = dXYZUVW + (size_t)6 * j;
    dX[i] += d[0]; dY[i] += d[1]; dZ[i] += d[2];
    dX[j] -= d[0]; dY[j] -= d[1]; dZ[j] -= d[2];
    dU[i] += d[3]; dV[i] += d[4]; dW[i] += d[5];
    dU[j] -= d[3]; dV[j] -= d[4]; dW[j] -= d[5];
  }
}

}  // extern "C"
