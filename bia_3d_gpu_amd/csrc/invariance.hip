// invariance.hip -- energy / momentum / angular-momentum invariants (reference
// bia/invariance.cu:15-110, CPU twins bia/Invariance_CPU.f90, bia/Invariance_QPU.f90).
//
//   KE = sum |v|^2, PE = eps * sum_{i<j} 1/|r_i - r_j|, L = sum r x v, sums of x..w
//
// The O(N^2) potential term is computed over 64x64 tiles of the i<j triangle (one wave per
// tile, lane = row, column particle wave-uniform).  Every lane accumulates in double-double
// (TwoSum), tile partials are written to a buffer and summed in a fixed order (256 blocks over
// contiguous chunks, then one wave), so the result
// is deterministic and close to the quad-precision sum of Invariance_QPU.f90 rather than to
// the plain double sum.  The ten O(N) sums use the same double-double scheme.
#include "bia_runtime.hpp"

namespace bia {

struct DD {
  double hi, lo;
};
__device__ __forceinline__ void dd_add(DD& a, double x) {  // Neumaier/TwoSum accumulation
  double s = a.hi + x;
  double bb = s - a.hi;
  double e = (a.hi - (s - bb)) + (x - bb);
  a.hi = s;
  a.lo += e;
}
__device__ __forceinline__ void dd_add(DD& a, const DD& b) {
  dd_add(a, b.hi);
  a.lo += b.lo;
}
__device__ __forceinline__ DD dd_shfl_xor(const DD& a, int off) {
  return DD{__shfl_xor(a.hi, off, 64), __shfl_xor(a.lo, off, 64)};
}
__device__ __forceinline__ DD dd_wave_sum(DD a) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) dd_add(a, dd_shfl_xor(a, off));
  return a;
}

// potential tiles: partial[t] = sum over the tile's pairs of 1/r (double-double)
__global__ __launch_bounds__(64) void k_inv_pe_tiles(int n, int ld, const double* __restrict__ st,
                                                     const int2* __restrict__ tiles, DD* __restrict__ part) {
  const int lane = threadIdx.x;
  const int2 t = tiles[blockIdx.x];
  const int I = t.x, J = t.y;
  const int i = I * 64 + lane;
  const bool ok = i < n;
  const double xi = ok ? st[i] : 0, yi = ok ? st[ld + i] : 0, zi = ok ? st[2 * (size_t)ld + i] : 0;
  DD s{0, 0};
  const int j0 = J * 64, jn = min(64, n - j0);
  for (int jj = 0; jj < jn; ++jj) {
    const int j = j0 + jj;
    const double xj = st[j], yj = st[ld + j], zj = st[2 * (size_t)ld + j];
    if (ok && (I != J || lane < jj)) {
      double dx = xi - xj, dy = yi - yj, dz = zi - zj;
      dd_add(s, 1.0 / sqrt(dx * dx + dy * dy + dz * dz));
    }
  }
  s = dd_wave_sum(s);
  if (lane == 0) part[blockIdx.x] = s;
}

// ten linear sums per block: Xa Ya Za Ua Va Wa Lx Ly Lz KE
__global__ __launch_bounds__(256) void k_inv_linear(int n, int ld, const double* __restrict__ st,
                                                    DD* __restrict__ part) {
  __shared__ DD red[4][10];
  DD a[10];
#pragma unroll
  for (int k = 0; k < 10; k++) a[k] = DD{0, 0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double x = st[i], y = st[ld + i], z = st[2 * (size_t)ld + i];
    double u = st[3 * (size_t)ld + i], v = st[4 * (size_t)ld + i], w = st[5 * (size_t)ld + i];
    dd_add(a[0], x); dd_add(a[1], y); dd_add(a[2], z);
    dd_add(a[3], u); dd_add(a[4], v); dd_add(a[5], w);
    dd_add(a[6], y * w - z * v);
    dd_add(a[7], z * u - x * w);
    dd_add(a[8], x * v - y * u);
    dd_add(a[9], u * u + v * v + w * w);
  }
  const int wv = threadIdx.x / 64, lane = threadIdx.x % 64;
#pragma unroll
  for (int k = 0; k < 10; k++) {
    DD r = dd_wave_sum(a[k]);
    if (lane == 0) red[wv][k] = r;
  }
  __syncthreads();
  if (threadIdx.x < 10) {
    DD r = red[0][threadIdx.x];
    for (int w2 = 1; w2 < 4; w2++) dd_add(r, red[w2][threadIdx.x]);
    part[(size_t)blockIdx.x * 10 + threadIdx.x] = r;
  }
}

// first stage of the potential's reduction: block b sums tile partials [b*chunk, (b+1)*chunk) in a
// fixed order (thread stride, then the block's waves in order) into out[b]
__global__ __launch_bounds__(256) void k_inv_pe_reduce(int ntile, int chunk, const DD* __restrict__ pe,
                                                       DD* __restrict__ out) {
  __shared__ DD red[4];
  DD a{0, 0};
  const int b0 = blockIdx.x * chunk, b1 = min(ntile, b0 + chunk);
  for (int b = b0 + (int)threadIdx.x; b < b1; b += blockDim.x) dd_add(a, pe[b]);
  a = dd_wave_sum(a);
  const int wv = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0) red[wv] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    DD r = red[0];
    for (int w2 = 1; w2 < 4; w2++) dd_add(r, red[w2]);
    out[blockIdx.x] = r;
  }
}

// single-wave final reduction in a fixed order; out = KE PE Lx Ly Lz Ua Va Wa Xa Ya Za
__global__ __launch_bounds__(64) void k_inv_final(int nlin, const DD* __restrict__ lin, int ntile,
                                                  const DD* __restrict__ pe, double eps, double* __restrict__ out) {
  const int lane = threadIdx.x;
  DD a[11];
#pragma unroll
  for (int k = 0; k < 11; k++) a[k] = DD{0, 0};
  for (int b = lane; b < nlin; b += 64)
#pragma unroll
    for (int k = 0; k < 10; k++) dd_add(a[k], lin[(size_t)b * 10 + k]);
  for (int b = lane; b < ntile; b += 64) dd_add(a[10], pe[b]);
#pragma unroll
  for (int k = 0; k < 11; k++) a[k] = dd_wave_sum(a[k]);
  if (lane == 0) {
    double v[11];
    for (int k = 0; k < 11; k++) v[k] = a[k].hi + a[k].lo;
    // eps * PE in double-double: product error term via fma
    double p = eps * a[10].hi;
    double pe_ = p + (fma(eps, a[10].hi, -p) + eps * a[10].lo);
    out[0] = v[9];
    out[1] = pe_;
    out[2] = v[6]; out[3] = v[7]; out[4] = v[8];
    out[5] = v[3]; out[6] = v[4]; out[7] = v[5];
    out[8] = v[0]; out[9] = v[1]; out[10] = v[2];
  }
}

static int invariants_locked(bia_ctx* ctx, int n, double eps, const double* d_state, int ld, double* d_out) {
  if (n < 1 || ld < n || !d_state || !d_out) return BIA_ERR_ARG;
  const int nt = (n + 63) / 64;
  const size_t ntiles = (size_t)nt * (nt + 1) / 2;
  const int nlin = 256, npe = 256;
  BIA_TRY(ctx, ensure_tiles(ctx, nt));
  BIA_TRY(ctx, ctx->inv_part.ensure((ntiles + (size_t)nlin * 10 + npe) * sizeof(DD)));
  DD* pe = ctx->inv_part.as<DD>();
  DD* lin = pe + ntiles;
  DD* pe2 = lin + (size_t)nlin * 10;
  hipLaunchKernelGGL(k_inv_pe_tiles, dim3((unsigned)ntiles), dim3(64), 0, ctx->stream, n, ld, d_state,
                     ctx->tiles.as<int2>(), pe);
  BIA_HIP_TRY(ctx, hipGetLastError());
  const int chunk = (int)((ntiles + npe - 1) / npe);
  hipLaunchKernelGGL(k_inv_pe_reduce, dim3(npe), dim3(256), 0, ctx->stream, (int)ntiles, chunk, pe, pe2);
  BIA_HIP_TRY(ctx, hipGetLastError());
  hipLaunchKernelGGL(k_inv_linear, dim3(nlin), dim3(256), 0, ctx->stream, n, ld, d_state, lin);
  BIA_HIP_TRY(ctx, hipGetLastError());
  hipLaunchKernelGGL(k_inv_final, dim3(1), dim3(64), 0, ctx->stream, nlin, lin, npe, pe2, eps, d_out);
  BIA_HIP_TRY(ctx, hipGetLastError());
  return BIA_OK;
}

static int invariants_host_locked(bia_ctx* ctx, int n, double eps, const double* const in[6], double* out) {
  BIA_TRY(ctx, ctx->inv_state.ensure((size_t)6 * n * sizeof(double) + 16 * sizeof(double)));
  double* d = ctx->inv_state.as<double>();
  double* dout = d + (size_t)6 * n;
  for (int c = 0; c < 6; c++)
    BIA_HIP_TRY(ctx, hipMemcpyAsync(d + (size_t)c * n, in[c], (size_t)n * sizeof(double), hipMemcpyHostToDevice,
                                    ctx->stream));
  BIA_TRY(ctx, invariants_locked(ctx, n, eps, d, n, dout));
  BIA_HIP_TRY(ctx, hipMemcpyAsync(out, dout, 11 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  BIA_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return BIA_OK;
}

}  // namespace bia

using namespace bia;

extern "C" {

int bia_invariants_dev(bia_ctx* ctx, int n, double eps, const double* d_state, int ld, double* d_out) {
  if (!ctx) return BIA_ERR_ARG;
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->last_error[0] = 0;
  BIA_HIP_TRY(ctx, hipSetDevice(ctx->device));
  return invariants_locked(ctx, n, eps, d_state, ld, d_out);
}

int bia_invariants_host(bia_ctx* ctx, int n, double eps, const double* X, const double* Y, const double* Z,
                        const double* U, const double* V, const double* W, double* out) {
  if (!ctx || !X || !Y || !Z || !U || !V || !W || !out || n < 1) return BIA_ERR_ARG;
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->last_error[0] = 0;
  BIA_HIP_TRY(ctx, hipSetDevice(ctx->device));
  const double* in[6] = {X, Y, Z, U, V, W};
  return invariants_host_locked(ctx, n, eps, in, out);
}

void invariance_gpu_(int N, double eps_th, double* X, double* Y, double* Z, double* U, double* V, double* W,
                     double* KE, double* PE, double* Lx, double* Ly, double* Lz, double* Ua, double* Va,
                     double* Wa, double* Xa, double* Ya, double* Za) {
  bia_ctx* ctx = default_ctx();
  double o[11];
  int s = bia_invariants_host(ctx, N, eps_th, X, Y, Z, U, V, W, o);
  if (s != BIA_OK) legacy_fail("invariance_gpu_", s, ctx);
  *KE = o[0]; *PE = o[1];
  *Lx = o[2]; *Ly = o[3]; *Lz = o[4];
  *Ua = o[5]; *Va = o[6]; *Wa = o[7];
  *Xa = o[8]; *Ya = o[9]; *Za = o[10];
}

}  // extern "C"
