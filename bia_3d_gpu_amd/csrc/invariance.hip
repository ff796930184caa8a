// invariance.hip -- energy / momentum / angular-momentum invariants (reference
// bia/invariance.cu:15-110, CPU twins bia/Invariance_CPU.f90, bia/Invariance_QPU.f90).
//
//   KE = sum |v|^2, PE = eps * sum_{i<j} 1/|r_i - r_j|, L = sum r x v, sums of x..w
//
// The O(N^2) potential term is computed over 64x64 tiles of the i<j triangle (one wave per
// tile, lane = row, column particle wave-uniform); a list of systems side by side in one state
// (an ensemble) is one launch of each kernel, tiles of all systems in one grid.  Every lane accumulates in double-double
// (TwoSum), tile partials are written to a buffer and summed in a fixed order (256 blocks over
// contiguous chunks, then one wave), so the result
// is deterministic and close to the quad-precision sum of Invariance_QPU.f90 rather than to
// the plain double sum.  The ten O(N) sums use the same double-double scheme.
#include <algorithm>
#include <climits>
#include <cstring>

#include "bia_pair.hpp"
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

// One system of a list: its columns in the state and where its partial sums go.
struct InvSys {
  long long col;        // first column of the system in the [6, ld] state
  int n;                // particles
  long long tile0;      // first of its potential tiles in the global tile numbering
  int nt;               // 64-wide particle tiles (ntiles = nt (nt + 1) / 2)
  int chunk, npe, pe0;  // first reduction stage: npe blocks of chunk tiles, results at pe0
  int nlin, lin0;       // linear sums: nlin blocks, results at lin0 (10 each)
  double eps;
};

// (I, J) of tile t of the row-major upper triangle of nt tiles (row I starts at I nt - I (I-1)/2)
__device__ __forceinline__ int2 tri_tile(long long t, int nt) {
  const double b = 2.0 * nt + 1;
  int I = (int)((b - sqrt(b * b - 8.0 * (double)t)) * 0.5);
  I = max(0, min(I, nt - 1));
  auto start = [nt](long long r) { return r * nt - r * (r - 1) / 2; };
  while (I > 0 && start(I) > t) --I;
  while (I + 1 < nt && start(I + 1) <= t) ++I;
  return make_int2(I, I + (int)(t - start(I)));
}

// the system owning global tile t: the last s with sys[s].tile0 <= t
__device__ __forceinline__ int sys_of_tile(const InvSys* __restrict__ sys, int nsys, long long t) {
  int lo = 0, hi = nsys - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sys[mid].tile0 <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// potential tile t: part[t] = sum over the tile's pairs of 1/r, or of pw_i pw_j / r with W
// (double-double)
template <bool W>
__device__ __forceinline__ void pe_tile(long long t, int lane, int nsys, const InvSys* __restrict__ sys, int ld,
                                        const double* __restrict__ state, const double* __restrict__ pw,
                                        DD* __restrict__ part) {
  const InvSys S = sys[sys_of_tile(sys, nsys, t)];
  const int2 IJ = tri_tile(t - S.tile0, S.nt);
  const int I = IJ.x, J = IJ.y, n = S.n;
  const double* st = state + S.col;
  const int i = I * 64 + lane;
  const bool ok = i < n;
  const double xi = ok ? st[i] : 0, yi = ok ? st[ld + i] : 0, zi = ok ? st[2 * (size_t)ld + i] : 0;
  const double wi = (W && ok) ? pw[S.col + i] : 1.0;
  DD s{0, 0};
  const int j0 = J * 64, jn = min(64, n - j0);
  for (int jj = 0; jj < jn; ++jj) {
    const int j = j0 + jj;
    const double xj = st[j], yj = st[ld + j], zj = st[2 * (size_t)ld + j];
    if (ok && (I != J || lane < jj)) {
      double dx = xi - xj, dy = yi - yj, dz = zi - zj;
      const double r2 = dx * dx + dy * dy + dz * dz;
      // 1/r within an ulp from the hardware estimate and one correction (the reference takes
      // rsqrt, bia/invariance.cu:23); coincident particles give +inf as 1/sqrt(0) does
      const double ir = r2 > 0 ? fast_rsqrt(r2) : __builtin_huge_val();
      dd_add(s, W ? (wi * pw[S.col + j]) * ir : ir);
    }
  }
  s = dd_wave_sum(s);
  if (lane == 0) part[t] = s;
}

// all potential tiles of a list of systems, one wave per tile, grid-stride (the grid is capped
// below 2^32 work-items)
template <bool W>
__global__ __launch_bounds__(64) void k_inv_pe_tiles(int nsys, long long ntiles, const InvSys* __restrict__ sys,
                                                     int ld, const double* __restrict__ state,
                                                     const double* __restrict__ pw, DD* __restrict__ part) {
  const int lane = threadIdx.x;
  for (long long t = blockIdx.x; t < ntiles; t += gridDim.x) pe_tile<W>(t, lane, nsys, sys, ld, state, pw, part);
}

// ten linear sums per block (Xa Ya Za Ua Va Wa Lx Ly Lz KE), each term weighted by kw_i with W;
// block b of a system takes particles b*256 + thread, stepping by kLinStride, so the partition
// does not depend on how many blocks of the grid are empty
constexpr int kLinBlocks = 256;
constexpr int kLinStride = kLinBlocks * 256;
template <bool W>
__global__ __launch_bounds__(256) void k_inv_linear(const InvSys* __restrict__ sys, int ld,
                                                    const double* __restrict__ state, const double* __restrict__ kw,
                                                    DD* __restrict__ part) {
  const InvSys S = sys[blockIdx.y];
  if ((int)blockIdx.x >= S.nlin) return;
  const double* st = state + S.col;
  const int n = S.n;
  __shared__ DD red[4][10];
  DD a[10];
#pragma unroll
  for (int k = 0; k < 10; k++) a[k] = DD{0, 0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += kLinStride) {
    double x = st[i], y = st[ld + i], z = st[2 * (size_t)ld + i];
    double u = st[3 * (size_t)ld + i], v = st[4 * (size_t)ld + i], w = st[5 * (size_t)ld + i];
    if (W) {
      const double m = kw[S.col + i];
      dd_add(a[0], m * x); dd_add(a[1], m * y); dd_add(a[2], m * z);
      dd_add(a[3], m * u); dd_add(a[4], m * v); dd_add(a[5], m * w);
      dd_add(a[6], m * (y * w - z * v));
      dd_add(a[7], m * (z * u - x * w));
      dd_add(a[8], m * (x * v - y * u));
      dd_add(a[9], m * (u * u + v * v + w * w));
    } else {
      dd_add(a[0], x); dd_add(a[1], y); dd_add(a[2], z);
      dd_add(a[3], u); dd_add(a[4], v); dd_add(a[5], w);
      dd_add(a[6], y * w - z * v);
      dd_add(a[7], z * u - x * w);
      dd_add(a[8], x * v - y * u);
      dd_add(a[9], u * u + v * v + w * w);
    }
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
    part[((size_t)S.lin0 + blockIdx.x) * 10 + threadIdx.x] = r;
  }
}

// first stage of the potential's reduction: block b of a system sums its tile partials
// [b*chunk, (b+1)*chunk) in a fixed order (thread stride, then the block's waves in order)
__global__ __launch_bounds__(256) void k_inv_pe_reduce(const InvSys* __restrict__ sys, const DD* __restrict__ pe,
                                                       DD* __restrict__ out) {
  const InvSys S = sys[blockIdx.y];
  if ((int)blockIdx.x >= S.npe) return;
  __shared__ DD red[4];
  DD a{0, 0};
  const long long ntiles = (long long)S.nt * (S.nt + 1) / 2;
  const long long b0 = (long long)blockIdx.x * S.chunk, b1 = min(ntiles, b0 + S.chunk);
  for (long long b = b0 + threadIdx.x; b < b1; b += blockDim.x) dd_add(a, pe[S.tile0 + b]);
  a = dd_wave_sum(a);
  const int wv = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0) red[wv] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    DD r = red[0];
    for (int w2 = 1; w2 < 4; w2++) dd_add(r, red[w2]);
    out[S.pe0 + blockIdx.x] = r;
  }
}

// one wave per system, final reduction in a fixed order; out[11 s ..] = KE PE Lx Ly Lz Ua Va Wa Xa Ya Za
__global__ __launch_bounds__(64) void k_inv_final(const InvSys* __restrict__ sys, const DD* __restrict__ lin,
                                                  const DD* __restrict__ pe, double* __restrict__ out_all) {
  const InvSys S = sys[blockIdx.x];
  double* out = out_all + 11 * (size_t)blockIdx.x;
  const int lane = threadIdx.x;
  DD a[11];
#pragma unroll
  for (int k = 0; k < 11; k++) a[k] = DD{0, 0};
  for (int b = lane; b < S.nlin; b += 64)
#pragma unroll
    for (int k = 0; k < 10; k++) dd_add(a[k], lin[((size_t)S.lin0 + b) * 10 + k]);
  for (int b = lane; b < S.npe; b += 64) dd_add(a[10], pe[S.pe0 + b]);
#pragma unroll
  for (int k = 0; k < 11; k++) a[k] = dd_wave_sum(a[k]);
  if (lane == 0) {
    double v[11];
    for (int k = 0; k < 11; k++) v[k] = a[k].hi + a[k].lo;
    // eps * PE in double-double: product error term via fma
    const double eps = S.eps;
    double p = eps * a[10].hi;
    double pe_ = p + (fma(eps, a[10].hi, -p) + eps * a[10].lo);
    out[0] = v[9];
    out[1] = pe_;
    out[2] = v[6]; out[3] = v[7]; out[4] = v[8];
    out[5] = v[3]; out[6] = v[4]; out[7] = v[5];
    out[8] = v[0]; out[9] = v[1]; out[10] = v[2];
  }
}

// Descriptors of the systems (uploaded when the sizes or eps change).  The first reduction stage
// of a system uses ceil(ntiles / chunk) blocks with chunk = ceil(ntiles / 256), and the linear sums
// min(256, ceil(n / 256)) blocks: the non-empty blocks of a fixed 256-block partition, so a
// system's result does not depend on the other systems of the list.
static int ensure_inv_plan(bia_ctx* ctx, int nsys, const int* sizes, const double* eps) {
  std::vector<long long> key(2 * (size_t)nsys);
  for (int s = 0; s < nsys; s++) {
    key[2 * s] = sizes[s];
    memcpy(&key[2 * s + 1], &eps[s], 8);
  }
  if (key == ctx->inv_key) return BIA_OK;
  std::vector<InvSys> h(nsys);
  long long col = 0, tile0 = 0;
  int pe0 = 0, lin0 = 0, maxnpe = 0, maxnlin = 0;
  for (int s = 0; s < nsys; s++) {
    InvSys& S = h[s];
    S.col = col;
    S.n = sizes[s];
    S.nt = (S.n + 63) / 64;
    const long long ntiles = (long long)S.nt * (S.nt + 1) / 2;
    S.tile0 = tile0;
    S.chunk = (int)((ntiles + kLinBlocks - 1) / kLinBlocks);
    S.npe = (int)((ntiles + S.chunk - 1) / S.chunk);
    S.pe0 = pe0;
    S.nlin = std::min(kLinBlocks, (S.n + 255) / 256);
    S.lin0 = lin0;
    S.eps = eps[s];
    col += S.n;
    tile0 += ntiles;
    pe0 += S.npe;
    lin0 += S.nlin;
    maxnpe = std::max(maxnpe, S.npe);
    maxnlin = std::max(maxnlin, S.nlin);
  }
  if (tile0 > INT_MAX) return BIA_ERR_ARG;  // tile counts are held in int (about 4.2 million particles)
  ctx->inv_key.clear();  // invalid until the upload below succeeds
  // the copy is synchronous and not ordered after kernels still in flight on a non-blocking
  // stream, which read these descriptors: drain the context's stream first
  BIA_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  BIA_TRY(ctx, ctx->inv_sys.ensure(h.size() * sizeof(InvSys)));
  BIA_HIP_TRY(ctx, hipMemcpy(ctx->inv_sys.p, h.data(), h.size() * sizeof(InvSys), hipMemcpyHostToDevice));
  ctx->inv_ntiles = (int)tile0;
  ctx->inv_npe = pe0;
  ctx->inv_nlin = lin0;
  ctx->inv_maxnpe = maxnpe;
  ctx->inv_maxnlin = maxnlin;
  ctx->inv_key = std::move(key);
  return BIA_OK;
}

// invariants of nsys systems side by side in a [6, ld] state (columns in the order of sizes);
// kw, pw: per-particle weights of the linear sums and of the potential (device, same columns as
// the state), or NULL for the reference's unweighted sums
static int invariants_locked(bia_ctx* ctx, int nsys, const int* sizes, const double* eps, const double* d_state,
                             int ld, double* d_out, const double* kw = nullptr, const double* pw = nullptr) {
  if (nsys < 1 || !sizes || !eps || !d_state || !d_out) return BIA_ERR_ARG;
  long long total = 0;
  for (int s = 0; s < nsys; s++) {
    if (sizes[s] < 1) return BIA_ERR_ARG;
    total += sizes[s];
  }
  if (total > ld || nsys > 65535) return BIA_ERR_ARG;  // nsys is the grid's y extent
  BIA_TRY(ctx, ensure_inv_plan(ctx, nsys, sizes, eps));
  const InvSys* sys = ctx->inv_sys.as<InvSys>();
  BIA_TRY(ctx, ctx->inv_part.ensure(((size_t)ctx->inv_ntiles + (size_t)ctx->inv_nlin * 10 + ctx->inv_npe) *
                                    sizeof(DD)));
  DD* pe = ctx->inv_part.as<DD>();
  DD* lin = pe + ctx->inv_ntiles;
  DD* pe2 = lin + (size_t)ctx->inv_nlin * 10;
  const long long nt = ctx->inv_ntiles;
  const unsigned grid = (unsigned)std::min<long long>(nt, 0xffffffffLL / 64);
  if (pw)
    hipLaunchKernelGGL(k_inv_pe_tiles<true>, dim3(grid), dim3(64), 0, ctx->stream, nsys, nt, sys, ld, d_state, pw,
                       pe);
  else
    hipLaunchKernelGGL(k_inv_pe_tiles<false>, dim3(grid), dim3(64), 0, ctx->stream, nsys, nt, sys, ld, d_state,
                       nullptr, pe);
  BIA_HIP_TRY(ctx, hipGetLastError());
  hipLaunchKernelGGL(k_inv_pe_reduce, dim3(ctx->inv_maxnpe, nsys), dim3(256), 0, ctx->stream, sys, pe, pe2);
  BIA_HIP_TRY(ctx, hipGetLastError());
  if (kw)
    hipLaunchKernelGGL(k_inv_linear<true>, dim3(ctx->inv_maxnlin, nsys), dim3(256), 0, ctx->stream, sys, ld, d_state,
                       kw, lin);
  else
    hipLaunchKernelGGL(k_inv_linear<false>, dim3(ctx->inv_maxnlin, nsys), dim3(256), 0, ctx->stream, sys, ld,
                       d_state, nullptr, lin);
  BIA_HIP_TRY(ctx, hipGetLastError());
  hipLaunchKernelGGL(k_inv_final, dim3(nsys), dim3(64), 0, ctx->stream, sys, lin, pe2, d_out);
  BIA_HIP_TRY(ctx, hipGetLastError());
  return BIA_OK;
}

static int invariants_host_locked(bia_ctx* ctx, int n, double eps, const double* const in[6], double* out) {
  StreamDrain drain{ctx};  // the copies below read and write the caller's host arrays
  BIA_TRY(ctx, ctx->inv_state.ensure((size_t)6 * n * sizeof(double) + 16 * sizeof(double)));
  double* d = ctx->inv_state.as<double>();
  double* dout = d + (size_t)6 * n;
  for (int c = 0; c < 6; c++)
    BIA_HIP_TRY(ctx, hipMemcpyAsync(d + (size_t)c * n, in[c], (size_t)n * sizeof(double), hipMemcpyHostToDevice,
                                    ctx->stream));
  BIA_TRY(ctx, invariants_locked(ctx, 1, &n, &eps, d, n, dout));
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
  return invariants_locked(ctx, 1, &n, &eps, d_state, ld, d_out);
}

int bia_invariants_batched_dev(bia_ctx* ctx, int nsys, const int* sizes, const double* eps, const double* d_state,
                               int ld, double* d_out) {
  if (!ctx) return BIA_ERR_ARG;
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->last_error[0] = 0;
  BIA_HIP_TRY(ctx, hipSetDevice(ctx->device));
  return invariants_locked(ctx, nsys, sizes, eps, d_state, ld, d_out);
}

int bia_invariants_coupled_batched_dev(bia_ctx* ctx, int coupling, double coup, int nsys, const int* sizes,
                                       const double* d_state, int ld, const double* d_mass, const double* d_charge,
                                       double* d_out) {
  if (!ctx) return BIA_ERR_ARG;
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->last_error[0] = 0;
  if (nsys < 1 || nsys > 65535 ||
      !(coupling == BIA_COUPLING_GRAVITY ? d_mass != nullptr : coupling == BIA_COUPLING_COULOMB && d_charge))
    return BIA_ERR_ARG;
  BIA_HIP_TRY(ctx, hipSetDevice(ctx->device));
  // KE + PE = 2 E: gravity PE = -2 G sum m_i m_j / r, Coulomb PE = 2 k sum q_i q_j / r
  const std::vector<double> eps(nsys, coupling == BIA_COUPLING_GRAVITY ? -2.0 * coup : 2.0 * coup);
  return invariants_locked(ctx, nsys, sizes, eps.data(), d_state, ld, d_out, d_mass,
                           coupling == BIA_COUPLING_GRAVITY ? d_mass : d_charge);
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
