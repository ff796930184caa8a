// bia_step.hip -- whole-step pair kernels (ALG / RKF45) for one system or a batch of systems,
// the multi-step driver (hipGraph), the launch autotuner, and the reference-compatible
// per-particle entry point find_dxyuv_kernel_.
//
// Tiling.  The i<j triangle of each system is cut into 64x64 tiles (I <= J), and the row tiles
// into blocks of kBlk = 4 (block R: row tiles 4R .. 4R+3, particles 256R .. 256R+255).  A work item
// is (block R, segment k): the column tiles J in [4R + 4k, 4R + 4k + 4).  One workgroup of four
// waves takes an item; wave w holds row tile I = 4R + w, lane l row particle i = 64 I + l in
// registers, and the four waves walk the item's column tiles together: for each J, every wave
// with a tile (I <= J) computes it, its column sums go to its own LDS row, and after a barrier
// the four rows are added in wave order into one column partial per (block, column).  The row
// sums stay in registers across the item's four column tiles.  So every slot of the per-system
// partial buffer has exactly one writer:
//   block R, column j >= 256 R : colpart(R)[c][j]   (the block's four row tiles, added in order)
//   block R, item k, row i     : rowpart(R)[k][c][i]
// with no atomics and nothing to clear, and a second kernel sums, for particle i of block R, the
// column partials of blocks 0..R and then its own block's row partials, in that fixed order
// (deterministic).  The drift v dt is added there.  Against one partial per tile and side, a
// block of four row tiles and a segment of four column tiles write a quarter of the partials.
//
// Batching.  The items of all systems form one list (R, k, system); the pair kernel is a
// persistent grid-stride launch over it, so a batch of systems of uneven N is one launch.
//
// Couplings.  SCALAR: the reference's r'' = eps r/|r|^3 with eps per system and the equal-mass
// half split (find_dXYUV_kernel.cu:510-515).  GRAVITY: eps_ij = -G (m_i + m_j).  COULOMB:
// eps_ij = k q_i q_j (1/m_i + 1/m_j).  With masses the relative change is split by the centre of
// mass: m_j/(m_i+m_j) to i and m_i/(m_i+m_j) to j.
#include <string.h>

#include <algorithm>
#include <cstddef>
#include <cmath>

#include "bia_pair.hpp"
#include "bia_runtime.hpp"

namespace bia {

constexpr int kWave = 64;
constexpr int kBlk = 4;                 // row tiles of a block = waves of a pair-kernel workgroup
constexpr int kGroup = kWave * kBlk;    // threads of a pair-kernel workgroup = particles of a block
constexpr int kStatSlots = 4096;  // statistics slots of the pair kernel (a power of two)
constexpr int kMaxGrid = (int)(0xffffffffu / kGroup);  // largest grid of pair-kernel workgroups

// Work items and partial buffer of one system of nt tiles (npad = 64 nt, nb = ceil(nt / 4) blocks),
// with segments of seg column tiles (seg divides kBlk): block R has nseg(R) = ceil(nt / seg) -
// (4 / seg) R items, the segments of its column tiles J in [4R, nt).  Block R holds its column
// partials colpart[6][W_R] for the columns j in [256 R, npad), W_R = npad - 256 R, then its row
// partials rowpart[nseg(R)][6][256] (one per item).  off(R) is where block R starts, off(nb) the
// system's total (with seg = 4 about 12 N nt / 8 doubles: a quarter of one partial per tile and
// side).
struct Layout {
  int nt, npad, nb, seg, ns0, step;  // ns0 = ceil(nt / seg) items in block 0, step = 4 / seg fewer per block
  __host__ __device__ constexpr Layout(int nt_, int seg_)
      : nt(nt_), npad(nt_ * kWave), nb((nt_ + kBlk - 1) / kBlk), seg(seg_), ns0((nt_ + seg_ - 1) / seg_),
        step(kBlk / seg_) {}
  __host__ __device__ constexpr long long w(int R) const { return (long long)npad - (long long)kGroup * R; }
  __host__ __device__ constexpr int nseg(int R) const { return ns0 - step * R; }
  __host__ __device__ constexpr long long items_before(int R) const {
    return (long long)R * ns0 - (long long)step * R * (R - 1) / 2;
  }
  __host__ __device__ constexpr long long off(int R) const {
    const long long r = R;
    return 6 * (r * npad - (long long)(kGroup / 2) * r * (r - 1)) + 6LL * kGroup * items_before(R);
  }
};

// segment length of a system: the longest (fewest partials) that still gives it 8192 items or more
// (about eight four-wave workgroups per CU), else single column tiles (small systems: a segment
// of four tiles would leave most CUs idle and serialise four tiles per wave)
constexpr int kMinItems = 8192;
__host__ __device__ constexpr int choose_seg(int nt) {
  return Layout(nt, 4).items_before(Layout(nt, 4).nb) >= kMinItems ? 4
         : Layout(nt, 2).items_before(Layout(nt, 2).nb) >= kMinItems ? 2 : 1;
}

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
  // Kepler modes: this tile's pair counts and evaluations per branch in 16-bit fields of tc and
  // ti (a lane takes at most 64 pairs of a tile and at most 60 evaluations per pair, so a field
  // holds at most 3840), folded into cnt/its once per tile
  unsigned long long tc, ti;
  __device__ void clear() {
    for (int k = 0; k < 4; k++) cnt[k] = its[k] = 0;
    for (int k = 0; k < 8; k++) fl[k] = 0;
    tc = ti = 0;
  }
  __device__ void add(int br, int it, unsigned f) {
    for (int k = 0; k < 4; k++) {
      cnt[k] += (br == k);
      its[k] += (br == k) ? (unsigned)it : 0u;
    }
    for (int k = 0; k < 8; k++) fl[k] += (f >> k) & 1u;
  }
  // br in [0, 4), it <= 60; flags are rare on these paths and go straight to their totals
  __device__ void add_tile(int br, int it, unsigned f) {
    const unsigned sh = 16u * (unsigned)br;
    tc += 1ull << sh;
    ti += (unsigned long long)(unsigned)it << sh;
    if (f)
      for (int k = 0; k < 8; k++) fl[k] += (f >> k) & 1u;
  }
  __device__ void fold_tile() {
    for (int k = 0; k < 4; k++) {
      cnt[k] += (unsigned)(tc >> (16 * k)) & 0xffffu;
      its[k] += (unsigned)(ti >> (16 * k)) & 0xffffu;
    }
    tc = ti = 0;
  }
  // the workgroup's totals into its slot: workgroup b uses slot b mod kStatSlots, which spreads a
  // launch's atomics over kStatSlots x 16 addresses instead of 16, and the four waves' totals are
  // added in LDS first (red: kBlk x 16), so a workgroup makes 16 atomics, not 64, all on one
  // cache line; k_stats_fold adds the slots up.  Every thread of the workgroup calls this.
  __device__ void flush(unsigned long long* slots, unsigned (*red)[16], int w) {
    unsigned v[16];
    for (int k = 0; k < 4; k++) {
      v[k] = wave_sum_u(cnt[k]);
      v[4 + k] = wave_sum_u(its[k]);
    }
    for (int k = 0; k < 8; k++) v[8 + k] = wave_sum_u(fl[k]);
    if (threadIdx.x % kWave == 0)
      for (int k = 0; k < 16; k++) red[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 16) {
      unsigned long long t = 0;
      for (int q = 0; q < kBlk; q++) t += red[q][threadIdx.x];
      unsigned long long* out = slots + (size_t)(blockIdx.x & (kStatSlots - 1)) * BIA_STATS_LEN;
      if (t) atomicAdd(out + threadIdx.x, t);
    }
  }
};

// stats[k] += sum over the slots of counter k, and the slots back to zero; one block per counter
__global__ __launch_bounds__(256) void k_stats_fold(unsigned long long* __restrict__ slots,
                                                    unsigned long long* __restrict__ stats) {
  __shared__ unsigned long long red[256];
  const int k = blockIdx.x;
  unsigned long long s = 0;
  for (int i = threadIdx.x; i < kStatSlots; i += 256) {
    unsigned long long* p = slots + (size_t)i * BIA_STATS_LEN + k;
    s += *p;
    *p = 0;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) stats[k] += red[0];
}

__global__ __launch_bounds__(256) void k_zero_u64(unsigned long long* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

enum Coupling : int { kScalar = 0, kGravity = 1, kCoulomb = 2 };

// Tile partials are written once and read once, by k_reduce after the launch: streaming stores,
// so that 3 GB of them do not evict the particle states the tiles re-read from L2.
__device__ __forceinline__ void put_partial(double* p, double v) { __builtin_nontemporal_store(v, p); }

// LDS add without return (ds_add_f64): the add runs in the LDS unit, not on the VALU.  The LDS
// operations of one wave complete in issue order, so a column sum built from one add per step
// is the sequential sum in step order, rounded like v_add_f64.
__device__ __forceinline__ void lds_add(double* p, double x) {
  __hip_atomic_fetch_add(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// OCC: minimum waves per SIMD asked of the register allocator (1 = its own choice; 4 = at most 128
// VGPRs; 5 = at most 96, for the Kepler-drift modes); picked per shape by the autotuner.
//
// One workgroup of four waves per work item (block R, segment k); wave w holds row tile
// I = 4R + w, lane l row i = 64 I + l.  For each column tile J of the segment, every wave with a
// tile (I <= J) computes it:
//  * reference solvers (plain modes, RKF45): divergent and register-heavy, so the columns are
//    walked in order (j wave-uniform, its state through scalar loads) and each column sum is a
//    wave reduction, kept by lane j; the wave's LDS residual history is free between tiles, and
//    its first six rows take the tile's column sums;
//  * Kepler modes: the column tile's states are staged in LDS by the whole workgroup and at step
//    k lane l takes column j = (l + k) mod 64, so the 64 lanes work on 64 different columns and no
//    column needs a cross-lane reduction: each lane adds its pair's column share into the
//    column's sum in its wave's LDS row (lds_add; one address per lane per step, the steps in
//    order).  A diagonal tile (I = J) takes steps k = 1..32 (at k = 32 lanes 0..31 only), which
//    covers each unordered pair once.
// Then the four waves' column sums are added in wave order into the block's column partial of J.
// Stage: the items [t0, t1) of the list this launch takes; staged (one system), the partials of
// blocks g0 and up, relative to block g0 (k_reduce_stage sums them in the order k_reduce would).
struct Stage {
  int t0, t1, g0, staged;
};

template <int METHOD, int MODE, int COUP, bool STATS, int OCC>
__global__ __launch_bounds__(kGroup, OCC) void k_tiles(Stage stg, int lds, const int4* __restrict__ items,
                                                  const SysDesc* __restrict__ sys, const double* __restrict__ st,
                                                  const double* __restrict__ mass, const double* __restrict__ charge,
                                                  double coup, double dt, int itr, double* __restrict__ part,
                                                  unsigned long long* __restrict__ stats, LearnTable T, RkfTol tol,
                                                  PairLog lg) {
  constexpr bool REF = reference_solvers(MODE);
  // per wave: the reference iteration's residual history (reference solvers; its first six rows
  // take the wave's column sums between tiles), or the column sums (Kepler modes)
  __shared__ double wrow[kBlk][REF ? kMaxItr : 6][kWave];
  __shared__ double cst[6][kWave];  // column tile state (Kepler modes)
  __shared__ double cmq[3][kWave];  // column masses, charges and reciprocal masses (couplings only)
  // the wave's index, in a scalar register (it is the same for the wave's 64 lanes)
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x % kWave;
  Hist hist{&wrow[w][0][lane], kWave};
  PairStats ps;
  if (STATS) ps.clear();
  for (int t = stg.t0 + blockIdx.x; t < stg.t1; t += gridDim.x) {
    const int4 item = items[t];
    const SysDesc S = sys[item.z];
    const int R = item.x, seg = item.y;
    const int I = kBlk * R + w;
    const bool wave_ok = I < S.nt;  // wave-uniform: the last block may hold fewer than four row tiles
    const int il = I * kWave + lane;  // local row index
    const bool row_ok = wave_ok && il < S.n;
    const int gi = S.off + il;
    double ri[6];
#pragma unroll
    for (int c = 0; c < 6; c++) ri[c] = row_ok ? st[(size_t)c * lds + gi] : 0.0;
    double mi = 1, qi = 0, rmi = 1;
    if (COUP != kScalar) {
      mi = (row_ok && mass) ? mass[gi] : 1.0;
      if (COUP == kCoulomb) {
        qi = row_ok ? charge[gi] : 0.0;
        rmi = 1 / mi;
      }
    }
    const Layout L(S.nt, S.seg);
    double* B = part + (stg.staged ? L.off(R) - L.off(stg.g0) : S.poff + L.off(R));
    const long long WR = L.w(R);
    const int Ja = kBlk * R + L.seg * seg, Jb = min(Ja + L.seg, S.nt);
    double acc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 1
    for (int J = Ja; J < Jb; ++J) {
      const int j0 = J * kWave;
      const int jn = min(kWave, S.n - j0);
      const bool tile = wave_ok && I <= J;  // wave-uniform
      // the last wave with a tile in column tile J
      const int wlast = min(min(kBlk - 1, J - kBlk * R), S.nt - 1 - kBlk * R);
      if constexpr (REF) {
        if (tile) {
          double col[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 1
          for (int jj = 0; jj < jn; ++jj) {
            const int gj = S.off + j0 + jj;
            double rj[6];
#pragma unroll
            for (int c = 0; c < 6; c++) rj[c] = st[(size_t)c * lds + gj];
            double d[6] = {0, 0, 0, 0, 0, 0};
            double wi = 1, wj = 1;
            if (row_ok && (I != J || lane < jj)) {
              double eps = S.eps;
              if (COUP != kScalar) {
                const double mj = mass ? mass[gj] : 1.0;
                if (COUP == kGravity) eps = -coup * (mi + mj);
                else eps = coup * qi * charge[gj] * (1 / mi + 1 / mj);
                wi = 2 * mj / (mi + mj);
                wj = 2 * mi / (mi + mj);
              }
              double rel[6];
#pragma unroll
              for (int c = 0; c < 6; c++) rel[c] = ri[c] - rj[c];
              int br = 0, it = 0;
              unsigned fl = 0;
              double res = 0;
              pair_eval<METHOD, MODE>(eps, dt, itr, rel, d, &br, &it, &fl, hist, T, tol, &res);
              if (STATS) {
                ps.add(br, it, fl);
                if (lg.count && (fl & kLogMask)) log_pair(lg, item.z, il, j0 + jj, br, it, fl, res);
              }
            }
#pragma unroll
            for (int c = 0; c < 6; c++) {
              acc[c] += wi * d[c];
              const double sum = wave_sum(wj * d[c]);
              if (lane == jj) col[c] = sum;
            }
          }
          // the history is free between tiles: its first rows take the column sums
#pragma unroll
          for (int c = 0; c < 6; c++) wrow[w][c][lane] = col[c];
        }
        __syncthreads();
      } else {
        __syncthreads();  // the previous column tile's readers are done with cst, cmq and wrow
        // the thread index from the wave index and the lane, through an empty asm statement: the
        // compiler then forms it here instead of keeping it in a register across the pair loop
        int t0 = w * kWave + lane;
        asm volatile("" : "+v"(t0));
        for (int q = t0; q < 6 * kWave; q += kGroup) {
          const int c = q / kWave, l = q % kWave;
          cst[c][l] = l < jn ? st[(size_t)c * lds + S.off + j0 + l] : 0.0;
        }
        if (COUP != kScalar && w == 0) {
          const bool ok = lane < jn;
          const double mj = (ok && mass) ? mass[S.off + j0 + lane] : 1.0;
          cmq[0][lane] = mj;
          cmq[1][lane] = (ok && COUP == kCoulomb) ? charge[S.off + j0 + lane] : 0.0;
          if (COUP == kCoulomb) cmq[2][lane] = 1 / mj;
        }
#pragma unroll
        for (int c = 0; c < 6; c++) wrow[w][c][lane] = 0.0;
        __syncthreads();
        if (tile) {
          const bool diag = (I == J);
          const int k0 = diag ? 1 : 0, k1 = diag ? 33 : kWave;
#pragma unroll 1
          for (int k = k0; k < k1; ++k) {
            const int jl = (lane + k) & (kWave - 1);
            const bool active = row_ok & (jl < jn) & (!diag | (k < 32) | (lane < 32));
            if (active) {
              double d[6];
              double wi = 1, wj = 1;  // multiples of the half deltas credited to i and debited from j
              double eps = S.eps;
              if (COUP != kScalar) {
                const double mj = cmq[0][jl];
                if (COUP == kGravity) eps = -coup * (mi + mj);
                else eps = coup * qi * cmq[1][jl] * (rmi + cmq[2][jl]);
                // the centre-of-mass split 2 m_j/(m_i+m_j), 2 m_i/(m_i+m_j) from one reciprocal
                // (within an ulp); equal masses split exactly in half, as the scalar coupling does
                const double rs = fast_rcp(mi + mj);
                wi = (mi == mj) ? 1.0 : 2 * mj * rs;
                wj = (mi == mj) ? 1.0 : 2 * mi * rs;
              }
              double rel[6];
#pragma unroll
              for (int c = 0; c < 6; c++) rel[c] = ri[c] - cst[c][jl];
              int br = 0, it = 0;
              unsigned fl = 0;
              double res = 0;
              pair_eval<METHOD, MODE>(eps, dt, itr, rel, d, &br, &it, &fl, hist, T, tol, &res);
              if (STATS) {
                ps.add_tile(br, it, fl);
                if (lg.count && (fl & kLogMask)) log_pair(lg, item.z, il, j0 + jl, br, it, fl, res);
              }
#pragma unroll
              for (int c = 0; c < 6; c++) {
                if (COUP == kScalar) {
                  acc[c] += d[c];
                  lds_add(&wrow[w][c][jl], d[c]);
                } else {
                  acc[c] += wi * d[c];
                  lds_add(&wrow[w][c][jl], wj * d[c]);
                }
              }
            }
          }
          if (STATS) ps.fold_tile();
        }
        __syncthreads();  // the column sums are complete
      }
      // the waves' column sums of tile column J, added in wave order, debited from the columns.
      // The thread index goes through an empty asm statement so that the compiler computes these
      // addresses here, not once before the loops (where they would hold registers across the
      // pair loop and cost the Kepler modes a wave per SIMD)
      int q0 = w * kWave + lane;
      asm volatile("" : "+v"(q0));
      for (int q = q0; q < 6 * kWave; q += kGroup) {
        const int c = q / kWave, j = q % kWave;
        double v = wrow[0][c][j];
        for (int ww = 1; ww <= wlast; ++ww) v += wrow[ww][c][j];
        put_partial(B + c * WR + (j0 - kGroup * R) + j, -v);
      }
      if (REF) __syncthreads();  // the history rows are the residual history again
    }
    if (wave_ok) {
      int l = lane;  // as above: the row partials' addresses are computed here
      asm volatile("" : "+v"(l));
      double* rp = B + 6 * WR + (size_t)seg * 6 * kGroup + w * kWave + l;
#pragma unroll
      for (int c = 0; c < 6; c++) put_partial(rp + c * kGroup, acc[c]);
    }
  }
  if (STATS) {
    __shared__ unsigned red[kBlk][16];
    ps.flush(stats, red, w);
  }
}

// Particle i (local index il of block R = il / 256) of a system: the column partials of blocks
// 0..R at i in block order, then the row partials of block R's items in item order (+ v dt for
// positions).  One workgroup per 256-particle chunk {system, first local index}, which is one
// block; one grid row per component.
__global__ __launch_bounds__(256) void k_reduce(const int2* __restrict__ chunks, const SysDesc* __restrict__ sys,
                                                int lds, int ldd, double dt, const double* __restrict__ st,
                                                const double* __restrict__ part, double* __restrict__ delta) {
  const int2 ch = chunks[blockIdx.x];
  const SysDesc S = sys[ch.x];
  const int il = ch.y + threadIdx.x;
  if (il >= S.n) return;
  const int c = blockIdx.y;
  const int R = ch.y / kGroup;
  const Layout L(S.nt, S.seg);
  const double* P = part + S.poff;
  double s = 0;
  // 16 loads in flight per thread, then the adds in order (the sum is the sequential one)
  int r = 0;
  for (; r + 16 <= R + 1; r += 16) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; u++)
      v[u] = __builtin_nontemporal_load(P + L.off(r + u) + c * L.w(r + u) + (il - kGroup * (r + u)));
#pragma unroll
    for (int u = 0; u < 16; u++) s += v[u];
  }
  for (; r <= R; ++r) s += __builtin_nontemporal_load(P + L.off(r) + c * L.w(r) + (il - kGroup * r));
  const double* q = P + L.off(R) + 6 * L.w(R) + (size_t)c * kGroup + threadIdx.x;
  const int nk = L.nseg(R);
  int k = 0;
  for (; k + 16 <= nk; k += 16) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; u++) v[u] = __builtin_nontemporal_load(q + (size_t)(k + u) * 6 * kGroup);
#pragma unroll
    for (int u = 0; u < 16; u++) s += v[u];
  }
  for (; k < nk; ++k) s += __builtin_nontemporal_load(q + (size_t)k * 6 * kGroup);
  const int g = S.off + il;
  if (c < 3) s += st[(size_t)(3 + c) * lds + g] * dt;
  delta[(size_t)c * ldd + g] = s;
}

// Staged reduction of one system (stage of blocks [g0, g0 + gs), partials relative to block g0):
// particle i of block R >= g0 continues its running sum in delta (0 at the first stage) with the
// column partials of the stage's blocks g0 .. min(R, g0 + gs - 1), which continues k_reduce's
// block order; for a block of the stage, then its row partials in item order, after which that
// particle's sum is complete and gets its drift.  One workgroup per block from g0 on, one grid
// row per component.
__global__ __launch_bounds__(256) void k_reduce_stage(int n, int nt, int seg, int g0, int gs, int first, int lds, int ldd,
                                                      double dt, const double* __restrict__ st,
                                                      const double* __restrict__ part, double* __restrict__ delta) {
  const int R = g0 + blockIdx.x;
  const int il = R * kGroup + threadIdx.x;
  if (il >= n) return;
  const int c = blockIdx.y;
  const Layout L(nt, seg);
  const double* P = part - L.off(g0);  // block offsets relative to the stage's first
  double s = first ? 0.0 : delta[(size_t)c * ldd + il];
  const int rend = min(R, g0 + gs - 1);
  int r = g0;
  for (; r + 8 <= rend + 1; r += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++)
      v[u] = __builtin_nontemporal_load(P + L.off(r + u) + c * L.w(r + u) + (il - kGroup * (r + u)));
#pragma unroll
    for (int u = 0; u < 8; u++) s += v[u];
  }
  for (; r <= rend; ++r) s += __builtin_nontemporal_load(P + L.off(r) + c * L.w(r) + (il - kGroup * r));
  if (R < g0 + gs) {
    const double* q = P + L.off(R) + 6 * L.w(R) + (size_t)c * kGroup + threadIdx.x;
    const int nk = L.nseg(R);
    int k = 0;
    for (; k + 8 <= nk; k += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = __builtin_nontemporal_load(q + (size_t)(k + u) * 6 * kGroup);
#pragma unroll
      for (int u = 0; u < 8; u++) s += v[u];
    }
    for (; k < nk; ++k) s += __builtin_nontemporal_load(q + (size_t)k * 6 * kGroup);
    if (c < 3) s += st[(size_t)(3 + c) * lds + il] * dt;
  }
  delta[(size_t)c * ldd + il] = s;
}

// state += delta (the update of bia/bia.f90:203-204), SoA [6][ld], first n columns
__global__ __launch_bounds__(256) void k_update(int n, int lds, int ldd, double* __restrict__ st,
                                                const double* __restrict__ delta) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
#pragma unroll
  for (int c = 0; c < 6; c++) st[(size_t)c * lds + g] += delta[(size_t)c * ldd + g];
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

// ------------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------------
struct StepArgs {
  int method, mode, coupling;
  int nsys;
  const int* sizes;   // host, nsys
  const double* eps;  // host, nsys (SCALAR coupling), may be NULL for the others
  double coup;        // G or k
  double dt;
  int itr;
  const double* d_state;
  int ld;
  const double* d_mass;
  const double* d_charge;
  double* d_delta;
  unsigned long long* d_stats;
};

// autotuner cache key: the kernel variant (method, mode, coupling) and the list of sizes
static std::vector<int> tune_key(const StepArgs& a) {
  std::vector<int> k = {a.method, a.mode, a.coupling};
  k.insert(k.end(), a.sizes, a.sizes + a.nsys);
  return k;
}

// (grid, occupancy variant) the autotuner chose for this kernel variant and sizes, else (0, 1):
// one workgroup per tile and the compiler's own register allocation
static std::pair<int, int> tuned_for(const bia_ctx* ctx, const StepArgs& a) {
  auto it = ctx->tuned_grid.find(tune_key(a));
  return it != ctx->tuned_grid.end() ? it->second : std::pair<int, int>(0, 1);
}

// the captured multi-step graph, if any, is dropped: its kernel arguments point into buffers
// that are about to change
static void drop_graph(bia_ctx* ctx) {
  if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  ctx->graph_exec = nullptr;
  ctx->graph = nullptr;
  ctx->graph_key.clear();
}

// build (or reuse) the plan for this list of sizes; upload system descriptors when eps changed.
// Uploads are synchronous copies, which are not ordered after work in flight on a non-blocking
// stream: the context's stream is drained first, so no step still running reads a descriptor or
// tile list that is being overwritten.  While buffers are rebuilt the plan is marked empty, so a
// failure part-way leaves no plan that a later call could take for valid.
static int ensure_plan(bia_ctx* ctx, const StepArgs& a) {
  Plan& P = ctx->plan;
  std::vector<int> sizes(a.sizes, a.sizes + a.nsys);
  if (P.sizes != sizes || P.limit != ctx->workspace_limit) {
    Plan Q;
    Q.sizes = sizes;
    Q.limit = ctx->workspace_limit;
    Q.nsys = a.nsys;
    std::vector<int4> items;  // the work items (R, k, system)
    std::vector<int2> chunks;
    long long off = 0, poff = 0;
    for (int s = 0; s < a.nsys; s++) {
      const int n = sizes[s];
      if (n < 1) return BIA_ERR_ARG;
      const int nt = (n + kWave - 1) / kWave;
      const Layout L(nt, choose_seg(nt));
      SysDesc d{(int)off, n, nt, L.seg, poff, 0.0};
      Q.hsys.push_back(d);
      for (int R = 0; R < L.nb; R++)
        for (int k = 0; k < L.nseg(R); k++) items.push_back(make_int4(R, k, s, 0));
      for (int c0 = 0; c0 < n; c0 += kGroup) chunks.push_back(make_int2(s, c0));  // one chunk per block
      off += n;
      poff += L.off(L.nb);
    }
    if (off > 0x7fffffffLL || items.size() > 0x7fffffffULL) return BIA_ERR_ARG;
    Q.ntotal = (int)off;
    Q.nitems = (int)items.size();
    Q.nchunks = (int)chunks.size();
    Q.part_doubles = (size_t)poff;
    Q.stage_t0 = {0, Q.nitems};
    if (a.nsys == 1 && ctx->workspace_limit > 0 && Q.part_doubles * sizeof(double) > ctx->workspace_limit) {
      // stages of G blocks: the first stage's blocks hold the most partials (a block's partials
      // shrink with R), L.off(G) doubles, which the limit bounds (at least one block)
      const Layout L(Q.hsys[0].nt, Q.hsys[0].seg);
      int G = 1;
      while (G < L.nb && (size_t)L.off(G + 1) * sizeof(double) <= ctx->workspace_limit) ++G;
      if (G < L.nb) {
        Q.stage_blocks = G;
        Q.nstages = (L.nb + G - 1) / G;
        Q.stage_t0.clear();
        for (int g = 0; g <= Q.nstages; g++) Q.stage_t0.push_back((int)L.items_before(std::min(L.nb, g * G)));
        Q.part_doubles = (size_t)L.off(G);
      }
    }
    P = Plan();  // invalid until every upload below has succeeded
    drop_graph(ctx);
    ctx->plan_gen++;
    BIA_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    BIA_TRY(ctx, ctx->plan_items.ensure(items.size() * sizeof(int4)));
    BIA_TRY(ctx, ctx->plan_chunks.ensure(chunks.size() * sizeof(int2)));
    BIA_TRY(ctx, ctx->plan_sys.ensure(Q.hsys.size() * sizeof(SysDesc)));
    BIA_HIP_TRY(ctx, hipMemcpy(ctx->plan_items.p, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice));
    BIA_HIP_TRY(ctx, hipMemcpy(ctx->plan_chunks.p, chunks.data(), chunks.size() * sizeof(int2), hipMemcpyHostToDevice));
    for (auto& d : Q.hsys) d.eps = __builtin_nan("");  // force the descriptor upload below
    P = Q;
  }
  std::vector<SysDesc> upd = P.hsys;
  bool dirty = false;
  for (int s = 0; s < a.nsys; s++) {
    const double e = a.eps ? a.eps[s] : 0.0;
    if (!(upd[s].eps == e)) {
      upd[s].eps = e;
      dirty = true;
    }
  }
  if (dirty) {
    BIA_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (auto& d : P.hsys) d.eps = __builtin_nan("");  // unknown on the device until the copy succeeds
    BIA_HIP_TRY(ctx, hipMemcpy(ctx->plan_sys.p, upd.data(), upd.size() * sizeof(SysDesc), hipMemcpyHostToDevice));
    P.hsys = upd;
  }
  if (P.part_doubles * sizeof(double) > ctx->part.bytes) {
    drop_graph(ctx);
    ctx->plan_gen++;
  }
  BIA_TRY(ctx, ctx->part.ensure(P.part_doubles * sizeof(double)));
  if ((a.d_stats || ctx->log_count) && !ctx->stat_slots.p) {
    const int n = kStatSlots * BIA_STATS_LEN;
    BIA_TRY(ctx, ctx->stat_slots.ensure((size_t)n * sizeof(unsigned long long)));
    hipLaunchKernelGGL(k_zero_u64, dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                       ctx->stat_slots.as<unsigned long long>(), n);
    BIA_HIP_TRY(ctx, hipGetLastError());
  }
  return BIA_OK;
}

// the learned-start table a mode reads (none for the plain and guess modes)
static LearnTable mode_table(const bia_ctx* ctx, int mode) {
  if (mode != BIA_MODE_LEARN && mode != BIA_MODE_PLAIN_LEARN) return LearnTable{nullptr, 0};
  const NetTable t = net_table(ctx, mode == BIA_MODE_LEARN ? 0 : 1);
  return LearnTable{t.t, t.n};
}

static int validate(bia_ctx* ctx, const StepArgs& a) {
  if (a.nsys < 1 || !a.sizes || !a.d_state || !a.d_delta) return BIA_ERR_ARG;
  if (a.method != BIA_METHOD_ALG && a.method != BIA_METHOD_RKF) return BIA_ERR_ARG;
  if (a.mode != BIA_MODE_PLAIN && a.mode != BIA_MODE_LEARN && a.mode != BIA_MODE_GUESS &&
      a.mode != BIA_MODE_PLAIN_LEARN)
    return BIA_ERR_ARG;
  if (a.mode != BIA_MODE_PLAIN && a.method != BIA_METHOD_ALG) return BIA_ERR_ARG;
  if (a.coupling < BIA_COUPLING_SCALAR || a.coupling > BIA_COUPLING_COULOMB) return BIA_ERR_ARG;
  if (a.coupling == BIA_COUPLING_SCALAR && !a.eps) return BIA_ERR_ARG;
  if (a.coupling == BIA_COUPLING_GRAVITY && !a.d_mass) return BIA_ERR_ARG;
  if (a.coupling == BIA_COUPLING_COULOMB && !a.d_charge) return BIA_ERR_ARG;
  if (a.itr < 0 || a.itr > kMaxItr) return BIA_ERR_ARG;
  long long tot = 0;
  for (int s = 0; s < a.nsys; s++) tot += a.sizes[s];
  if (a.ld < tot) return BIA_ERR_ARG;
  if (a.mode == BIA_MODE_LEARN && !net_table(ctx, 0).ready) {
    snprintf(ctx->last_error, sizeof(ctx->last_error), "learned mode needs a trained network (bia_mlp_build_table)");
    return BIA_ERR_STATE;
  }
  if (a.mode == BIA_MODE_PLAIN_LEARN && !net_table(ctx, 1).ready) {
    snprintf(ctx->last_error, sizeof(ctx->last_error),
             "reference solvers with learned starts need network 1 trained and tabulated (bia_mlp_select, "
             "bia_mlp_build_table)");
    return BIA_ERR_STATE;
  }
  return BIA_OK;
}

// the device plan a launch reads: the context's buffers, or the no-allocation scratch array
struct PlanView {
  const int4* items;
  const SysDesc* sys;
  const int2* chunks;
  double* part;
  int nitems, nchunks;
  bool context_plan;  // the context's own plan: statistics, failure records and autotuned grids apply
};

// enqueue pair kernel + reduction on ctx->stream (no synchronisation, no allocation once planned);
// view NULL: the context's plan
static int launch_step(bia_ctx* ctx, const StepArgs& a, int grid_override, int occ_override = 0,
                       const PlanView* view = nullptr) {
  const PlanView V = view ? *view
                          : PlanView{ctx->plan_items.as<int4>(), ctx->plan_sys.as<SysDesc>(), ctx->plan_chunks.as<int2>(),
                                     ctx->part.as<double>(), ctx->plan.nitems, ctx->plan.nchunks, true};
  const auto tuned = V.context_plan ? tuned_for(ctx, a) : std::pair<int, int>(0, 1);
  const int tuned_grid = tuned.first, tuned_occ = tuned.second;
  int occ = occ_override > 0 ? occ_override : tuned_occ;
  const LearnTable T = mode_table(ctx, a.mode);
  int grid = grid_override > 0 ? grid_override : (tuned_grid > 0 ? tuned_grid : V.nitems);
  // a dispatch holds fewer than 2^32 work-items; the pair kernel is grid-stride, so a grid capped
  // there still covers every item
  grid = std::min(grid, std::min(V.nitems, kMaxGrid));
  const int4* items = V.items;
  const SysDesc* sys = V.sys;
  double* part = V.part;
  // a failure log needs the statistics variant of the pair kernel, which writes the records; its
  // counters go to the caller's array or, without one, to the context's own
  const PairLog lg = V.context_plan ? PairLog{static_cast<PairRecord*>(ctx->log_rec), ctx->log_count, ctx->log_cap}
                                    : PairLog{nullptr, nullptr, 0};
  if (!V.context_plan && a.d_stats) return BIA_ERR_ARG;
  unsigned long long* stats_out = a.d_stats ? a.d_stats : lg.count ? ctx->log_stats.as<unsigned long long>() : nullptr;
  const bool st = stats_out != nullptr;
  // the statistics variant holds 20 more registers per lane: at the autotuner's 5 waves per SIMD
  // it spills (52-108 bytes per lane) and runs a third slower, at 4 it does not
  if (st && occ == 5 && occ_override <= 0) occ = 4;
  unsigned long long* slots = ctx->stat_slots.as<unsigned long long>();
  if (st && !slots) return BIA_ERR_STATE;  // ensure_plan allocates the slots for a step with statistics
  const RkfTol tol{ctx->rkf_rel, ctx->rkf_abs};
  const Plan& Pl = ctx->plan;
  const bool staged = V.context_plan && Pl.nstages > 1;
  const int nstages = staged ? Pl.nstages : 1;
  for (int g = 0; g < nstages; g++) {
  Stage stg{0, V.nitems, 0, 0};
  if (staged) {
    stg = Stage{Pl.stage_t0[g], Pl.stage_t0[g + 1], g * Pl.stage_blocks, 1};
    grid = std::min(grid_override > 0 ? grid_override : (tuned_grid > 0 ? tuned_grid : stg.t1 - stg.t0),
                    std::min(stg.t1 - stg.t0, kMaxGrid));
  }
  BIA_TRY(ctx, prof_begin(ctx));
#define L_OCC(M, MO, CP, S, O)                                                                                     \
  hipLaunchKernelGGL((k_tiles<M, MO, CP, S, O>), dim3(grid), dim3(kGroup), 0, ctx->stream, stg, a.ld, items, sys, \
                     a.d_state, a.d_mass, a.d_charge, a.coup, a.dt, a.itr, part, slots, T, tol, lg)
#define L_(M, MO, CP, S)                                    \
  do {                                                      \
    if (occ == 4) L_OCC(M, MO, CP, S, 4);                   \
    else if (M == 1 && !reference_solvers(MO) && occ == 5) L_OCC(M, MO, CP, S, 5); \
    else L_OCC(M, MO, CP, S, 1);                            \
  } while (0)
#define L_COUP(M, MO, S)                                                  \
  do {                                                                    \
    if (a.coupling == BIA_COUPLING_SCALAR) L_(M, MO, kScalar, S);         \
    else if (a.coupling == BIA_COUPLING_GRAVITY) L_(M, MO, kGravity, S);  \
    else L_(M, MO, kCoulomb, S);                                          \
  } while (0)
  if (a.method == BIA_METHOD_RKF) {
    if (st) L_COUP(0, kPlain, true);
    else L_COUP(0, kPlain, false);
  } else if (a.mode == BIA_MODE_PLAIN) {
    if (st) L_COUP(1, kPlain, true);
    else L_COUP(1, kPlain, false);
  } else if (a.mode == BIA_MODE_LEARN) {
    if (st) L_COUP(1, kLearn, true);
    else L_COUP(1, kLearn, false);
  } else if (a.mode == BIA_MODE_PLAIN_LEARN) {
    if (st) L_COUP(1, kPlainLearn, true);
    else L_COUP(1, kPlainLearn, false);
  } else {
    if (st) L_COUP(1, kGuess, true);
    else L_COUP(1, kGuess, false);
  }
#undef L_COUP
#undef L_
#undef L_OCC
  BIA_HIP_TRY(ctx, hipGetLastError());
  BIA_TRY(ctx, prof_end(ctx));
  if (staged) {  // this stage's partials into the running sums (the final deltas of its own rows)
    const SysDesc& S0 = Pl.hsys[0];
    const int nb = Layout(S0.nt, S0.seg).nb;
    const int gs = std::min(Pl.stage_blocks, nb - stg.g0);
    BIA_TRY(ctx, prof_begin(ctx, 1));
    hipLaunchKernelGGL(k_reduce_stage, dim3(nb - stg.g0, 6), dim3(kGroup), 0, ctx->stream, S0.n, S0.nt, S0.seg, stg.g0, gs,
                       g == 0 ? 1 : 0, a.ld, a.ld, a.dt, a.d_state, part, a.d_delta);
    BIA_HIP_TRY(ctx, hipGetLastError());
    BIA_TRY(ctx, prof_end(ctx, 1));
  }
  }  // stages
  if (st) {
    hipLaunchKernelGGL(k_stats_fold, dim3(BIA_STATS_LEN), dim3(256), 0, ctx->stream, slots, stats_out);
    BIA_HIP_TRY(ctx, hipGetLastError());
  }
  if (!staged) {
    BIA_TRY(ctx, prof_begin(ctx, 1));
    hipLaunchKernelGGL(k_reduce, dim3(V.nchunks, 6), dim3(kGroup), 0, ctx->stream, V.chunks, sys, a.ld, a.ld, a.dt,
                       a.d_state, part, a.d_delta);
    BIA_HIP_TRY(ctx, hipGetLastError());
    BIA_TRY(ctx, prof_end(ctx, 1));
  }
  return BIA_OK;
}

static int step_locked(bia_ctx* ctx, const StepArgs& a) {
  BIA_TRY(ctx, validate(ctx, a));
  BIA_TRY(ctx, ensure_plan(ctx, a));
  return launch_step(ctx, a, 0);
}

// Time candidate grid sizes of the persistent pair kernel on the caller's input (one step each,
// outputs written to d_delta) and keep the fastest for this list of sizes.
static int autotune_locked(bia_ctx* ctx, const StepArgs& a, float* best_ms, int* best_grid) {
  BIA_TRY(ctx, validate(ctx, a));
  BIA_TRY(ctx, ensure_plan(ctx, a));
  int ncu = 0;
  BIA_HIP_TRY(ctx, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device));
  const int nitems = ctx->plan.nitems;
  std::vector<int> cands = {nitems};
  for (int m : {2, 4, 8, 16})
    if (m * ncu < nitems) cands.push_back(m * ncu);
  hipEvent_t e0, e1;
  BIA_HIP_TRY(ctx, hipEventCreate(&e0));
  BIA_HIP_TRY(ctx, hipEventCreate(&e1));
  StepArgs b = a;
  b.d_stats = nullptr;
  const int saved_prof = ctx->profiling;
  ctx->profiling = 0;
  float best = 1e30f;
  int bestg = nitems, besto = 1;
  int status = BIA_OK;
  std::vector<int> occs = {1, 4};
  if (a.method == BIA_METHOD_ALG && (a.mode == BIA_MODE_LEARN || a.mode == BIA_MODE_GUESS)) occs.push_back(5);
  for (int o : occs)
    for (int g : cands) {
      if ((status = launch_step(ctx, b, g, o)) != BIA_OK) break;  // warm-up
      (void)hipEventRecord(e0, ctx->stream);
      if ((status = launch_step(ctx, b, g, o)) != BIA_OK) break;
      (void)hipEventRecord(e1, ctx->stream);
      (void)hipEventSynchronize(e1);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) {
        best = ms;
        bestg = g;
        besto = o;
      }
    }
  ctx->profiling = saved_prof;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (status != BIA_OK) return status;
  ctx->tuned_grid[tune_key(a)] = {bestg, besto};
  if (best_ms) *best_ms = best;
  if (best_grid) *best_grid = bestg * 10 + besto;
  return BIA_OK;
}

// nsteps of (step, state += delta) on the stream; with use_graph the step is captured once into a
// hipGraph (captured again when the arguments change) and replayed
static int run_steps_graph(bia_ctx* ctx, const StepArgs& a, double* d_state_rw, int nsteps);

static int run_steps_locked(bia_ctx* ctx, const StepArgs& a, double* d_state_rw, int nsteps, int use_graph) {
  BIA_TRY(ctx, validate(ctx, a));
  BIA_TRY(ctx, ensure_plan(ctx, a));
  const int ntot = ctx->plan.ntotal;
  if (use_graph) {
    // capture is not possible on the null stream: use a private blocking stream, which stays
    // ordered with null-stream work on both sides
    hipStream_t user = ctx->stream;
    if (!user) {
      if (!ctx->cap_stream) BIA_HIP_TRY(ctx, hipStreamCreate(&ctx->cap_stream));
      ctx->stream = ctx->cap_stream;
    }
    const int s = run_steps_graph(ctx, a, d_state_rw, nsteps);
    ctx->stream = user;
    return s;
  }
  {
    for (int k = 0; k < nsteps; k++) {
      BIA_TRY(ctx, launch_step(ctx, a, 0));
      hipLaunchKernelGGL(k_update, dim3((ntot + 255) / 256), dim3(256), 0, ctx->stream, ntot, a.ld, a.ld, d_state_rw,
                         a.d_delta);
      BIA_HIP_TRY(ctx, hipGetLastError());
    }
    return BIA_OK;
  }
}

static int run_steps_graph(bia_ctx* ctx, const StepArgs& a, double* d_state_rw, int nsteps) {
  const int ntot = ctx->plan.ntotal;
  const auto tuned = tuned_for(ctx, a);
  std::vector<long long> key = {a.method, a.mode, a.coupling, a.itr, (long long)(size_t)d_state_rw,
                                (long long)(size_t)a.d_delta, (long long)(size_t)a.d_mass,
                                (long long)(size_t)a.d_charge, a.ld, (long long)(size_t)ctx->stream, tuned.first, tuned.second,
                                (long long)(size_t)ctx->part.p,
                                (long long)(size_t)mode_table(ctx, a.mode).t, mode_table(ctx, a.mode).n,
                                (long long)(size_t)a.d_stats,
                                (long long)(size_t)ctx->stat_slots.p, ctx->plan_gen,
                                (long long)(size_t)ctx->plan_items.p, (long long)(size_t)ctx->plan_sys.p,
                                (long long)(size_t)ctx->plan_chunks.p, (long long)(size_t)ctx->log_rec,
                                (long long)(size_t)ctx->log_count, ctx->log_cap};
  for (const double v : {a.dt, a.coup, ctx->rkf_rel, ctx->rkf_abs}) {
    long long b;
    memcpy(&b, &v, 8);
    key.push_back(b);
  }
  for (int s : ctx->plan.sizes) key.push_back(s);
  if (key != ctx->graph_key || !ctx->graph_exec) {
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    ctx->graph_exec = nullptr;
    ctx->graph = nullptr;
    const int saved_prof = ctx->profiling;
    ctx->profiling = 0;  // no event records inside the graph
    BIA_HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int s = launch_step(ctx, a, 0);
    if (s == BIA_OK)
      hipLaunchKernelGGL(k_update, dim3((ntot + 255) / 256), dim3(256), 0, ctx->stream, ntot, a.ld, a.ld, d_state_rw,
                         a.d_delta);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    ctx->profiling = saved_prof;
    if (s != BIA_OK) {
      if (g) (void)hipGraphDestroy(g);
      return s;
    }
    BIA_HIP_TRY(ctx, e);
    ctx->graph = g;
    BIA_HIP_TRY(ctx, hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0));
    ctx->graph_key = key;
  }
  for (int k = 0; k < nsteps; k++) BIA_HIP_TRY(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream));
  return BIA_OK;
}

// ------------------------------------------------------------------------------------------------
// no-allocation step (bia_step_noalloc_dev).  The plan and the partials of one system of at most
// kNoAllocMaxN particles live in the file-scope device array `scratch` (32 MB, held by every GPU
// the library runs on until the process ends), so the call allocates nothing.  In host code
// `scratch` names a host-side stand-in, not device memory: its device address on each GPU comes
// from hipGetSymbolAddress, once per device at bia_create, and only that address is used.
// ------------------------------------------------------------------------------------------------
constexpr int kNoAllocMaxN = 8192;
constexpr int kNoAllocMaxNt = kNoAllocMaxN / kWave;
// sized for single-tile segments, which hold the most items and partials of any segment length
constexpr Layout kNoAllocLayout(kNoAllocMaxNt, 1);
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kOffItems = 0;
constexpr size_t kOffChunks =
    align256(kOffItems + (size_t)kNoAllocLayout.items_before(kNoAllocLayout.nb) * sizeof(int4));
constexpr size_t kOffSys = align256(kOffChunks + (size_t)(kNoAllocMaxN + kGroup - 1) / kGroup * sizeof(int2));
constexpr size_t kOffPart = align256(kOffSys + sizeof(SysDesc));
constexpr size_t kScratchBytes = kOffPart + (size_t)kNoAllocLayout.off(kNoAllocLayout.nb) * sizeof(double);

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each GPU

// with `device` current: look up and keep this GPU's address of the scratch array (once)
int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return BIA_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return BIA_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    (void)hipGetLastError();
    return BIA_ERR_HIP;
  }
  g_scratch[device] = ptr;
  return BIA_OK;
}

static void* scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch[device];
}

// the plan of one n-particle system, as ensure_plan lists it: items (R, k), k < nseg(R), block by
// block; reduction chunks of one block (256 particles); the system descriptor.  Rewritten on every
// call (a fill kernel: a copy or memset node would not be safe to capture in a graph).
__global__ __launch_bounds__(256) void k_noalloc_plan(int n, int nt, int seg, double eps, int4* __restrict__ items,
                                                      int2* __restrict__ chunks, SysDesc* __restrict__ sys) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const Layout L(nt, seg);
  if (t < L.items_before(L.nb)) {
    int R = 0, first = 0;  // block R holds items [first, first + nseg(R))
    while (first + L.nseg(R) <= t) {
      first += L.nseg(R);
      ++R;
    }
    items[t] = make_int4(R, t - first, 0, 0);
  }
  if (t < (n + kGroup - 1) / kGroup) chunks[t] = make_int2(0, kGroup * t);
  if (t == 0) *sys = SysDesc{0, n, nt, seg, 0LL, eps};
}

static_assert(sizeof(PairRecord) == sizeof(bia_pair_record) && offsetof(PairRecord, error) ==
                  offsetof(bia_pair_record, error), "the device record is the public one");

}  // namespace bia

using namespace bia;

#define BIA_LOCKED(ctx)                       \
  if (!(ctx)) return BIA_ERR_ARG;             \
  std::lock_guard<std::mutex> lk_((ctx)->mu); \
  (ctx)->last_error[0] = 0;                   \
  BIA_HIP_TRY((ctx), hipSetDevice((ctx)->device))

extern "C" {

int bia_step_dev(bia_ctx* ctx, int method, int mode, int n, double eps, double dt, int itr, const double* d_state,
                 int ld, double* d_delta, unsigned long long* d_stats) {
  BIA_LOCKED(ctx);
  StepArgs a{method, mode, BIA_COUPLING_SCALAR, 1, &n, &eps, 0.0, dt, itr, d_state, ld, nullptr, nullptr, d_delta,
             d_stats};
  return step_locked(ctx, a);
}

int bia_step_batched_dev(bia_ctx* ctx, int method, int mode, int coupling, int nsys, const int* sizes,
                         const double* eps, double coup, double dt, int itr, const double* d_state, int ld,
                         const double* d_mass, const double* d_charge, double* d_delta, unsigned long long* d_stats) {
  BIA_LOCKED(ctx);
  StepArgs a{method, mode, coupling, nsys, sizes, eps, coup, dt, itr, d_state, ld, d_mass, d_charge, d_delta, d_stats};
  return step_locked(ctx, a);
}

int bia_autotune_dev(bia_ctx* ctx, int method, int mode, int coupling, int nsys, const int* sizes, const double* eps,
                     double coup, double dt, int itr, const double* d_state, int ld, const double* d_mass,
                     const double* d_charge, double* d_delta, float* best_ms, int* best_grid) {
  BIA_LOCKED(ctx);
  StepArgs a{method, mode, coupling, nsys, sizes, eps, coup, dt, itr, d_state, ld, d_mass, d_charge, d_delta, nullptr};
  return autotune_locked(ctx, a, best_ms, best_grid);
}

int bia_run_steps_dev(bia_ctx* ctx, int method, int mode, int coupling, int nsys, const int* sizes, const double* eps,
                      double coup, double dt, int itr, double* d_state, int ld, const double* d_mass,
                      const double* d_charge, double* d_delta, int nsteps, int use_graph, unsigned long long* d_stats) {
  BIA_LOCKED(ctx);
  if (nsteps < 0) return BIA_ERR_ARG;
  StepArgs a{method, mode, coupling, nsys, sizes, eps, coup, dt, itr, d_state, ld, d_mass, d_charge, d_delta, d_stats};
  return run_steps_locked(ctx, a, d_state, nsteps, use_graph);
}

int bia_step_noalloc_dev(bia_ctx* ctx, int method, int mode, int n, double eps, double dt, int itr,
                         const double* d_state, int ld, double* d_delta) {
  BIA_LOCKED(ctx);
  if (n < 1 || n > kNoAllocMaxN) {
    snprintf(ctx->last_error, sizeof(ctx->last_error), "n = %d: the no-allocation step takes 1 <= n <= %d", n,
             kNoAllocMaxN);
    return BIA_ERR_ARG;
  }
  StepArgs a{method, mode, BIA_COUPLING_SCALAR, 1, &n, &eps, 0.0, dt, itr, d_state, ld, nullptr, nullptr, d_delta,
             nullptr};
  BIA_TRY(ctx, validate(ctx, a));
  unsigned char* base = static_cast<unsigned char*>(scratch_on(ctx->device));
  if (!base) {
    snprintf(ctx->last_error, sizeof(ctx->last_error), "no scratch address for device %d", ctx->device);
    return BIA_ERR_STATE;
  }
  const int nt = (n + kWave - 1) / kWave;
  const Layout L(nt, choose_seg(nt));
  const int nitems = (int)L.items_before(L.nb), nchunks = (n + kGroup - 1) / kGroup;
  int4* items = reinterpret_cast<int4*>(base + kOffItems);
  int2* chunks = reinterpret_cast<int2*>(base + kOffChunks);
  SysDesc* sys = reinterpret_cast<SysDesc*>(base + kOffSys);
  double* part = reinterpret_cast<double*>(base + kOffPart);
  hipLaunchKernelGGL(k_noalloc_plan, dim3((nitems + 255) / 256), dim3(256), 0, ctx->stream, n, nt, L.seg, eps, items,
                     chunks, sys);
  BIA_HIP_TRY(ctx, hipGetLastError());
  const PlanView V{items, sys, chunks, part, nitems, nchunks, false};
  return launch_step(ctx, a, 0, 0, &V);
}

int bia_set_workspace_limit(bia_ctx* ctx, unsigned long long bytes) {
  BIA_LOCKED(ctx);
  ctx->workspace_limit = (size_t)bytes;
  return BIA_OK;
}

int bia_plan_info(bia_ctx* ctx, int* nstages, unsigned long long* workspace_bytes) {
  BIA_LOCKED(ctx);
  if (ctx->plan.sizes.empty()) return BIA_ERR_STATE;
  if (nstages) *nstages = ctx->plan.nstages;
  if (workspace_bytes) *workspace_bytes = (unsigned long long)(ctx->plan.part_doubles * sizeof(double));
  return BIA_OK;
}

int bia_set_rkf_tolerance(bia_ctx* ctx, double relerr, double abserr) {
  BIA_LOCKED(ctx);
  // the error test divides by relerr |y| + abserr: both non-negative and finite, not both zero
  if (!(relerr >= 0 && abserr >= 0 && relerr + abserr > 0) || !std::isfinite(relerr + abserr)) return BIA_ERR_ARG;
  ctx->rkf_rel = relerr;
  ctx->rkf_abs = abserr;
  return BIA_OK;
}

int bia_set_failure_log(bia_ctx* ctx, bia_pair_record* d_records, int capacity, unsigned long long* d_count) {
  BIA_LOCKED(ctx);
  if (!d_count) {  // logging off
    ctx->log_rec = nullptr;
    ctx->log_count = nullptr;
    ctx->log_cap = 0;
    return BIA_OK;
  }
  if (capacity < 0 || (capacity > 0 && !d_records)) return BIA_ERR_ARG;
  BIA_TRY(ctx, ctx->log_stats.ensure(BIA_STATS_LEN * sizeof(unsigned long long)));
  ctx->log_rec = d_records;
  ctx->log_count = d_count;
  ctx->log_cap = capacity;
  return BIA_OK;
}

int bia_step_host(bia_ctx* ctx, int method, int mode, int n, double eps, double dt, int itr, const double* X,
                  const double* Y, const double* Z, const double* U, const double* V, const double* W, double* dX,
                  double* dY, double* dZ, double* dU, double* dV, double* dW, unsigned long long* stats) {
  if (!ctx || n < 1) return BIA_ERR_ARG;
  const double* in[6] = {X, Y, Z, U, V, W};
  double* out[6] = {dX, dY, dZ, dU, dV, dW};
  for (int c = 0; c < 6; c++)
    if (!in[c] || !out[c]) return BIA_ERR_ARG;
  BIA_LOCKED(ctx);
  StreamDrain drain{ctx};  // the copies below read and write the caller's host arrays
  const int ld = n;
  // device staging: [state 6n][delta 6n][stats]
  BIA_TRY(ctx, ctx->delta.ensure((size_t)12 * ld * sizeof(double) + BIA_STATS_LEN * 8));
  double* dst = ctx->delta.as<double>();
  double* ddl = dst + (size_t)6 * ld;
  unsigned long long* dsa = reinterpret_cast<unsigned long long*>(ddl + (size_t)6 * ld);
  for (int c = 0; c < 6; c++)
    BIA_HIP_TRY(ctx, hipMemcpyAsync(dst + (size_t)c * ld, in[c], (size_t)n * sizeof(double), hipMemcpyHostToDevice,
                                    ctx->stream));
  BIA_HIP_TRY(ctx, hipMemsetAsync(dsa, 0, BIA_STATS_LEN * 8, ctx->stream));
  StepArgs a{method, mode, BIA_COUPLING_SCALAR, 1, &n, &eps, 0.0, dt, itr, dst, ld, nullptr, nullptr, ddl, dsa};
  BIA_TRY(ctx, step_locked(ctx, a));
  for (int c = 0; c < 6; c++)
    BIA_HIP_TRY(ctx, hipMemcpyAsync(out[c], ddl + (size_t)c * ld, (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                                    ctx->stream));
  if (stats) BIA_HIP_TRY(ctx, hipMemcpyAsync(stats, dsa, BIA_STATS_LEN * 8, hipMemcpyDeviceToHost, ctx->stream));
  BIA_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return BIA_OK;
}

void find_dxyuv_kernel_(int METHOD, int N, int I, double EPS, double DT, int ITR, double* dX, double* dY, double* dZ,
                        double* dU, double* dV, double* dW, double* RXYZUVW, double* dXYZUVW) {
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
  // only rows j > i are read and produced (the reference copies both whole arrays each way and
  // leaves the other rows of dXYZUVW unspecified): they pass through pinned host buffers that the
  // kernel reads and writes directly, one launch and one synchronisation per call
  const int cnt = N - (i + 1);
  const size_t off = (size_t)6 * (i + 1), rb = (size_t)6 * cnt * sizeof(double);
  if (e == hipSuccess && cnt > 0) {
    memcpy(ctx->leg_in.host<double>() + off, RXYZUVW + off, rb);
    dim3 g((cnt + kWave - 1) / kWave), b(kWave);
    if (METHOD == 1)
      hipLaunchKernelGGL(k_legacy_row<1>, g, b, 0, ctx->stream, N, i, EPS, DT, ITR, ctx->leg_in.device<double>(),
                         ctx->leg_out.device<double>());
    else
      hipLaunchKernelGGL(k_legacy_row<0>, g, b, 0, ctx->stream, N, i, EPS, DT, ITR, ctx->leg_in.device<double>(),
                         ctx->leg_out.device<double>());
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) memcpy(dXYZUVW + off, ctx->leg_out.host<double>() + off, rb);
  }
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
