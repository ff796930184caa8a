// bia_runtime.hpp -- library context, error handling and cached device buffers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

#include <map>
#include <mutex>
#include <vector>

#include "../../include/bia_hip.h"

namespace bia {

constexpr int kProfSlots = 256;  // pair-kernel launches whose timing events are kept

// one system of a (possibly batched) step: particles [off, off + n) of the SoA state, nt = ceil(n/64)
// tiles, partial sums at part + poff (blocks of four row tiles, each its column partials and then
// its row partials; bia_step.hip, blk_off)
struct SysDesc {
  int off, n, nt, seg;  // seg: column tiles per pair-kernel work item (bia_step.hip, choose_seg)
  long long poff;
  double eps;
};

// cached launch plan for a list of system sizes: the pair kernel's work items (block of four row
// tiles, segment of four column tiles, system).  A single system whose partials would exceed the
// context's workspace limit runs in nstages stages of stage_blocks blocks each (stage g takes items
// [stage_t0[g], stage_t0[g + 1]) of the list); otherwise nstages = 1.
struct Plan {
  std::vector<int> sizes;
  size_t limit = 0;  // the workspace limit the plan was made for
  int nsys = 0, nitems = 0, nchunks = 0, ntotal = 0;
  size_t part_doubles = 0;
  int nstages = 1, stage_blocks = 0;
  std::vector<int> stage_t0;
  std::vector<SysDesc> hsys;  // host copy (eps updated per call)
};

// grow-only device buffer; reused across calls so steady-state steps allocate nothing
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t need) {
    if (need <= bytes) return BIA_OK;
    if (p) {
      if (hipFree(p) != hipSuccess) return BIA_ERR_HIP;
      p = nullptr;
      bytes = 0;
    }
    if (need == 0) return BIA_OK;
    if (hipMalloc(&p, need) != hipSuccess) {
      (void)hipGetLastError();
      return BIA_ERR_OOM;
    }
    bytes = need;
    return BIA_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Pinned host memory (hipHostMalloc) that kernels read and write directly through its device
// address: the per-particle reference entry point moves its small rows this way instead of with
// two copies per call.
struct HostBuf {
  void* p = nullptr;
  void* dev = nullptr;
  size_t bytes = 0;
  int ensure(size_t need) {
    if (need <= bytes) return BIA_OK;
    release();
    if (need == 0) return BIA_OK;
    if (hipHostMalloc(&p, need, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return BIA_ERR_OOM;
    }
    if (hipHostGetDevicePointer(&dev, p, 0) != hipSuccess) {
      (void)hipGetLastError();
      release();
      return BIA_ERR_HIP;
    }
    bytes = need;
    return BIA_OK;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = dev = nullptr;
    bytes = 0;
  }
  template <class T>
  T* host() const {
    return static_cast<T*>(p);
  }
  template <class T>
  T* device() const {
    return static_cast<T*>(dev);
  }
};

}  // namespace bia

struct bia_ctx {
  int device = 0;
  hipStream_t stream = nullptr;  // default: the null stream (torch's default stream too)
  hipStream_t cap_stream = nullptr;  // private blocking stream for graph capture when stream is null
  std::mutex mu;                 // one call at a time per context
  // step workspace; stat_slots: per-workgroup statistics slots of the pair kernel, zero between
  // launches (the fold kernel clears what it reads)
  bia::DevBuf part, state, delta, stat_slots;
  // batched plans: device tile list (int4 I, J, sys, 0), system descriptors, reduce chunks
  bia::Plan plan;
  bia::DevBuf plan_items, plan_sys, plan_chunks;
  std::map<std::vector<int>, std::pair<int, int>> tuned_grid;  // autotuner cache: (method, mode, coupling, sizes) -> (grid, occ)
  // multi-step driver graph; plan_gen counts the rebuilds of the plan's device buffers (the
  // graph's kernel arguments point into them, so a graph is only replayed on the generation it
  // was captured on)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<long long> graph_key;
  long long plan_gen = 0;
  // legacy per-row buffers
  bia::HostBuf leg_in, leg_out;  // rows of the per-particle reference entry point
  // invariance workspace: staging, partial sums, and the plan of the last list of systems
  // (device tile list int4 {sys, I, J, 0} and per-system descriptors, keyed by sizes and eps)
  bia::DevBuf inv_state, inv_part, inv_out, inv_tiles, inv_sys;
  std::vector<long long> inv_key;
  int inv_ntiles = 0, inv_npe = 0, inv_nlin = 0, inv_maxnpe = 0, inv_maxnlin = 0;
  // DIM workspace; dim_h: pinned staging of the step sizes of the captured steps
  bia::DevBuf dim_work, dim_ctl;
  bia::HostBuf dim_h;
  // learned initialisers: fp32 master weights + Adam moments, training workspace, tabulated
  // output.  Two networks: 0 starts the Kepler drift (BIA_MODE_LEARN), 1 the reference solvers
  // (BIA_MODE_PLAIN_LEARN).  The mlp_* fields hold the network bia_mlp_select chose (mlp_net),
  // the alt_* fields the other one.
  bia::DevBuf mlp_w, mlp_ws, mlp_table;
  int mlp_init = 0;   // weights present
  int mlp_ready = 0;  // table built from the current weights
  int table_n = 0;
  long long mlp_step = 0;
  int mlp_net = 0;
  bia::DevBuf alt_w, alt_table;
  int alt_init = 0, alt_ready = 0, alt_table_n = 0;
  long long alt_step = 0;
  // per-pair RKF45 tolerances (the reference's RELERR 0, ABSERR 1e-16 by default) and the
  // caller's record buffer of pairs the solvers could not finish (bia_set_failure_log);
  // log_stats: counters for steps that log without statistics of their own
  double rkf_rel = 0.0, rkf_abs = 1e-16;
  // largest tile-partial workspace of a step before a single system runs in stages (bytes, 0 =
  // no limit); bia_set_workspace_limit
  size_t workspace_limit = (size_t)1 << 30;
  void* log_rec = nullptr;
  unsigned long long* log_count = nullptr;
  int log_cap = 0;
  bia::DevBuf log_stats;
  // optional timing of the dominant (pair) kernel with HIP events on the context's stream
  // pair-kernel timing: event pairs recorded around the last kProfSlots launches (a ring), and
  // the launches recorded since profiling was switched on
  int profiling = 0;
  // ring 0: the pair kernel; ring 1: the tile-partial reduction kernel of the same steps
  hipEvent_t ev_begin[2][bia::kProfSlots] = {}, ev_end[2][bia::kProfSlots] = {};
  long long ev_count[2] = {0, 0};
  char last_error[256] = {0};
};

#define BIA_HIP_TRY(ctx, expr)                                                                \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) {                                                                   \
      snprintf((ctx)->last_error, sizeof((ctx)->last_error), "%s:%d: %s: %s", __FILE__,       \
               __LINE__, #expr, hipGetErrorString(e_));                                       \
      return BIA_ERR_HIP;                                                                     \
    }                                                                                         \
  } while (0)

#define BIA_TRY(ctx, expr)                                                                   \
  do {                                                                                       \
    int s_ = (expr);                                                                         \
    if (s_ != BIA_OK) {                                                                      \
      if (!(ctx)->last_error[0])                                                             \
        snprintf((ctx)->last_error, sizeof((ctx)->last_error), "%s:%d: %s -> %d", __FILE__, \
                 __LINE__, #expr, s_);                                                       \
      return s_;                                                                             \
    }                                                                                        \
  } while (0)

namespace bia {
// the table of network k (0: the Kepler drift's start, 1: the reference solvers' starts) and
// whether it is built
struct NetTable {
  const float* t;
  int n, ready;
};
inline NetTable net_table(const bia_ctx* c, int k) {
  return (k == c->mlp_net) ? NetTable{c->mlp_table.as<float>(), c->table_n, c->mlp_ready}
                           : NetTable{c->alt_table.as<float>(), c->alt_table_n, c->alt_ready};
}
// default context used by the legacy Fortran-callable symbols (the reference ABI has no handle)
bia_ctx* default_ctx();
// abort with a message: the legacy symbols have no error return, so failures are loud
[[noreturn]] void legacy_fail(const char* where, int code, bia_ctx* ctx);
// record the profiling events around the pair kernel (no-ops unless profiling is on)
// with `device` current: keep the device address of the no-allocation step's scratch array
// (bia_step.hip), once per device
int noalloc_scratch_init(int device);
int prof_begin(bia_ctx* ctx, int ring = 0);
int prof_end(bia_ctx* ctx, int ring = 0);
// waits for the context's stream when it goes out of scope, so that a call returning early on
// an error leaves no copy pending from or into host memory that the caller may then release
struct StreamDrain {
  bia_ctx* ctx;
  ~StreamDrain() { (void)hipStreamSynchronize(ctx->stream); }
};
}  // namespace bia
