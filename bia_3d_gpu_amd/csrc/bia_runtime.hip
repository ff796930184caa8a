// bia_runtime.hip -- context lifetime, error strings, the default context of the legacy ABI.
#include <algorithm>

#include "bia_runtime.hpp"

namespace bia {

static bia_ctx* g_default = nullptr;
static std::mutex g_default_mu;

bia_ctx* default_ctx() {
  std::lock_guard<std::mutex> lk(g_default_mu);
  if (!g_default) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    bia_ctx* c = nullptr;
    int s = bia_create(&c, dev);
    if (s != BIA_OK) legacy_fail("bia default context", s, c);
    g_default = c;
  }
  return g_default;
}

void legacy_fail(const char* where, int code, bia_ctx* ctx) {
  fprintf(stderr, "libbia_hip: %s failed: %s%s%s\n", where, bia_error_string(code),
          (ctx && ctx->last_error[0]) ? " -- " : "", (ctx && ctx->last_error[0]) ? ctx->last_error : "");
  fflush(stderr);
  abort();
}

int prof_begin(bia_ctx* ctx, int ring) {
  if (!ctx->profiling) return BIA_OK;
  BIA_HIP_TRY(ctx, hipEventRecord(ctx->ev_begin[ring][ctx->ev_count[ring] % kProfSlots], ctx->stream));
  return BIA_OK;
}

int prof_end(bia_ctx* ctx, int ring) {
  if (!ctx->profiling) return BIA_OK;
  BIA_HIP_TRY(ctx, hipEventRecord(ctx->ev_end[ring][ctx->ev_count[ring] % kProfSlots], ctx->stream));
  ctx->ev_count[ring]++;
  return BIA_OK;
}

// durations of the last min(max, kProfSlots) launches recorded in a ring, oldest first
static int ring_ms(bia_ctx* c, int ring, float* ms, int max, int* n) {
  const long long avail = std::min<long long>(c->ev_count[ring], kProfSlots);
  const int cnt = (int)std::min<long long>(avail, max);
  for (int i = 0; i < cnt; i++) {
    const int k = (int)((c->ev_count[ring] - cnt + i) % kProfSlots);
    BIA_HIP_TRY(c, hipEventSynchronize(c->ev_end[ring][k]));
    BIA_HIP_TRY(c, hipEventElapsedTime(ms + i, c->ev_begin[ring][k], c->ev_end[ring][k]));
  }
  *n = cnt;
  return BIA_OK;
}


}  // namespace bia

extern "C" {

int bia_version(void) { return 10000; }

const char* bia_error_string(int status) {
  switch (status) {
    case BIA_OK: return "ok";
    case BIA_ERR_ARG: return "invalid argument";
    case BIA_ERR_HIP: return "HIP runtime error";
    case BIA_ERR_OOM: return "device out of memory";
    case BIA_ERR_STATE: return "invalid call sequence";
    case BIA_ERR_RKF: return "RKF45 integration failed";
    case BIA_ERR_IO: return "record file unreadable, unwritable or malformed";
    default: return "unknown status";
  }
}

const char* bia_last_error(bia_ctx* ctx) { return ctx ? ctx->last_error : ""; }

int bia_create(bia_ctx** out, int device) {
  if (!out) return BIA_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return BIA_ERR_HIP;
  if (device < 0 || device >= ndev) return BIA_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return BIA_ERR_HIP;
  // the no-allocation step's scratch address on this device; a failure is reported by that
  // entry point (BIA_ERR_STATE), the others do not need it
  (void)bia::noalloc_scratch_init(device);
  bia_ctx* c = new bia_ctx();
  c->device = device;
  c->stream = nullptr;
  *out = c;
  return BIA_OK;
}

int bia_destroy(bia_ctx* c) {
  if (!c) return BIA_ERR_ARG;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  c->leg_in.release();
  c->leg_out.release();
  c->dim_h.release();
  bia::DevBuf* bufs[] = {&c->part, &c->state, &c->delta, &c->stat_slots, &c->log_stats, &c->inv_state, &c->inv_part,
                         &c->inv_out,
                         &c->inv_tiles, &c->inv_sys, &c->dim_work, &c->dim_ctl,
                         &c->mlp_w, &c->mlp_ws, &c->mlp_table, &c->alt_w, &c->alt_table};
  for (auto* b : bufs) b->release();
  c->plan_items.release();
  c->plan_sys.release();
  c->plan_chunks.release();
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->cap_stream) (void)hipStreamDestroy(c->cap_stream);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  for (int r = 0; r < 2; r++)
    for (int k = 0; k < bia::kProfSlots; k++) {
      if (c->ev_begin[r][k]) (void)hipEventDestroy(c->ev_begin[r][k]);
      if (c->ev_end[r][k]) (void)hipEventDestroy(c->ev_end[r][k]);
    }
  delete c;
  return BIA_OK;
}

int bia_set_stream(bia_ctx* c, void* stream) {
  if (!c) return BIA_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  c->stream = static_cast<hipStream_t>(stream);
  return BIA_OK;
}

int bia_set_profiling(bia_ctx* c, int on) {
  if (!c) return BIA_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  BIA_HIP_TRY(c, hipSetDevice(c->device));
  if (on && !c->ev_begin[0][0]) {
    for (int r = 0; r < 2; r++)
      for (int k = 0; k < bia::kProfSlots; k++) {
        BIA_HIP_TRY(c, hipEventCreate(&c->ev_begin[r][k]));
        BIA_HIP_TRY(c, hipEventCreate(&c->ev_end[r][k]));
      }
  }
  c->profiling = on ? 1 : 0;
  c->ev_count[0] = c->ev_count[1] = 0;
  return BIA_OK;
}

int bia_last_kernel_ms(bia_ctx* c, float* ms) {
  if (!c || !ms) return BIA_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  if (c->ev_count[0] == 0) return BIA_ERR_STATE;
  const int k = (int)((c->ev_count[0] - 1) % bia::kProfSlots);
  BIA_HIP_TRY(c, hipEventSynchronize(c->ev_end[0][k]));
  BIA_HIP_TRY(c, hipEventElapsedTime(ms, c->ev_begin[0][k], c->ev_end[0][k]));
  return BIA_OK;
}

int bia_kernel_ms(bia_ctx* c, float* ms, int max, int* n) {
  if (!c || !ms || !n || max < 0) return BIA_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  return bia::ring_ms(c, 0, ms, max, n);
}

int bia_reduce_kernel_ms(bia_ctx* c, float* ms, int max, int* n) {
  if (!c || !ms || !n || max < 0) return BIA_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  return bia::ring_ms(c, 1, ms, max, n);
}

int bia_synchronize(bia_ctx* c) {
  if (!c) return BIA_ERR_ARG;
  BIA_HIP_TRY(c, hipSetDevice(c->device));
  BIA_HIP_TRY(c, hipStreamSynchronize(c->stream));
  return BIA_OK;
}

void sincos_(double* x, double* s, double* c) { ::sincos(*x, s, c); }

}  // extern "C"
