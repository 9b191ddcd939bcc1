// o2d_runtime.hip - device state, buffer pool, errors and the small runtime entry points.
#include "o2d_runtime.h"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

namespace o2d {

namespace {
thread_local char g_err[1024] = "";
Device g_dev[kMaxDevices];
std::mutex g_init_mu;
}  // namespace

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

void clear_error() { g_err[0] = '\0'; }

int device(int32_t ord, Device** out) {
  if (ord < 0 || ord >= kMaxDevices) return fail(O2D_ERR_ARG, "device ordinal %d out of range", ord);
  int count = 0;
  O2D_HIP(hipGetDeviceCount(&count));
  if (ord >= count) return fail(O2D_ERR_ARG, "device %d not present (%d visible)", ord, count);
  O2D_HIP(hipSetDevice(ord));
  Device* d = &g_dev[ord];
  if (!d->ready) {
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (!d->ready) {
      hipDeviceProp_t prop;
      O2D_HIP(hipGetDeviceProperties(&prop, ord));
      d->num_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
      // default priority on purpose: streams of different priorities are never mixed here
      O2D_HIP(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
      d->ordinal = ord;
      d->ready = true;
    }
  }
  *out = d;
  return 0;
}

int slot(Device* d, Slot which, size_t bytes, hipStream_t s, void** out) {
  if (bytes == 0) bytes = 256;
  if (d->cap[which] < bytes) {
    if (d->buf[which]) {
      // the old buffer may still be read by work queued on any stream of this device
      (void)s;
      O2D_HIP(hipDeviceSynchronize());
      O2D_HIP(hipFree(d->buf[which]));
      d->buf[which] = nullptr;
      d->cap[which] = 0;
    }
    size_t want = bytes + bytes / 4;  // headroom so slowly growing inputs do not reallocate
    want = (want + 4095) & ~size_t(4095);
    void* p = nullptr;
    O2D_HIP(hipMalloc(&p, want));
    d->buf[which] = p;
    d->cap[which] = want;
  }
  *out = d->buf[which];
  return 0;
}

int take(Device* d, Workspace* ws, Slot which, size_t bytes, hipStream_t s, void** out) {
  if (!ws) return slot(d, which, bytes, s, out);
  if (bytes == 0) bytes = 256;
  const size_t off = (ws->used + 255) & ~size_t(255);
  if (off > ws->cap || bytes > ws->cap - off)
    return fail(O2D_ERR_SIZE, "the %zu-byte workspace is too small: %zu bytes are in use and "
                              "%zu more are needed", ws->cap, off, bytes);
  *out = ws->base + off;
  ws->used = off + bytes;
  return 0;
}

int check_env(const o2d_env* env) {
  if (!env) return fail(O2D_ERR_ARG, "env is NULL");
  if (env->core < O2D_CORE_RM || env->core > O2D_CORE_V3)
    return fail(O2D_ERR_CORE, "unknown core %d", env->core);
  if (env->precision != O2D_PLAIN_FP64 && env->precision != O2D_SPLIT_FP32)
    return fail(O2D_ERR_PRECISION, "unknown precision form %d", env->precision);
  if (env->results != O2D_VELONLY && env->results != O2D_VELANDVORT &&
      env->results != O2D_VELANDSHEAR)
    return fail(O2D_ERR_RESULTS, "unsupported result type %d", env->results);
  return 0;
}

void release_blas(Device* d);  // bem.hip

namespace {
struct TimedPair {
  int cls;
  hipEvent_t a, b;
};
bool g_timing = false;
std::mutex g_timing_mu;
std::vector<TimedPair> g_pending;
std::vector<hipEvent_t> g_event_pool;
double g_total_ms[O2D_KT_COUNT] = {};
int64_t g_launches[O2D_KT_COUNT] = {};

hipEvent_t take_event() {
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// resolve every pending pair (waits for its events) into the per-class totals
int drain_pending() {
  for (const TimedPair& p : g_pending) {
    float ms = 0.0f;
    O2D_HIP(hipEventSynchronize(p.b));
    O2D_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    g_total_ms[p.cls] += ms;
    g_launches[p.cls] += 1;
    g_event_pool.push_back(p.a);
    g_event_pool.push_back(p.b);
  }
  g_pending.clear();
  return 0;
}
}  // namespace

KernelTimer::KernelTimer(int c, hipStream_t st) : cls(c), s(st) {
  if (!g_timing) return;
  std::lock_guard<std::mutex> lk(g_timing_mu);
  a = take_event();
  b = take_event();
  // a launch whose start event cannot be recorded is simply not timed
  if (!a || !b || hipEventRecord(a, s) != hipSuccess) {
    if (a) g_event_pool.push_back(a);
    if (b) g_event_pool.push_back(b);
    a = b = nullptr;
  }
}

void KernelTimer::stop() {
  if (!a || !b) return;
  const bool ok = hipEventRecord(b, s) == hipSuccess;
  std::lock_guard<std::mutex> lk(g_timing_mu);
  if (ok) {
    g_pending.push_back({cls, a, b});
  } else {
    g_event_pool.push_back(a);
    g_event_pool.push_back(b);
  }
  a = b = nullptr;
}

}  // namespace o2d

using namespace o2d;

extern "C" {

int o2d_hip_init(int32_t ord) {
  clear_error();
  Device* d = nullptr;
  return device(ord, &d);
}

int o2d_hip_release(void) {
  clear_error();
  for (int i = 0; i < kMaxDevices; ++i) {
    Device* d = &g_dev[i];
    if (!d->ready) continue;
    O2D_HIP(hipSetDevice(i));
    // work on caller streams may still read the buffers
    O2D_HIP(hipDeviceSynchronize());
    release_blas(d);
    for (int s = 0; s < kNumSlots; ++s) {
      if (d->buf[s]) O2D_HIP(hipFree(d->buf[s]));
      d->buf[s] = nullptr;
      d->cap[s] = 0;
    }
    O2D_HIP(hipStreamDestroy(d->stream));
    *d = Device();
  }
  return 0;
}

const char* o2d_hip_last_error_string(void) { return g_err; }

int o2d_hip_timing_enable(int32_t enable) {
  clear_error();
  g_timing = enable != 0;
  return 0;
}

int o2d_hip_timing_read(int32_t cls, double* total_ms, int64_t* launches) {
  clear_error();
  if (cls < 1 || cls >= O2D_KT_COUNT) return fail(O2D_ERR_ARG, "unknown kernel class %d", cls);
  if (!total_ms || !launches) return fail(O2D_ERR_ARG, "NULL output");
  std::lock_guard<std::mutex> lk(g_timing_mu);
  O2D_TRY(drain_pending());
  *total_ms = g_total_ms[cls];
  *launches = g_launches[cls];
  g_total_ms[cls] = 0.0;
  g_launches[cls] = 0;
  return 0;
}

int32_t o2d_hip_abi_version(void) { return O2D_HIP_ABI_VERSION; }

int32_t o2d_hip_flops_per_pair(int32_t core, int32_t point_targets, int32_t panel) {
  // src/Kernels.h:81,95 (10 + core), src/CoreFunc.h flops_tp/tv_nograds, src/Kernels.h:312,362
  if (panel == 1) return 39;
  if (panel == 2) return 47;
  static const int tp[6] = {0, 3, 7, 5, 6, 8};
  static const int tv[6] = {0, 5, 9, 7, 9, 12};
  if (core < O2D_CORE_RM || core > O2D_CORE_V3) return -1;
  return 10 + (point_targets ? tp[core] : tv[core]);
}

}  // extern "C"
