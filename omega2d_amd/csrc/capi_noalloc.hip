// capi_noalloc.hip - particles on points for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// o2d_hip_points_on_points_noalloc_dev_d is o2d_hip_points_on_points_dev_d with every
// temporary of the sum (Morton keys and orders, the radix sort's storage, staged sources and
// targets, source-chunk partials) taken from `scratch`, one fixed array per device, instead of
// the library's grow-only buffer pool. Up to kNoAllocMaxPoints sources and targets; the number
// of source chunks is capped so that their partials fit what is left of the array. The call
// allocates nothing: no hipMalloc, no hipFree, no allocating library routine.
//
// Every call on a device shares that device's array and leaves its contents behind (nothing in
// it needs to start at zero: every piece is written before it is read). Calls that overlap on
// one device - on two streams, or from two host threads - race on the array.
#include <mutex>

#include "nbody.h"
#include "o2d_runtime.h"

namespace o2d {

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per device
constexpr int64_t kNoAllocMaxPoints = 262144;       // sources, and targets, per call

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// the device address of `scratch` on each device (the array's name in host code is only a
// host-side handle: kernels and copies get these addresses)
void* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

// with device `ord` current
int scratch_on(int32_t ord, void** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[ord]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(O2D_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", ord,
                  hipGetErrorString(e));
    g_scratch[ord] = p;
  }
  *out = g_scratch[ord];
  return 0;
}

}  // namespace

}  // namespace o2d

using namespace o2d;

extern "C" {

int o2d_hip_points_on_points_noalloc_dev_d(const o2d_env* env, int64_t ns, const double* sx,
                                           const double* sy, const double* sstr,
                                           const double* srad, int64_t nt, const double* tx,
                                           const double* ty, const double* trad, double* tu,
                                           double* tv, double* tw, double* te) {
  clear_error();
  O2D_TRY(check_pp(env, ns, sx, sy, sstr, srad, nt, tx, ty, tu, tv, tw, te));
  if (ns > kNoAllocMaxPoints || nt > kNoAllocMaxPoints)
    return fail(O2D_ERR_SIZE,
                "the no-allocation entry takes at most %lld sources and %lld targets (got %lld "
                "and %lld); use o2d_hip_points_on_points_dev_d",
                (long long)kNoAllocMaxPoints, (long long)kNoAllocMaxPoints, (long long)ns,
                (long long)nt);
  Device* d;
  O2D_TRY(device(env->device, &d));
  void* base;
  O2D_TRY(scratch_on(env->device, &base));
  Workspace ws;
  ws.base = static_cast<char*>(base);
  ws.cap = kScratchBytes;
  double* aux = env->results == O2D_VELANDVORT ? tw : (env->results == O2D_VELANDSHEAR ? te : nullptr);
  return nbody_dev(d, dev_stream(env), env->core, env->precision, env->results, ns, sx, sy, sstr,
                   srad, nt, tx, ty, trad, tu, tv, aux, &ws);
}

}  // extern "C"
