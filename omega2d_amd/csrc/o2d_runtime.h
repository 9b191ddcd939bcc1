// o2d_runtime.h - per-device state, buffer pool and error reporting shared by every entry point.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "omega2d_hip.h"

namespace o2d {

constexpr int kMaxDevices = 64;

// Buffer-pool slots. Each slot is one grow-only device allocation owned by the device state.
enum Slot : int {
  kSrcX, kSrcY, kSrcS, kSrcR,
  kTrgX, kTrgY, kTrgR,
  kOutU, kOutV, kOutW, kOutE,
  kStageF0, kStageF1, kStageF2, kStageF3, kStageF4, kStageF5, kStageF6, kStageF7,
  kStageF8, kStageF9, kStageF10,
  kPartials,
  kKeysA, kKeysB, kOrderA, kOrderB, kSortTemp, kBBox,
  kTileSrc, kTileCent,
  kTrgKeys, kTrgOrder, kTrgPos, kTrgW, kSrc64,
  kPanX, kPanY, kPanIdx, kPanVs, kPanSs,
  kPanGeom,
  kMatA, kRhs,
  kLU, kPivots, kInfo,
  kRkPtrs, kRkStage,
  kGradX, kGradY, kGradS, kGradR,
  kNearPart, kNodeNorm, kPanRec, kCutTable, kNearCounts, kNearTiles,
  kInv, kSolveTmp,
  kTreeKeys, kTreeOrder, kTreeTOrder, kTreeSrc, kTreeCells, kTreeMom,
  kPan2X, kPan2Y, kPan2Idx, kPanGeom2,
  kLuPairs, kPanClus, kPanClusG, kPanClusS, kVortSrc,
  kNumSlots
};

struct Device {
  bool ready = false;
  int ordinal = -1;
  int num_cu = 256;
  hipStream_t stream = nullptr;
  void* buf[kNumSlots] = {};
  size_t cap[kNumSlots] = {};
  // LU factors of the last matrix given to o2d_hip_bem_factor_*
  int64_t lu_n = 0;
  bool lu_ready = false;
  // solves made with the current factors; from the second one on, the explicit inverse
  // (built once from the factors) turns each solve into one matrix-vector product
  int64_t lu_solves = 0;
  bool inv_ready = false;
  void* blas = nullptr;  // rocblas_handle
  // the LU's second stream (the trailing update beside the next panel), its rocBLAS handle and
  // two events; created on first use
  hipStream_t aux = nullptr;
  void* blas_aux = nullptr;
  hipEvent_t aux_ev[2] = {};
};

// Error handling: every entry point returns fail(code, fmt, ...) on failure.
int fail(int code, const char* fmt, ...);
void clear_error();

// Make `ord` current, creating its state on first use.
int device(int32_t ord, Device** out);
// Stream of a synchronous host-array call: the caller's, else the library's own.
inline hipStream_t stream_of(const o2d_env* env, Device* d) {
  return env->stream ? static_cast<hipStream_t>(env->stream) : d->stream;
}
// Stream of a device-pointer call: exactly the caller's (NULL is the device's null stream).
inline hipStream_t dev_stream(const o2d_env* env) { return static_cast<hipStream_t>(env->stream); }
// Device buffer for `slot` with at least `bytes`; grows (after draining `s`) when too small.
int slot(Device* d, Slot which, size_t bytes, hipStream_t s, void** out);

// A fixed device buffer handed out in consecutive 256-byte-aligned pieces, for calls that
// must not allocate.
struct Workspace {
  char* base = nullptr;
  size_t cap = 0;
  size_t used = 0;
  // bytes still free after aligning the next piece
  size_t left() const {
    const size_t off = (used + 255) & ~size_t(255);
    return off < cap ? cap - off : 0;
  }
};
// Temporary device memory: `slot` when ws is null, else the next piece of *ws (an error when
// the workspace is too small; nothing is allocated).
int take(Device* d, Workspace* ws, Slot which, size_t bytes, hipStream_t s, void** out);

int check_env(const o2d_env* env);

// The Morton order of n points (x, y device arrays): *order (n entries, in the kOrderB slot, so
// valid until that slot is next used) lists the point indices along the curve. morton.hip.
int morton_order16(Device* d, hipStream_t s, int64_t n, const double* x, const double* y,
                   const uint32_t** order);

// Optional per-kernel timing: KernelTimer t(cls, s); ... launch ...; t.stop();
// records a HIP event pair on s around the launch when timing is enabled.
struct KernelTimer {
  KernelTimer(int cls, hipStream_t s);
  void stop();
  int cls;
  hipStream_t s;
  hipEvent_t a = nullptr, b = nullptr;
};

}  // namespace o2d

#define O2D_HIP(expr)                                                                      \
  do {                                                                                     \
    hipError_t o2d_e_ = (expr);                                                            \
    if (o2d_e_ != hipSuccess)                                                              \
      return ::o2d::fail(O2D_ERR_HIP, "%s failed: %s (%s:%d)", #expr,                      \
                         hipGetErrorString(o2d_e_), __FILE__, __LINE__);                   \
  } while (0)

#define O2D_TRY(expr)            \
  do {                           \
    int o2d_rc_ = (expr);        \
    if (o2d_rc_ != 0) return o2d_rc_; \
  } while (0)

#define O2D_LAUNCH_CHECK()                                                                 \
  do {                                                                                     \
    hipError_t o2d_e_ = hipGetLastError();                                                 \
    if (o2d_e_ != hipSuccess)                                                              \
      return ::o2d::fail(O2D_ERR_HIP, "kernel launch failed: %s (%s:%d)",                  \
                         hipGetErrorString(o2d_e_), __FILE__, __LINE__);                   \
  } while (0)
