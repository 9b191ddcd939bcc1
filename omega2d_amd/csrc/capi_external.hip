// capi_external.hip - the reference's link-time hook for an external particle velocity solver,
// external_vel_solver_f_ / _tr_f_ / _d_ (declared at src/Influence.h:31-38, called at
// src/Influence.h:74-98 when the environment is not internal and EXTERNAL_VEL_SOLVE is defined).
//
// A reference build linked against this library gets the GPU sum through its own hook, with no
// source change. The hook passes no core, precision or device, so they come from the process
// environment, read on the first call:
//   O2D_EXT_CORE       rm | gauss | wl | v2 | v3   (default v2, the reference's compiled-in
//                                                    core, src/CoreFunc.h:38)
//   O2D_EXT_PRECISION  split | fp64                (default split for the float entries,
//                                                    fp64 for the double entry)
//   O2D_EXT_DEVICE     GPU ordinal                 (default 0)
// Velocities are accumulated unscaled into tu, tv, as the internal loops do; the caller zeroes
// and finalizes them. The return value is the flop count of the reference's convention
// (nt (2 + flops per pair * ns)); on an error it is negative and the reason goes to stderr.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "o2d_runtime.h"

namespace o2d {
namespace {

struct ExtConfig {
  bool read = false;
  bool ok = false;
  int32_t core = O2D_CORE_V2;
  int32_t precision = 0;  // 0: the entry's default
  int32_t device = 0;
};

ExtConfig g_ext;

bool read_ext_config() {
  if (g_ext.read) return g_ext.ok;
  g_ext.read = true;
  g_ext.ok = false;
  if (const char* c = std::getenv("O2D_EXT_CORE")) {
    static const struct {
      const char* name;
      int32_t core;
    } kCores[] = {{"rm", O2D_CORE_RM}, {"gauss", O2D_CORE_GAUSSIAN},
                  {"gaussian", O2D_CORE_GAUSSIAN}, {"wl", O2D_CORE_WL},
                  {"v2", O2D_CORE_V2}, {"v3", O2D_CORE_V3}};
    bool found = false;
    for (const auto& k : kCores)
      if (std::strcmp(c, k.name) == 0) {
        g_ext.core = k.core;
        found = true;
      }
    if (!found) {
      std::fprintf(stderr, "external_vel_solver: O2D_EXT_CORE=%s is not rm|gauss|wl|v2|v3\n", c);
      return false;
    }
  }
  if (const char* p = std::getenv("O2D_EXT_PRECISION")) {
    if (std::strcmp(p, "split") == 0) g_ext.precision = O2D_SPLIT_FP32;
    else if (std::strcmp(p, "fp64") == 0) g_ext.precision = O2D_PLAIN_FP64;
    else {
      std::fprintf(stderr, "external_vel_solver: O2D_EXT_PRECISION=%s is not split|fp64\n", p);
      return false;
    }
  }
  if (const char* d = std::getenv("O2D_EXT_DEVICE")) {
    char* end = nullptr;
    const long v = std::strtol(d, &end, 10);
    if (end == d || *end != '\0' || v < 0 || v > 1023) {
      std::fprintf(stderr, "external_vel_solver: O2D_EXT_DEVICE=%s is not a device ordinal\n", d);
      return false;
    }
    g_ext.device = (int32_t)v;
  }
  g_ext.ok = true;
  return true;
}

float ext_flops(int32_t core, bool point_targets, int ns, int nt) {
  const int per = o2d_hip_flops_per_pair(core, point_targets ? 1 : 0, 0);
  return (float)nt * (2.0f + (float)per * (float)ns);
}

float ext_fail(const char* fn, int rc) {
  std::fprintf(stderr, "%s: status %d: %s\n", fn, rc, o2d_hip_last_error_string());
  return -1.0f;
}

// shared body of the three entries; T is the storage type
template <class T>
float ext_call(const char* fn, const int* ns, const T* sx, const T* sy, const T* ss,
               const T* sr, const int* nt, const T* tx, const T* ty, const T* tr, T* tu, T* tv) {
  if (!read_ext_config()) return -1.0f;
  if (!ns || !nt) return ext_fail(fn, fail(O2D_ERR_ARG, "NULL count"));
  if (*ns < 0 || *nt < 0) return ext_fail(fn, fail(O2D_ERR_SIZE, "negative count"));
  if (*ns == 0 || *nt == 0) return 0.0f;
  const int32_t def = sizeof(T) == sizeof(float) ? O2D_SPLIT_FP32 : O2D_PLAIN_FP64;
  const o2d_env env = {g_ext.core, g_ext.precision ? g_ext.precision : def, O2D_VELONLY,
                       g_ext.device, nullptr};
  int rc;
  if constexpr (sizeof(T) == sizeof(float))
    rc = o2d_hip_points_on_points_f(&env, *ns, sx, sy, ss, sr, *nt, tx, ty, tr, tu, tv, nullptr,
                                    nullptr);
  else
    rc = o2d_hip_points_on_points_d(&env, *ns, sx, sy, ss, sr, *nt, tx, ty, tr, tu, tv, nullptr,
                                    nullptr);
  if (rc != O2D_OK) return ext_fail(fn, rc);
  return ext_flops(g_ext.core, tr == nullptr, *ns, *nt);
}

}  // namespace
}  // namespace o2d

using namespace o2d;

extern "C" {

float external_vel_solver_f_(int* ns, const float* sx, const float* sy, const float* ss,
                             const float* sr, int* nt, const float* tx, const float* ty, float* tu,
                             float* tv) {
  return ext_call<float>("external_vel_solver_f_", ns, sx, sy, ss, sr, nt, tx, ty, nullptr, tu,
                         tv);
}

float external_vel_solver_tr_f_(int* ns, const float* sx, const float* sy, const float* ss,
                                const float* sr, int* nt, const float* tx, const float* ty,
                                const float* tr, float* tu, float* tv) {
  if (nt && *nt > 0 && !tr) {
    clear_error();
    return ext_fail("external_vel_solver_tr_f_", fail(O2D_ERR_ARG, "NULL target radii"));
  }
  return ext_call<float>("external_vel_solver_tr_f_", ns, sx, sy, ss, sr, nt, tx, ty, tr, tu,
                         tv);
}

float external_vel_solver_d_(int* ns, const double* sx, const double* sy, const double* ss,
                             const double* sr, int* nt, const double* tx, const double* ty,
                             double* tu, double* tv) {
  return ext_call<double>("external_vel_solver_d_", ns, sx, sy, ss, sr, nt, tx, ty, nullptr, tu,
                          tv);
}

}  // extern "C"
