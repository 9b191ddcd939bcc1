// capi_reflect.hip - C ABI of the particle/panel proximity operations (src/Reflect.h):
// o2d_hip_nearest_distances_*, o2d_hip_reflect_*, o2d_hip_clear_inner_*.
#include <cstring>

#include "nbody.h"
#include "o2d_runtime.h"

#pragma clang fp contract(off)

namespace o2d {

namespace {

constexpr int kModeNearest = 0, kModeReflect = 1, kModeClear = 2;

int check_geometry(int64_t nnodes, const void* nx, const void* ny, int64_t np, const void* idx,
                   int64_t nt, const void* tx, const void* ty) {
  if (nnodes < 0 || np < 0 || nt < 0) return fail(O2D_ERR_SIZE, "negative count");
  if (np > (1LL << 30) || nnodes > 0xffffffffLL) return fail(O2D_ERR_SIZE, "too many panels");
  if (nt > 0x7fffffffLL) return fail(O2D_ERR_SIZE, "too many targets");
  if (np > 0 && nnodes < 1) return fail(O2D_ERR_SIZE, "panels need at least one node");
  if (np > 0 && (!nx || !ny || !idx)) return fail(O2D_ERR_ARG, "NULL panel array");
  if (nt > 0 && (!tx || !ty)) return fail(O2D_ERR_ARG, "NULL target array");
  return 0;
}

// node indices in range and no zero-length panel (src/Reflect.h:131 asserts on one)
int check_panels_host(int64_t nnodes, const double* nx, const double* ny, int64_t np,
                      const uint32_t* idx) {
  for (int64_t j = 0; j < np; ++j) {
    const uint32_t a = idx[2 * j], b = idx[2 * j + 1];
    if ((int64_t)a >= nnodes || (int64_t)b >= nnodes)
      return fail(O2D_ERR_SIZE, "panel %lld names node %u or %u, but there are %lld nodes",
                  (long long)j, a, b, (long long)nnodes);
    const double bx = nx[b] - nx[a], by = ny[b] - ny[a];
    if (bx * bx + by * by == 0.0)
      return fail(O2D_ERR_ARG, "panel %lld has zero length", (long long)j);
  }
  return 0;
}

// host-array form shared by the three operations
int near_host(const o2d_env* env, int mode, int method, double cutoff_mult, double ips,
              int64_t nnodes, const double* nx, const double* ny, int64_t np,
              const uint32_t* idx, const double* pnx, const double* pny, int64_t nt, double* tx,
              double* ty, double* ts, const double* tr, double* dist, double* value,
              int64_t* counts_out) {
  Device* d;
  O2D_TRY(device(env->device, &d));
  hipStream_t s = stream_of(env, d);
  double *dnx = nullptr, *dny = nullptr, *dpnx = nullptr, *dpny = nullptr;
  double *dtx = nullptr, *dty = nullptr, *dts = nullptr, *dtr = nullptr, *ddist = nullptr;
  uint32_t* didx = nullptr;
  if (np > 0) {
    O2D_TRY(upload<double>(d, kPanX, nx, nnodes, s, &dnx));
    O2D_TRY(upload<double>(d, kPanY, ny, nnodes, s, &dny));
    O2D_TRY(upload<uint32_t>(d, kPanIdx, idx, 2 * np, s, &didx));
    if (mode != kModeNearest) {
      O2D_TRY(upload<double>(d, kPanVs, pnx, np, s, &dpnx));
      O2D_TRY(upload<double>(d, kPanSs, pny, np, s, &dpny));
    }
  }
  O2D_TRY(upload<double>(d, kTrgX, tx, nt, s, &dtx));
  O2D_TRY(upload<double>(d, kTrgY, ty, nt, s, &dty));
  if (ts) O2D_TRY(upload<double>(d, kOutV, ts, nt, s, &dts));
  if (tr) O2D_TRY(upload<double>(d, kTrgR, tr, nt, s, &dtr));
  if (dist) {
    void* p;
    O2D_TRY(slot(d, kOutU, sizeof(double) * (nt > 0 ? nt : 1), s, &p));
    ddist = static_cast<double*>(p);
  }
  void* cp;
  O2D_TRY(slot(d, kNearCounts, 64, s, &cp));
  int64_t* dcounts = static_cast<int64_t*>(cp);
  double* dvalue = reinterpret_cast<double*>(dcounts + 4);
  O2D_TRY(near_dev(d, s, mode, method, cutoff_mult, ips, nnodes, dnx, dny, np, didx, dpnx, dpny,
                   nt, dtx, dty, dts, dtr, ddist, mode == kModeNearest ? nullptr : dvalue,
                   mode == kModeNearest ? nullptr : dcounts));
  int64_t hc[5] = {0, 0, 0, 0, 0};
  if (mode == kModeNearest) {
    O2D_TRY(download<double>(dist, ddist, nt, s));
  } else {
    O2D_TRY(download<double>(tx, dtx, nt, s));
    O2D_TRY(download<double>(ty, dty, nt, s));
    if (ts) O2D_TRY(download<double>(ts, dts, nt, s));
    O2D_HIP(hipMemcpyAsync(hc, dcounts, sizeof(hc), hipMemcpyDeviceToHost, s));
  }
  O2D_HIP(hipStreamSynchronize(s));
  if (mode != kModeNearest) {
    double v;
    std::memcpy(&v, &hc[4], sizeof(v));
    if (value) *value = v;
    if (counts_out) {
      counts_out[0] = hc[0];
      counts_out[1] = hc[1];
      counts_out[2] = hc[2];
    }
  }
  return 0;
}

}  // namespace

}  // namespace o2d

using namespace o2d;

extern "C" {

int o2d_hip_nearest_distances_dev_d(const o2d_env* env, int64_t nnodes, const double* nx,
                                    const double* ny, int64_t np, const uint32_t* idx,
                                    int64_t nt, const double* tx, const double* ty,
                                    double* dist) {
  clear_error();
  O2D_TRY(check_env(env));
  O2D_TRY(check_geometry(nnodes, nx, ny, np, idx, nt, tx, ty));
  if (nt > 0 && !dist) return fail(O2D_ERR_ARG, "NULL dist");
  if (nt == 0) return 0;
  Device* d;
  O2D_TRY(device(env->device, &d));
  return near_dev(d, dev_stream(env), kModeNearest, 0, 0.0, 0.0, nnodes, nx, ny, np, idx,
                  nullptr, nullptr, nt, const_cast<double*>(tx), const_cast<double*>(ty),
                  nullptr, nullptr, dist, nullptr, nullptr);
}

int o2d_hip_nearest_distances_d(const o2d_env* env, int64_t nnodes, const double* nx,
                                const double* ny, int64_t np, const uint32_t* idx, int64_t nt,
                                const double* tx, const double* ty, double* dist) {
  clear_error();
  O2D_TRY(check_env(env));
  O2D_TRY(check_geometry(nnodes, nx, ny, np, idx, nt, tx, ty));
  if (nt > 0 && !dist) return fail(O2D_ERR_ARG, "NULL dist");
  if (np > 0) O2D_TRY(check_panels_host(nnodes, nx, ny, np, idx));
  if (nt == 0) return 0;
  return near_host(env, kModeNearest, 0, 0.0, 0.0, nnodes, nx, ny, np, idx, nullptr, nullptr,
                   nt, const_cast<double*>(tx), const_cast<double*>(ty), nullptr, nullptr, dist,
                   nullptr, nullptr);
}

int o2d_hip_reflect_dev_d(const o2d_env* env, int64_t nnodes, const double* nx,
                          const double* ny, int64_t np, const uint32_t* idx, const double* pnx,
                          const double* pny, int64_t nt, double* tx, double* ty,
                          int64_t* counts) {
  clear_error();
  O2D_TRY(check_env(env));
  O2D_TRY(check_geometry(nnodes, nx, ny, np, idx, nt, tx, ty));
  if (np > 0 && (!pnx || !pny)) return fail(O2D_ERR_ARG, "NULL panel normals");
  Device* d;
  O2D_TRY(device(env->device, &d));
  return near_dev(d, dev_stream(env), kModeReflect, 0, 0.0, 0.0, nnodes, nx, ny, np, idx, pnx,
                  pny, nt, tx, ty, nullptr, nullptr, nullptr, nullptr, counts);
}

int o2d_hip_reflect_d(const o2d_env* env, int64_t nnodes, const double* nx, const double* ny,
                      int64_t np, const uint32_t* idx, const double* pnx, const double* pny,
                      int64_t nt, double* tx, double* ty, int64_t* num_reflected) {
  clear_error();
  O2D_TRY(check_env(env));
  O2D_TRY(check_geometry(nnodes, nx, ny, np, idx, nt, tx, ty));
  if (np > 0 && (!pnx || !pny)) return fail(O2D_ERR_ARG, "NULL panel normals");
  if (np > 0) O2D_TRY(check_panels_host(nnodes, nx, ny, np, idx));
  if (num_reflected) *num_reflected = 0;
  if (nt == 0 || np == 0) return 0;
  int64_t c[3];
  O2D_TRY(near_host(env, kModeReflect, 0, 0.0, 0.0, nnodes, nx, ny, np, idx, pnx, pny, nt, tx,
                    ty, nullptr, nullptr, nullptr, nullptr, c));
  if (num_reflected) *num_reflected = c[0];
  return 0;
}

int o2d_hip_clear_inner_dev_d(const o2d_env* env, int32_t method, double cutoff_mult,
                              double ips, int64_t nnodes, const double* nx, const double* ny,
                              int64_t np, const uint32_t* idx, const double* pnx,
                              const double* pny, int64_t nt, double* tx, double* ty,
                              double* tstr, const double* trad, double* circ_removed,
                              int64_t* counts) {
  clear_error();
  O2D_TRY(check_env(env));
  if (method != 0 && method != 1) return fail(O2D_ERR_ARG, "clear method must be 0 or 1");
  O2D_TRY(check_geometry(nnodes, nx, ny, np, idx, nt, tx, ty));
  if (np > 0 && (!pnx || !pny)) return fail(O2D_ERR_ARG, "NULL panel normals");
  if (method == 0 && trad && nt > 0 && !tstr)
    return fail(O2D_ERR_ARG, "method 0 on particles needs their strengths");
  Device* d;
  O2D_TRY(device(env->device, &d));
  return near_dev(d, dev_stream(env), kModeClear, method, cutoff_mult, ips, nnodes, nx, ny, np,
                  idx, pnx, pny, nt, tx, ty, trad ? tstr : nullptr, trad, nullptr, circ_removed,
                  counts);
}

int o2d_hip_clear_inner_d(const o2d_env* env, int32_t method, double cutoff_mult, double ips,
                          int64_t nnodes, const double* nx, const double* ny, int64_t np,
                          const uint32_t* idx, const double* pnx, const double* pny, int64_t nt,
                          double* tx, double* ty, double* tstr, const double* trad,
                          double* circ_removed, int64_t* num_cropped) {
  clear_error();
  O2D_TRY(check_env(env));
  if (method != 0 && method != 1) return fail(O2D_ERR_ARG, "clear method must be 0 or 1");
  O2D_TRY(check_geometry(nnodes, nx, ny, np, idx, nt, tx, ty));
  if (np > 0 && (!pnx || !pny)) return fail(O2D_ERR_ARG, "NULL panel normals");
  if (method == 0 && trad && nt > 0 && !tstr)
    return fail(O2D_ERR_ARG, "method 0 on particles needs their strengths");
  if (np > 0) O2D_TRY(check_panels_host(nnodes, nx, ny, np, idx));
  if (circ_removed) *circ_removed = 0.0;
  if (num_cropped) *num_cropped = 0;
  if (nt == 0 || np == 0) return 0;
  int64_t c[3];
  double v = 0.0;
  O2D_TRY(near_host(env, kModeClear, method, cutoff_mult, ips, nnodes, nx, ny, np, idx, pnx,
                    pny, nt, tx, ty, trad ? tstr : nullptr, trad, nullptr, &v, c));
  if (circ_removed) *circ_removed = v;
  if (num_cropped) *num_cropped = c[0];
  if (c[1] > 0)
    return fail(O2D_ERR_STATE, "%lld targets still moving after 64 clearing passes",
                (long long)c[1]);
  return 0;
}

}  // extern "C"
