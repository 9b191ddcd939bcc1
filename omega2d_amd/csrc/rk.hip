// rk.hip - fp64 Runge-Kutta position update, ElementBase::move (src/ElementBase.h:276-359).
//
// Stage weights combine in the reference's order: one stage x += (dt*w0)*u0; two or three
// stages u = w0*u0 + w1*u1 (+ w2*u2), x += dt*u. Every operation is an explicitly rounded IEEE
// double operation (no fused multiply-add), so the update is bit-identical to a plain C loop.
#include "nbody.h"
#include "o2d_runtime.h"

namespace o2d {

namespace {

constexpr int kBlock = 256;

struct Stages {
  const double* u[3];
  const double* v[3];
  double w[3];
};

__global__ __launch_bounds__(kBlock) void rk_kernel(int64_t n, double dt, int nstage, Stages st,
                                                    double* __restrict__ x,
                                                    double* __restrict__ y) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  // plain operators on purpose: the pragma keeps exactly these from being fused
  if (nstage == 1) {
    const double f = dt * st.w[0];
    x[i] = x[i] + f * st.u[0][i];
    y[i] = y[i] + f * st.v[0][i];
    return;
  }
  double u = st.w[0] * st.u[0][i] + st.w[1] * st.u[1][i];
  double v = st.w[0] * st.v[0][i] + st.w[1] * st.v[1][i];
  if (nstage == 3) {
    u = u + st.w[2] * st.u[2][i];
    v = v + st.w[2] * st.v[2][i];
  }
  x[i] = x[i] + dt * u;
  y[i] = y[i] + dt * v;
}

}  // namespace

int rk_move_dev(Device* d, hipStream_t s, int64_t n, double dt, int nstage, const double* w,
                const double* const* u, const double* const* v, double* x, double* y) {
  (void)d;
  if (n == 0) return 0;
  Stages st = {};
  for (int k = 0; k < nstage; ++k) {
    st.u[k] = u[k];
    st.v[k] = v[k];
    st.w[k] = w[k];
  }
  rk_kernel<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(n, dt, nstage, st, x, y);
  O2D_LAUNCH_CHECK();
  return 0;
}

static int check_rk(const o2d_env* env, int64_t n, int32_t nstage, const double* w,
                    const double* const* u, const double* const* v, double* x, double* y) {
  if (!env) return fail(O2D_ERR_ARG, "env is NULL");
  if (nstage < 1 || nstage > 3) return fail(O2D_ERR_RESULTS, "nstage must be 1, 2 or 3");
  if (n < 0) return fail(O2D_ERR_SIZE, "negative count");
  if (!w || !u || !v) return fail(O2D_ERR_ARG, "NULL weights or stage arrays");
  if (n > 0) {
    if (!x || !y) return fail(O2D_ERR_ARG, "NULL positions");
    for (int k = 0; k < nstage; ++k)
      if (!u[k] || !v[k]) return fail(O2D_ERR_ARG, "NULL stage velocity %d", k);
  }
  return 0;
}

}  // namespace o2d

using namespace o2d;

extern "C" {

// u and v are host arrays of nstage device pointers
int o2d_hip_rk_move_dev_d(const o2d_env* env, int64_t n, double dt, int32_t nstage,
                          const double* w, const double* const* u, const double* const* v,
                          double* x, double* y) {
  clear_error();
  O2D_TRY(check_rk(env, n, nstage, w, u, v, x, y));
  Device* d;
  O2D_TRY(device(env->device, &d));
  return rk_move_dev(d, dev_stream(env), n, dt, nstage, w, u, v, x, y);
}

int o2d_hip_rk_move_d(const o2d_env* env, int64_t n, double dt, int32_t nstage, const double* w,
                      const double* const* u, const double* const* v, double* x, double* y) {
  clear_error();
  O2D_TRY(check_rk(env, n, nstage, w, u, v, x, y));
  if (n == 0) return 0;
  Device* d;
  O2D_TRY(device(env->device, &d));
  hipStream_t s = stream_of(env, d);
  static const Slot us[3] = {kStageF0, kStageF1, kStageF2};
  static const Slot vs[3] = {kStageF3, kStageF4, kStageF5};
  const double* du[3];
  const double* dv[3];
  for (int k = 0; k < nstage; ++k) {
    double *a, *b;
    O2D_TRY(upload<double>(d, us[k], u[k], n, s, &a));
    O2D_TRY(upload<double>(d, vs[k], v[k], n, s, &b));
    du[k] = a;
    dv[k] = b;
  }
  double *dx, *dy;
  O2D_TRY(upload<double>(d, kTrgX, x, n, s, &dx));
  O2D_TRY(upload<double>(d, kTrgY, y, n, s, &dy));
  O2D_TRY(rk_move_dev(d, s, n, dt, nstage, w, du, dv, dx, dy));
  O2D_TRY(download<double>(x, dx, n, s));
  O2D_TRY(download<double>(y, dy, n, s));
  O2D_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // extern "C"
