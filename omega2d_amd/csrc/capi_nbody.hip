// capi_nbody.hip - C ABI for particles on points (host and device pointer forms) and the
// float/double staging shared by the synchronous entry points.
#include "nbody.h"
#include "o2d_runtime.h"

namespace o2d {

namespace {

__global__ void f2d_kernel(int64_t n, const float* __restrict__ a, double* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}

__global__ void d2f_kernel(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

}  // namespace

int upload_f2d(Device* d, Slot stage, Slot dst, const float* h, int64_t n, hipStream_t s,
               double** out) {
  float* f;
  O2D_TRY(upload<float>(d, stage, h, n, s, &f));
  void* p;
  O2D_TRY(slot(d, dst, sizeof(double) * (n > 0 ? n : 1), s, &p));
  if (n > 0) {
    f2d_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, f, static_cast<double*>(p));
    O2D_LAUNCH_CHECK();
  }
  *out = static_cast<double*>(p);
  return 0;
}

int download_d2f(Device* d, Slot stage, float* h, const double* dptr, int64_t n, hipStream_t s) {
  if (n <= 0) return 0;
  void* p;
  O2D_TRY(slot(d, stage, sizeof(float) * n, s, &p));
  d2f_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, dptr, static_cast<float*>(p));
  O2D_LAUNCH_CHECK();
  return download<float>(h, static_cast<const float*>(p), n, s);
}

int check_pp(const o2d_env* env, int64_t ns, const void* sx, const void* sy,
                    const void* ss, const void* sr, int64_t nt, const void* tx, const void* ty,
                    const void* tu, const void* tv, const void* tw, const void* te) {
  O2D_TRY(check_env(env));
  if (ns < 0 || nt < 0) return fail(O2D_ERR_SIZE, "negative count (ns=%lld, nt=%lld)",
                                    (long long)ns, (long long)nt);
  if (nt > 0x7fffffffLL) return fail(O2D_ERR_SIZE, "nt=%lld exceeds 2^31-1", (long long)nt);
  if (ns > 0 && (!sx || !sy || !ss || !sr)) return fail(O2D_ERR_ARG, "NULL source array");
  if (nt > 0 && (!tx || !ty || !tu || !tv)) return fail(O2D_ERR_ARG, "NULL target array");
  if (nt > 0 && env->results == O2D_VELANDVORT && !tw)
    return fail(O2D_ERR_ARG, "results=VELANDVORT needs tw");
  if (nt > 0 && env->results == O2D_VELANDSHEAR && !te)
    return fail(O2D_ERR_ARG, "results=VELANDSHEAR needs te");
  return 0;
}

}  // namespace o2d

using namespace o2d;

extern "C" {

int o2d_hip_points_on_points_dev_d(const o2d_env* env, int64_t ns, const double* sx,
                                   const double* sy, const double* sstr, const double* srad,
                                   int64_t nt, const double* tx, const double* ty,
                                   const double* trad, double* tu, double* tv, double* tw,
                                   double* te) {
  clear_error();
  O2D_TRY(check_pp(env, ns, sx, sy, sstr, srad, nt, tx, ty, tu, tv, tw, te));
  Device* d;
  O2D_TRY(device(env->device, &d));
  double* aux = env->results == O2D_VELANDVORT ? tw : (env->results == O2D_VELANDSHEAR ? te : nullptr);
  return nbody_dev(d, dev_stream(env), env->core, env->precision, env->results, ns, sx, sy, sstr,
                   srad, nt, tx, ty, trad, tu, tv, aux);
}

int o2d_hip_points_on_points_d(const o2d_env* env, int64_t ns, const double* sx, const double* sy,
                               const double* sstr, const double* srad, int64_t nt,
                               const double* tx, const double* ty, const double* trad,
                               double* tu, double* tv, double* tw, double* te) {
  clear_error();
  O2D_TRY(check_pp(env, ns, sx, sy, sstr, srad, nt, tx, ty, tu, tv, tw, te));
  if (ns == 0 || nt == 0) return 0;
  Device* d;
  O2D_TRY(device(env->device, &d));
  hipStream_t s = stream_of(env, d);
  double *dsx, *dsy, *dss, *dsr, *dtx, *dty, *dtr = nullptr, *du, *dv, *dw = nullptr;
  O2D_TRY(upload<double>(d, kSrcX, sx, ns, s, &dsx));
  O2D_TRY(upload<double>(d, kSrcY, sy, ns, s, &dsy));
  O2D_TRY(upload<double>(d, kSrcS, sstr, ns, s, &dss));
  O2D_TRY(upload<double>(d, kSrcR, srad, ns, s, &dsr));
  // a collection acting on itself (the usual particle case) passes the same host arrays as
  // sources and targets: their device copies serve both
  const bool self = nt == ns;
  if (self && tx == sx) dtx = dsx;
  else O2D_TRY(upload<double>(d, kTrgX, tx, nt, s, &dtx));
  if (self && ty == sy) dty = dsy;
  else O2D_TRY(upload<double>(d, kTrgY, ty, nt, s, &dty));
  if (trad && self && trad == srad) dtr = dsr;
  else if (trad) O2D_TRY(upload<double>(d, kTrgR, trad, nt, s, &dtr));
  O2D_TRY(upload<double>(d, kOutU, tu, nt, s, &du));
  O2D_TRY(upload<double>(d, kOutV, tv, nt, s, &dv));
  double* haux = env->results == O2D_VELANDVORT ? tw : (env->results == O2D_VELANDSHEAR ? te : nullptr);
  if (haux) O2D_TRY(upload<double>(d, kOutW, haux, nt, s, &dw));
  O2D_TRY(nbody_dev(d, s, env->core, env->precision, env->results, ns, dsx, dsy, dss, dsr, nt, dtx,
                    dty, dtr, du, dv, dw));
  O2D_TRY(download<double>(tu, du, nt, s));
  O2D_TRY(download<double>(tv, dv, nt, s));
  if (haux) O2D_TRY(download<double>(haux, dw, nt, s));
  O2D_HIP(hipStreamSynchronize(s));
  return 0;
}

int o2d_hip_points_on_points_f(const o2d_env* env, int64_t ns, const float* sx, const float* sy,
                               const float* sstr, const float* srad, int64_t nt, const float* tx,
                               const float* ty, const float* trad, float* tu, float* tv,
                               float* tw, float* te) {
  clear_error();
  O2D_TRY(check_pp(env, ns, sx, sy, sstr, srad, nt, tx, ty, tu, tv, tw, te));
  if (ns == 0 || nt == 0) return 0;
  Device* d;
  O2D_TRY(device(env->device, &d));
  hipStream_t s = stream_of(env, d);
  double *dsx, *dsy, *dss, *dsr, *dtx, *dty, *dtr = nullptr, *du, *dv, *dw = nullptr;
  O2D_TRY(upload_f2d(d, kStageF0, kSrcX, sx, ns, s, &dsx));
  O2D_TRY(upload_f2d(d, kStageF1, kSrcY, sy, ns, s, &dsy));
  O2D_TRY(upload_f2d(d, kStageF2, kSrcS, sstr, ns, s, &dss));
  O2D_TRY(upload_f2d(d, kStageF3, kSrcR, srad, ns, s, &dsr));
  const bool self = nt == ns;
  if (self && tx == sx) dtx = dsx;
  else O2D_TRY(upload_f2d(d, kStageF4, kTrgX, tx, nt, s, &dtx));
  if (self && ty == sy) dty = dsy;
  else O2D_TRY(upload_f2d(d, kStageF5, kTrgY, ty, nt, s, &dty));
  if (trad && self && trad == srad) dtr = dsr;
  else if (trad) O2D_TRY(upload_f2d(d, kStageF6, kTrgR, trad, nt, s, &dtr));
  O2D_TRY(upload_f2d(d, kStageF7, kOutU, tu, nt, s, &du));
  O2D_TRY(upload_f2d(d, kStageF8, kOutV, tv, nt, s, &dv));
  float* haux = env->results == O2D_VELANDVORT ? tw : (env->results == O2D_VELANDSHEAR ? te : nullptr);
  if (haux) O2D_TRY(upload_f2d(d, kStageF9, kOutW, haux, nt, s, &dw));
  O2D_TRY(nbody_dev(d, s, env->core, env->precision, env->results, ns, dsx, dsy, dss, dsr, nt, dtx,
                    dty, dtr, du, dv, dw));
  O2D_TRY(download_d2f(d, kStageF7, tu, du, nt, s));
  O2D_TRY(download_d2f(d, kStageF8, tv, dv, nt, s));
  if (haux) O2D_TRY(download_d2f(d, kStageF9, haux, dw, nt, s));
  O2D_HIP(hipStreamSynchronize(s));
  return 0;
}

int o2d_hip_morton_order_dev_d(const o2d_env* env, int64_t n, const double* x, const double* y,
                               double x0, double y0, double inv_extent, uint32_t* keys,
                               uint32_t* order) {
  clear_error();
  if (!env) return fail(O2D_ERR_ARG, "env is NULL");
  if (n < 0 || n > 0x7fffffffLL) return fail(O2D_ERR_SIZE, "bad count %lld", (long long)n);
  if (n > 0 && (!x || !y || !keys || !order)) return fail(O2D_ERR_ARG, "NULL array");
  Device* d;
  O2D_TRY(device(env->device, &d));
  return morton_order(d, dev_stream(env), n, x, y, false, x0, y0, inv_extent, keys, order);
}

}  // extern "C"
