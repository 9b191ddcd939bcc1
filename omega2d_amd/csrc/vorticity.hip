// vorticity.hip - Gaussian vorticity of particles at targets, points_affect_points_vorticity
// (src/InfluenceVort.h:47-126, the x86 loop :103-117):
//   tw_i += sum_j 2 s_j exp(-d_ij^2 / r_j^2) / r_j^2      (unscaled; finalize_vels adds 1/2pi)
// in fp64. Sources are staged through LDS as (x, y, 1/r^2, 2 s/r^2); each lane owns two
// targets. A pair with d^2/r^2 > 746 contributes exactly +-0 (exp underflows to 0 in double), so
// its exponential is skipped without changing the sum. Few targets split the source range into
// chunks whose fp64 partials are summed in chunk order.
#include "nbody.h"
#include "o2d_runtime.h"

// workgroups per CU the source chunks aim for (128: 17.1 -> 13.8 ms for 200k particles on
// themselves, against 8)
#ifndef O2D_WANT_VORT
#define O2D_WANT_VORT 128
#endif

namespace o2d {

namespace {

constexpr int kBlock = 256;
constexpr int kTile = 256;
constexpr int kTpl = 2;
constexpr double kUnderflow = 746.0;  // exp(-x) == 0 in double for every x above this

struct VSrc {
  double x, y, rr, c;  // position, 1/r^2, 2 s / r^2
};

__global__ __launch_bounds__(kBlock) void vort_kernel(int64_t ns, const double* __restrict__ sx,
                                                      const double* __restrict__ sy,
                                                      const double* __restrict__ ss,
                                                      const double* __restrict__ sr, int64_t nt,
                                                      const double* __restrict__ tx,
                                                      const double* __restrict__ ty,
                                                      int64_t chunk, double* __restrict__ out,
                                                      int direct) {
  __shared__ VSrc tile[kTile];
  const int64_t j0 = (int64_t)blockIdx.y * chunk;
  const int64_t j1 = (j0 + chunk < ns) ? j0 + chunk : ns;
  double px[kTpl], py[kTpl], acc[kTpl];
  int64_t ti[kTpl];
#pragma unroll
  for (int t = 0; t < kTpl; ++t) {
    ti[t] = ((int64_t)blockIdx.x * kTpl + t) * kBlock + threadIdx.x;
    const bool ok = ti[t] < nt;
    px[t] = ok ? tx[ti[t]] : 0.0;
    py[t] = ok ? ty[ti[t]] : 0.0;
    acc[t] = 0.0;
  }
  for (int64_t base = j0; base < j1; base += kTile) {
    const int n = (int)((j1 - base) < kTile ? (j1 - base) : kTile);
    __syncthreads();
    if ((int)threadIdx.x < n) {
      const int64_t j = base + threadIdx.x;
      const double r2 = sr[j] * sr[j];
      const double rr = 1.0 / r2;
      tile[threadIdx.x] = VSrc{sx[j], sy[j], rr, 2.0 * ss[j] * rr};
    }
    __syncthreads();
    for (int k = 0; k < n; ++k) {
      const VSrc q = tile[k];
#pragma unroll
      for (int t = 0; t < kTpl; ++t) {
        const double dx = px[t] - q.x, dy = py[t] - q.y;
        const double x = (dx * dx + dy * dy) * q.rr;
        if (x < kUnderflow) acc[t] += q.c * exp(-x);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < kTpl; ++t) {
    if (ti[t] >= nt) continue;
    if (direct)
      out[ti[t]] += acc[t];
    else
      out[(int64_t)blockIdx.y * nt + ti[t]] = acc[t];
  }
}

__global__ __launch_bounds__(kBlock) void vort_reduce_kernel(int64_t nchunk, int64_t nt,
                                                             const double* __restrict__ part,
                                                             double* __restrict__ tw) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= nt) return;
  double s = 0.0;
  for (int64_t c = 0; c < nchunk; ++c) s += part[c * nt + i];
  tw[i] += s;
}

}  // namespace

int vorticity_dev(Device* d, hipStream_t s, int64_t ns, const double* sx, const double* sy,
                  const double* ss, const double* sr, int64_t nt, const double* tx,
                  const double* ty, double* tw) {
  if (ns == 0 || nt == 0) return 0;
  const int64_t twgs = (nt + kBlock * kTpl - 1) / (kBlock * kTpl);
  const int64_t ntiles = (ns + kTile - 1) / kTile;
  const int64_t want = O2D_WANT_VORT * (int64_t)d->num_cu;
  int64_t nchunk = twgs >= want ? 1 : (want + twgs - 1) / twgs;
  if (nchunk > ntiles) nchunk = ntiles;
  if (nchunk > 65535) nchunk = 65535;
  const int64_t per = (ntiles + nchunk - 1) / nchunk;
  nchunk = (ntiles + per - 1) / per;
  const int64_t chunk = per * kTile;
  double* out = tw;
  const int direct = nchunk == 1;
  if (!direct) {
    void* p;
    O2D_TRY(slot(d, kPartials, sizeof(double) * nchunk * nt, s, &p));
    out = static_cast<double*>(p);
  }
  {
    KernelTimer kt(O2D_KT_VORT, s);
    vort_kernel<<<dim3((unsigned)twgs, (unsigned)nchunk), kBlock, 0, s>>>(
        ns, sx, sy, ss, sr, nt, tx, ty, chunk, out, direct);
    kt.stop();
    O2D_LAUNCH_CHECK();
  }
  if (!direct) {
    vort_reduce_kernel<<<(unsigned)((nt + kBlock - 1) / kBlock), kBlock, 0, s>>>(nchunk, nt, out,
                                                                                 tw);
    O2D_LAUNCH_CHECK();
  }
  return 0;
}

}  // namespace o2d

using namespace o2d;

extern "C" {

static int check_vort(const o2d_env* env, int64_t ns, const void* sx, const void* sy,
                      const void* ss, const void* sr, int64_t nt, const void* tx,
                      const void* ty, const void* tw) {
  O2D_TRY(check_env(env));  // any precision form: the sum is always fp64
  if (ns < 0 || nt < 0) return fail(O2D_ERR_SIZE, "negative count");
  if (ns > 0x7fffffffLL || nt > 0x7fffffffLL) return fail(O2D_ERR_SIZE, "count too large");
  if (ns > 0 && (!sx || !sy || !ss || !sr)) return fail(O2D_ERR_ARG, "NULL source array");
  if (nt > 0 && (!tx || !ty || !tw)) return fail(O2D_ERR_ARG, "NULL target array");
  return 0;
}

int o2d_hip_points_vorticity_dev_d(const o2d_env* env, int64_t ns, const double* sx,
                                   const double* sy, const double* sstr, const double* srad,
                                   int64_t nt, const double* tx, const double* ty, double* tw) {
  clear_error();
  O2D_TRY(check_vort(env, ns, sx, sy, sstr, srad, nt, tx, ty, tw));
  if (ns == 0 || nt == 0) return 0;
  Device* d;
  O2D_TRY(device(env->device, &d));
  return vorticity_dev(d, dev_stream(env), ns, sx, sy, sstr, srad, nt, tx, ty, tw);
}

int o2d_hip_points_vorticity_d(const o2d_env* env, int64_t ns, const double* sx,
                               const double* sy, const double* sstr, const double* srad,
                               int64_t nt, const double* tx, const double* ty, double* tw) {
  clear_error();
  O2D_TRY(check_vort(env, ns, sx, sy, sstr, srad, nt, tx, ty, tw));
  if (ns == 0 || nt == 0) return 0;
  Device* d;
  O2D_TRY(device(env->device, &d));
  hipStream_t s = stream_of(env, d);
  double *dsx, *dsy, *dss, *dsr, *dtx, *dty, *dtw;
  O2D_TRY(upload<double>(d, kSrcX, sx, ns, s, &dsx));
  O2D_TRY(upload<double>(d, kSrcY, sy, ns, s, &dsy));
  O2D_TRY(upload<double>(d, kSrcS, sstr, ns, s, &dss));
  O2D_TRY(upload<double>(d, kSrcR, srad, ns, s, &dsr));
  O2D_TRY(upload<double>(d, kTrgX, tx, nt, s, &dtx));
  O2D_TRY(upload<double>(d, kTrgY, ty, nt, s, &dty));
  O2D_TRY(upload<double>(d, kOutW, tw, nt, s, &dtw));
  O2D_TRY(vorticity_dev(d, s, ns, dsx, dsy, dss, dsr, nt, dtx, dty, dtw));
  O2D_TRY(download<double>(tw, dtw, nt, s));
  O2D_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // extern "C"
