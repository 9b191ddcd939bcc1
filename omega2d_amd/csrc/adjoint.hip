// adjoint.hip - vector-Jacobian product of the point-target velocity sum (no reference
// counterpart). For the unscaled sum
//   tu_k = -sum_j s_j K(d_kj^2; r_j) dy_kj,  tv_k = sum_j s_j K dx_kj,   dx_kj = tx_k - sx_j
// and upstream gradients gu_k, gv_k, every source j receives (bbb = 2 dK/d(d^2))
//   dJ/ds_j = sum_k K (gv dx - gu dy)                       (the transposed Biot-Savart sum)
//   dJ/dx_j = s_j sum_k [gu bbb dx dy - gv (K + bbb dx^2)]
//   dJ/dy_j = s_j sum_k [gu (K + bbb dy^2) - gv bbb dx dy]
//   dJ/dr_j = s_j sum_k dK/dr (gv dx - gu dy)
// The interaction is recomputed pair by pair and never stored. One lane per source, the point
// targets (with their upstream gradients) staged through LDS; fp64 throughout. When there are
// too few sources to fill the GPU the target range is split into chunks whose partials are
// summed in a fixed order.
#include <algorithm>

#include "nbody.h"
#include "o2d_math.h"
#include "o2d_runtime.h"

namespace o2d {

namespace {

// workgroups per CU the target chunks aim for (32: 4% faster than 8 on 250k sources x 4,096
// probes, where the 16 probe tiles cap the split)
#ifndef O2D_WANT_ADJ
#define O2D_WANT_ADJ 32
#endif

constexpr int kBlock = 256;
constexpr int kTile = 256;
constexpr int kTpl = 2;

struct alignas(16) Probe {
  double x, y, gu, gv;
};

template <int CORE, int TPL>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8)))
void adjoint_kernel(
    int64_t ns, const double* __restrict__ sx, const double* __restrict__ sy,
    const double* __restrict__ ss, const double* __restrict__ sr, int64_t nt,
    const double* __restrict__ tx, const double* __restrict__ ty, const double* __restrict__ gu,
    const double* __restrict__ gv, int64_t chunk, double* __restrict__ ox, double* __restrict__ oy,
    double* __restrict__ os, double* __restrict__ orr, int direct) {
  __shared__ Probe tile[kTile];
  __shared__ double4 box[kBlock / 64];
  const int tid = threadIdx.x;
  const int64_t jbase = (int64_t)blockIdx.x * (kBlock * TPL) + tid;
  double px[TPL], py[TPL], pr[TPL], ax[TPL], ay[TPL], as[TPL], ar[TPL];
#pragma unroll
  for (int t = 0; t < TPL; ++t) {
    const int64_t j = jbase + (int64_t)t * kBlock;
    const int64_t jj = j < ns ? j : ns - 1;
    px[t] = sx[jj];
    py[t] = sy[jj];
    pr[t] = sr[jj];
    ax[t] = ay[t] = as[t] = ar[t] = 0.0;
  }
  const int64_t kb = (int64_t)blockIdx.y * chunk;
  const int64_t ke = kb + chunk < nt ? kb + chunk : nt;
  for (int64_t k0 = kb; k0 < ke; k0 += kTile) {
    const int64_t k = k0 + tid;
    // past the chunk's end: a copy of its last probe with zero upstream gradient, which
    // contributes exactly 0 (a probe anywhere else could coincide with a source and make 0 * NaN)
    const int64_t kk = k < ke ? k : ke - 1;
    Probe p;
    p.x = tx[kk];
    p.y = ty[kk];
    p.gu = k < ke ? gu[kk] : 0.0;
    p.gv = k < ke ? gv[kk] : 0.0;
    __syncthreads();
    tile[tid] = p;
    if constexpr (CORE == O2D_CORE_GAUSSIAN) {
      // bounding box of this tile's probes (padding repeats a probe of the tile)
      double b0 = p.x, b1 = p.x, b2 = p.y, b3 = p.y;
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) {
        b0 = fmin(b0, __shfl_xor(b0, m));
        b1 = fmax(b1, __shfl_xor(b1, m));
        b2 = fmin(b2, __shfl_xor(b2, m));
        b3 = fmax(b3, __shfl_xor(b3, m));
      }
      if ((tid & 63) == 0) box[tid / 64] = make_double4(b0, b1, b2, b3);
    }
    __syncthreads();
    if constexpr (CORE == O2D_CORE_GAUSSIAN) {
      double bx0 = box[0].x, bx1 = box[0].y, by0 = box[0].z, by1 = box[0].w;
#pragma unroll
      for (int w = 1; w < kBlock / 64; ++w) {
        bx0 = fmin(bx0, box[w].x);
        bx1 = fmax(bx1, box[w].y);
        by0 = fmin(by0, box[w].z);
        by1 = fmax(by1, box[w].w);
      }
      // every probe of the tile beyond the compact cutoff of this lane's sources: then
      // K = 1/d^2, bbb = -2/d^4 and dK/dsigma = 0 for every pair, with no branch per pair.
      // With z = dx + i dy, t = 1/z and G = gv - i gu the sums above reduce to
      //   dJ/ds += Re(G t),   (dJ/dx, dJ/dy) / s += (Re(G t^2), -Im(G t^2))
      bool far = true;
#pragma unroll
      for (int t = 0; t < TPL; ++t) {
        const double ex = fmax(fmax(bx0 - px[t], px[t] - bx1), 0.0);
        const double ey = fmax(fmax(by0 - py[t], py[t] - by1), 0.0);
        far = far && (ex * ex + ey * ey > 16.0 * pr[t] * pr[t] * (1.0 + 1e-12));
      }
      if (__all(far)) {
#pragma unroll 4
        for (int q = 0; q < kTile; ++q) {
          const Probe pr_ = tile[q];
#pragma unroll
          for (int t = 0; t < TPL; ++t) {
            const double dx = pr_.x - px[t];
            const double dy = pr_.y - py[t];
            const double kk = rcp64(__builtin_fma(dx, dx, dy * dy));
            const double tr = dx * kk, ti = -dy * kk;  // t = 1/z
            as[t] = __builtin_fma(pr_.gv, tr, __builtin_fma(pr_.gu, ti, as[t]));
            const double sr2 = __builtin_fma(tr, tr, -ti * ti), si2 = 2.0 * tr * ti;  // t^2
            ax[t] = __builtin_fma(pr_.gv, sr2, __builtin_fma(pr_.gu, si2, ax[t]));
            ay[t] = __builtin_fma(pr_.gu, sr2, __builtin_fma(-pr_.gv, si2, ay[t]));
          }
        }
        continue;
      }
    }
#pragma unroll 2
    for (int q = 0; q < kTile; ++q) {
      const Probe pr_ = tile[q];
#pragma unroll
      for (int t = 0; t < TPL; ++t) {
        const double dx = pr_.x - px[t];
        const double dy = pr_.y - py[t];
        const double d2 = __builtin_fma(dx, dx, dy * dy);
        double kk, bbb, dks;
        core_adj64<CORE>(dx, dy, d2, pr[t], kk, bbb, dks);
        const double w = __builtin_fma(pr_.gv, dx, -pr_.gu * dy);
        as[t] = __builtin_fma(kk, w, as[t]);
        ar[t] = __builtin_fma(dks, w, ar[t]);
        const double bxy = bbb * dx * dy;
        ax[t] += __builtin_fma(pr_.gu, bxy, -pr_.gv * __builtin_fma(bbb, dx * dx, kk));
        ay[t] += __builtin_fma(pr_.gu, __builtin_fma(bbb, dy * dy, kk), -pr_.gv * bxy);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < TPL; ++t) {
    const int64_t j = jbase + (int64_t)t * kBlock;
    if (j >= ns) continue;
    const double s = ss[j];
    const double gx = s * ax[t], gy = s * ay[t], gsv = as[t], gr = s * ar[t];
    if (direct) {
      ox[j] += gx;
      oy[j] += gy;
      os[j] += gsv;
      orr[j] += gr;
    } else {
      const int64_t o = (int64_t)blockIdx.y * ns + j;
      ox[o] = gx;
      oy[o] = gy;
      os[o] = gsv;
      orr[o] = gr;
    }
  }
}

__global__ __launch_bounds__(kBlock) void reduce4_kernel(int64_t nchunk, int64_t n,
                                                         const double* __restrict__ part,
                                                         double* __restrict__ ox,
                                                         double* __restrict__ oy,
                                                         double* __restrict__ os,
                                                         double* __restrict__ orr) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  double a = 0.0, b = 0.0, c = 0.0, e = 0.0;
  const int64_t stride = nchunk * n;
  for (int64_t k = 0; k < nchunk; ++k) {
    a += part[k * n + j];
    b += part[stride + k * n + j];
    c += part[2 * stride + k * n + j];
    e += part[3 * stride + k * n + j];
  }
  ox[j] += a;
  oy[j] += b;
  os[j] += c;
  orr[j] += e;
}

template <int CORE>
void launch_adj(dim3 grid, hipStream_t s, int64_t ns, const double* sx, const double* sy,
                const double* ss, const double* sr, int64_t nt, const double* tx,
                const double* ty, const double* gu, const double* gv, int64_t chunk, double* ox,
                double* oy, double* os, double* orr, int direct) {
  adjoint_kernel<CORE, kTpl><<<grid, kBlock, 0, s>>>(ns, sx, sy, ss, sr, nt, tx, ty, gu, gv, chunk,
                                                     ox, oy, os, orr, direct);
}

}  // namespace

int adjoint_dev(Device* d, hipStream_t s, int core, int64_t ns, const double* sx,
                const double* sy, const double* ss, const double* sr, int64_t nt,
                const double* tx, const double* ty, const double* gu, const double* gv,
                double* gx, double* gy, double* gs, double* gr) {
  if (ns == 0 || nt == 0) return 0;
  const int64_t swgs = (ns + kBlock * kTpl - 1) / (kBlock * kTpl);
  const int64_t ntiles = (nt + kTile - 1) / kTile;
  const int64_t want = O2D_WANT_ADJ * (int64_t)d->num_cu;
  int64_t nchunk = swgs >= want ? 1 : std::min<int64_t>((want + swgs - 1) / swgs, ntiles);
  const int64_t per = (ntiles + nchunk - 1) / nchunk;
  nchunk = (ntiles + per - 1) / per;
  const int64_t chunk = per * kTile;
  const int direct = nchunk == 1;
  double *ox = gx, *oy = gy, *os = gs, *orr = gr;
  if (!direct) {
    void* p;
    O2D_TRY(slot(d, kPartials, sizeof(double) * 4 * nchunk * ns, s, &p));
    ox = static_cast<double*>(p);
    oy = ox + nchunk * ns;
    os = oy + nchunk * ns;
    orr = os + nchunk * ns;
  }
  const dim3 grid((unsigned)swgs, (unsigned)nchunk);
  KernelTimer kt(O2D_KT_ADJOINT, s);
#define O2D_LADJ(C) launch_adj<C>(grid, s, ns, sx, sy, ss, sr, nt, tx, ty, gu, gv, chunk, ox, oy, os, orr, direct)
  switch (core) {
    case O2D_CORE_RM: O2D_LADJ(O2D_CORE_RM); break;
    case O2D_CORE_GAUSSIAN: O2D_LADJ(O2D_CORE_GAUSSIAN); break;
    case O2D_CORE_WL: O2D_LADJ(O2D_CORE_WL); break;
    case O2D_CORE_V2: O2D_LADJ(O2D_CORE_V2); break;
    default: O2D_LADJ(O2D_CORE_V3); break;
  }
#undef O2D_LADJ
  kt.stop();
  O2D_LAUNCH_CHECK();
  if (!direct) {
    reduce4_kernel<<<(unsigned)((ns + kBlock - 1) / kBlock), kBlock, 0, s>>>(nchunk, ns, ox, gx, gy,
                                                                             gs, gr);
    O2D_LAUNCH_CHECK();
  }
  return 0;
}

static int check_adj(const o2d_env* env, int64_t ns, const void* sx, const void* sy,
                     const void* ss, const void* sr, int64_t nt, const void* tx, const void* ty,
                     const void* gu, const void* gv, const void* gx, const void* gy,
                     const void* gs, const void* gr) {
  O2D_TRY(check_env(env));
  if (env->precision != O2D_PLAIN_FP64)
    return fail(O2D_ERR_PRECISION, "the adjoint is computed in the plain fp64 form only");
  if (env->results != O2D_VELONLY)
    return fail(O2D_ERR_RESULTS, "the adjoint covers the velocity sum only");
  if (ns < 0 || nt < 0 || ns > 0x7fffffffLL) return fail(O2D_ERR_SIZE, "bad count");
  if (ns > 0 && (!sx || !sy || !ss || !sr || !gx || !gy || !gs || !gr))
    return fail(O2D_ERR_ARG, "NULL source or gradient array");
  if (nt > 0 && (!tx || !ty || !gu || !gv)) return fail(O2D_ERR_ARG, "NULL target array");
  return 0;
}

}  // namespace o2d

using namespace o2d;

extern "C" {

int o2d_hip_points_on_points_adjoint_dev_d(const o2d_env* env, int64_t ns, const double* sx,
                                           const double* sy, const double* sstr,
                                           const double* srad, int64_t nt, const double* tx,
                                           const double* ty, const double* gu, const double* gv,
                                           double* gx, double* gy, double* gs, double* gr) {
  clear_error();
  O2D_TRY(check_adj(env, ns, sx, sy, sstr, srad, nt, tx, ty, gu, gv, gx, gy, gs, gr));
  Device* d;
  O2D_TRY(device(env->device, &d));
  return adjoint_dev(d, dev_stream(env), env->core, ns, sx, sy, sstr, srad, nt, tx, ty, gu, gv,
                     gx, gy, gs, gr);
}

int o2d_hip_points_on_points_adjoint_d(const o2d_env* env, int64_t ns, const double* sx,
                                       const double* sy, const double* sstr, const double* srad,
                                       int64_t nt, const double* tx, const double* ty,
                                       const double* gu, const double* gv, double* gx,
                                       double* gy, double* gs, double* gr) {
  clear_error();
  O2D_TRY(check_adj(env, ns, sx, sy, sstr, srad, nt, tx, ty, gu, gv, gx, gy, gs, gr));
  if (ns == 0 || nt == 0) return 0;
  Device* d;
  O2D_TRY(device(env->device, &d));
  hipStream_t s = stream_of(env, d);
  double *dsx, *dsy, *dss, *dsr, *dtx, *dty, *dgu, *dgv, *dgx, *dgy, *dgs, *dgr;
  O2D_TRY(upload<double>(d, kSrcX, sx, ns, s, &dsx));
  O2D_TRY(upload<double>(d, kSrcY, sy, ns, s, &dsy));
  O2D_TRY(upload<double>(d, kSrcS, sstr, ns, s, &dss));
  O2D_TRY(upload<double>(d, kSrcR, srad, ns, s, &dsr));
  O2D_TRY(upload<double>(d, kTrgX, tx, nt, s, &dtx));
  O2D_TRY(upload<double>(d, kTrgY, ty, nt, s, &dty));
  O2D_TRY(upload<double>(d, kOutU, gu, nt, s, &dgu));
  O2D_TRY(upload<double>(d, kOutV, gv, nt, s, &dgv));
  O2D_TRY(upload<double>(d, kGradX, gx, ns, s, &dgx));
  O2D_TRY(upload<double>(d, kGradY, gy, ns, s, &dgy));
  O2D_TRY(upload<double>(d, kGradS, gs, ns, s, &dgs));
  O2D_TRY(upload<double>(d, kGradR, gr, ns, s, &dgr));
  O2D_TRY(adjoint_dev(d, s, env->core, ns, dsx, dsy, dss, dsr, nt, dtx, dty, dgu, dgv, dgx, dgy, dgs,
                      dgr));
  O2D_TRY(download<double>(gx, dgx, ns, s));
  O2D_TRY(download<double>(gy, dgy, ns, s));
  O2D_TRY(download<double>(gs, dgs, ns, s));
  O2D_TRY(download<double>(gr, dgr, ns, s));
  O2D_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // extern "C"
