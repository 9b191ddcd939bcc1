// nbody.hip - particles on points, the direct Biot-Savart sum (src/Influence.h:56-416).
//
// Sources and targets are put in Morton order once per call (the stable order of 32-bit keys;
// a collection acting on itself sorts once), so that a wave's targets and a tile's sources are
// each compact whatever the caller's order, and results are scattered back to the targets'
// input order as they are written. The split form always sorts its sources (its groups need
// compact centroids) and sorts other targets for the Gaussian core; the fp64 form sorts both
// for the Gaussian's velocity sum. Those are the sums whose speed depends on locality: the
// Gaussian's wave-uniform far-field loops skip its in-core test. Every other sum keeps the
// input order and skips the sort.
//
// Plain form (fp64): every lane owns TPL targets and runs an unrolled loop over a 256-source
// tile staged in LDS with coalesced loads; all lanes of a wave read the same source (an LDS
// broadcast). Sums are fp64 per target.
//
// Split form (fp32 pair arithmetic): the sorted sources are cut into groups of kGroup
// sources, and stored as fp32 offsets from their group's fp64 centroid.
// Per (target, group) a lane forms its own fp64 offset to that centroid once and rounds it to
// fp32, so every pair difference dx = (t - c) - (s - c) is an fp32 difference of two small
// numbers. Each group's fp32 partial sums are folded into fp64 per-target accumulators.
//
// Both forms split the source range into chunks when there are too few target workgroups to
// fill the GPU; chunk partials are then reduced in a fixed order by a second kernel (no float
// atomics, so results are reproducible run to run).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>

#include "nbody.h"
#include "o2d_math.h"
#include "o2d_runtime.h"

namespace o2d {

namespace {

constexpr int kBlock = 256;  // 4 wave64 per workgroup
constexpr int kTile = 256;   // sources per LDS stage
constexpr int kGroup = 64;   // sources sharing one centroid in the split form
// launch shape, tuned on MI355X (overridable at build time for tuning sweeps)
#ifndef O2D_TPL64
#define O2D_TPL64 2
#endif
#ifndef O2D_TPL32
#define O2D_TPL32 4
#endif
// 5 waves per SIMD give the split kernels 96 VGPRs and no scratch; at 8 they spilled about
// 100 B per lane (written out to memory by every wave) and ran no faster
#ifndef O2D_WAVES32
#define O2D_WAVES32 5
#endif
// source chunks: enough workgroups for this many per CU (each chunk at least kMinChunkTiles
// tiles). Many short-lived workgroups keep every SIMD issuing to the end of the launch; with
// one resident round of long-lived ones, the 1M-particle launches ran 8-10% slower.
#ifndef O2D_WANT32
#define O2D_WANT32 128
#endif
#ifndef O2D_WANT64
#define O2D_WANT64 128
#endif
constexpr int kMinChunkTiles = 16;
constexpr int kStage = 4;  // source tiles brought into LDS per stage of the split kernel
constexpr int kTpl64 = O2D_TPL64;  // targets per lane, plain form
constexpr int kTpl32 = O2D_TPL32;  // targets per lane, split form

struct alignas(16) Src64 {
  double x, y, s, w;
};

// ---------------------------------------------------------------------------- plain fp64

// the pair loop over one staged tile; TR adds each target's radius term per pair
template <int CORE, int RES, bool TR, int TPL>
__device__ __forceinline__ void tile_pairs64(const Src64* __restrict__ tile, const double* px,
                                             const double* py, const double* pw, double* au,
                                             double* av, double* aw) {
#pragma unroll 4
  for (int k = 0; k < kTile; ++k) {
    const Src64 q = tile[k];
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
      const double dx = px[t] - q.x;
      const double dy = py[t] - q.y;
      const double d2 = __builtin_fma(dx, dx, dy * dy);
      const double w = TR ? wsum(q.w, pw[t]) : q.w;
      if constexpr (RES == O2D_VELONLY) {
        const double g = q.s * core_k64<CORE>(dx, dy, d2, w);
        au[t] = __builtin_fma(-g, dy, au[t]);
        av[t] = __builtin_fma(g, dx, av[t]);
      } else {
        double kk, bbb;
        core_kg64<CORE>(dx, dy, d2, w, kk, bbb);
        kk *= q.s;
        bbb *= q.s;
        au[t] = __builtin_fma(-kk, dy, au[t]);
        av[t] = __builtin_fma(kk, dx, av[t]);
        if constexpr (RES == O2D_VELANDVORT) {
          // (bbb dx^2 + K) - (-bbb dy^2 - K)   src/Kernels.h:176-178
          aw[t] += __builtin_fma(bbb, d2, 2.0 * kk);
        } else {
          // sqrt(dudx^2 + asym^2/4)           src/Kernels.h:225-230
          const double dudx = -bbb * dx * dy;
          const double asym = bbb * (dx * dx - dy * dy);
          aw[t] += sqrt(__builtin_fma(dudx, dudx, 0.25 * asym * asym));
        }
      }
    }
  }
}

// src: the staged sources in Morton order; tq, tw: the targets' positions and radius terms in
// Morton order (tw unused unless TR); torder[i]: the input index of sorted target i (torder
// null: the targets are in input order)
template <int CORE, int RES, bool TR, int TPL>
__global__ __launch_bounds__(kBlock) void nbody64_kernel(
    int64_t ns, const Src64* __restrict__ src, int64_t nt, const double2* __restrict__ tq,
    const double* __restrict__ tw, const uint32_t* __restrict__ torder, int64_t chunk,
    double* __restrict__ ou, double* __restrict__ ov, double* __restrict__ ow, int direct) {
  // the Gaussian's velocity sum takes a loop without the per-pair cutoff test for a tile that
  // lies beyond the compact cutoff of every target of the wave (d^2 > 16 w for all its pairs)
  constexpr bool kFar = CORE == O2D_CORE_GAUSSIAN && RES == O2D_VELONLY;
  __shared__ Src64 tile[kTile];
  __shared__ double tbox[kBlock / 64][5];  // per wave: x min/max, y min/max, largest radius term
  const int tid = threadIdx.x;
  const int64_t tbase = (int64_t)blockIdx.x * (kBlock * TPL) + tid;

  double px[TPL], py[TPL], pw[TPL], au[TPL], av[TPL], aw[TPL];
#pragma unroll
  for (int t = 0; t < TPL; ++t) {
    const int64_t i = tbase + (int64_t)t * kBlock;
    const int64_t ii = i < nt ? i : nt - 1;
    const double2 p = tq[ii];
    px[t] = p.x;
    py[t] = p.y;
    pw[t] = TR ? tw[ii] : 0.0;
    au[t] = av[t] = aw[t] = 0.0;
  }
  // when every target of the block has the same radius term w0, it is added to each source's
  // term once as the tile is staged, and the pair loop runs without the per-pair add (the same
  // fp64 sum, so the same result)
  double w0 = 0.0;
  bool uni = false;
  if constexpr (TR) {
    const int64_t i0 = (int64_t)blockIdx.x * (kBlock * TPL);
    w0 = tw[i0 < nt ? i0 : nt - 1];
    bool same = true;
#pragma unroll
    for (int t = 0; t < TPL; ++t) same = same && pw[t] == w0;
    uni = __syncthreads_and(same) != 0;
  }

  const int64_t jb = (int64_t)blockIdx.y * chunk;
  const int64_t je = jb + chunk < ns ? jb + chunk : ns;
  for (int64_t j0 = jb; j0 < je; j0 += kTile) {
    const int64_t j = j0 + tid;
    Src64 s = src[j < je ? j : je - 1];
    if (j >= je) s.s = 0.0;  // zero strength at the chunk's last source: contributes exactly 0
    if (uni) s.w = wsum(s.w, w0);
    if constexpr (kFar) {
      double x0 = s.x, x1 = s.x, y0 = s.y, y1 = s.y, wm = s.w;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        x0 = fmin(x0, __shfl_xor(x0, o, 64));
        x1 = fmax(x1, __shfl_xor(x1, o, 64));
        y0 = fmin(y0, __shfl_xor(y0, o, 64));
        y1 = fmax(y1, __shfl_xor(y1, o, 64));
        wm = fmax(wm, __shfl_xor(wm, o, 64));
      }
      __syncthreads();
      if ((tid & 63) == 0) {
        double* b = tbox[tid >> 6];
        b[0] = x0;
        b[1] = x1;
        b[2] = y0;
        b[3] = y1;
        b[4] = wm;
      }
    } else {
      __syncthreads();
    }
    tile[tid] = s;
    __syncthreads();
    if constexpr (kFar) {
      double x0 = tbox[0][0], x1 = tbox[0][1], y0 = tbox[0][2], y1 = tbox[0][3], wm = tbox[0][4];
#pragma unroll
      for (int k = 1; k < kBlock / 64; ++k) {
        x0 = fmin(x0, tbox[k][0]);
        x1 = fmax(x1, tbox[k][1]);
        y0 = fmin(y0, tbox[k][2]);
        y1 = fmax(y1, tbox[k][3]);
        wm = fmax(wm, tbox[k][4]);
      }
      bool far = true;
#pragma unroll
      for (int t = 0; t < TPL; ++t) {
        const double ex = fmax(fmax(x0 - px[t], px[t] - x1), 0.0);
        const double ey = fmax(fmax(y0 - py[t], py[t] - y1), 0.0);
        // the nearest pair's d^2 clears 16 w by far more than its rounding
        far = far && ex * ex + ey * ey > 16.0 * (uni ? wm : wm + pw[t]) * (1.0 + 1e-12);
      }
      if (__all(far)) {
#pragma unroll 4
        for (int k = 0; k < kTile; ++k) {
          const Src64 q = tile[k];
#pragma unroll
          for (int t = 0; t < TPL; ++t) {
            const double dx = px[t] - q.x;
            const double dy = py[t] - q.y;
            const double g = q.s * rcp64(__builtin_fma(dx, dx, dy * dy));
            au[t] = __builtin_fma(-g, dy, au[t]);
            av[t] = __builtin_fma(g, dx, av[t]);
          }
        }
        continue;
      }
    }
    if (uni)
      tile_pairs64<CORE, RES, false, TPL>(tile, px, py, pw, au, av, aw);
    else
      tile_pairs64<CORE, RES, TR, TPL>(tile, px, py, pw, au, av, aw);
  }

#pragma unroll
  for (int t = 0; t < TPL; ++t) {
    const int64_t i = tbase + (int64_t)t * kBlock;
    if (i >= nt) continue;
    if (direct) {
      const int64_t o = torder ? (int64_t)torder[i] : i;
      ou[o] += au[t];
      ov[o] += av[t];
      if (RES != O2D_VELONLY) ow[o] += aw[t];
    } else {
      const int64_t o = (int64_t)blockIdx.y * nt + i;
      ou[o] = au[t];
      ov[o] = av[t];
      if (RES != O2D_VELONLY) ow[o] = aw[t];
    }
  }
}

// ---------------------------------------------------------------------------- split fp32

// One source group: fp64 centroid, the largest fp32 offset length of its sources from it, and
// their largest radius term (rterm units of the core).
struct alignas(16) GroupInfo {
  double cx, cy;
  float rad, wmax, pad0, pad1;
};

// pair loop over one group; CHECKED = false drops the Gaussian's in-core test (the caller has
// shown that every pair of this group is beyond the compact cutoff, where K = 1/d^2). RES adds
// the vorticity or shear sum into fw (factors from core_kw32).
template <int CORE, int RES, bool TR, int TPL, int UNROLL, bool CHECKED>
__device__ __forceinline__ void group_pairs(const float4* __restrict__ q0, const float* ex,
                                            const float* ey, const float* pw, float* fu,
                                            float* fv, float* fw) {
#pragma unroll UNROLL
  for (int k = 0; k < kGroup; ++k) {
    const float4 q = q0[k];
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
      const float dx = ex[t] - q.x;
      const float dy = ey[t] - q.y;
      float kk;
      if constexpr (CORE == O2D_CORE_GAUSSIAN && !CHECKED) {
        kk = __builtin_amdgcn_rcpf(__builtin_fmaf(dx, dx, dy * dy));
        // beyond the cutoff the shear factor is K and the vorticity factor is dropped
        if constexpr (RES == O2D_VELANDSHEAR) fw[t] += __builtin_fabsf(q.z) * kk;
      } else if constexpr (RES == O2D_VELONLY) {
        const float w = TR ? q.w + pw[t] : q.w;
        kk = core_k32<CORE>(dx, dy, w);
      } else {
        const float w = TR ? q.w + pw[t] : q.w;
        float aux;
        kk = core_kw32<CORE, RES>(dx, dy, w, aux);
        fw[t] = __builtin_fmaf(RES == O2D_VELANDVORT ? q.z : __builtin_fabsf(q.z), aux, fw[t]);
      }
      const float gk = q.z * kk;
      fu[t] = __builtin_fmaf(-gk, dy, fu[t]);
      fv[t] = __builtin_fmaf(gk, dx, fv[t]);
    }
  }
}

// sources prepared by the split prep: fp32 (x - c, y - c, strength, radius term) in Morton
// order, padded to whole tiles; grp[g] describes group g.
//
// A workgroup owns 64 * TPL targets (TPL per lane), and its kWaves waves all hold those same
// targets: each stage brings kStage tiles (kStage * kTile sources) into LDS, and wave w takes
// groups w, w + kWaves, ... of the stage. After the chunk's last stage the waves' fp64 sums of a
// target are added in LDS in wave order, so one workgroup writes one partial per target per
// chunk: a quarter of the chunk partials that a workgroup of four waves with a target set each
// would write, for the same work per wave.
template <int CORE, int RES, bool TR, int TPL, int UNROLL>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(O2D_WAVES32, 8)))
void nbody32_kernel(
    const float4* __restrict__ src, const GroupInfo* __restrict__ grp, int64_t ntiles,
    int64_t nt, const double2* __restrict__ tq, const float* __restrict__ tw,
    const uint32_t* __restrict__ torder, int64_t tiles_per_chunk, double* __restrict__ ou,
    double* __restrict__ ov, double* __restrict__ ow, int direct) {
  constexpr int kGroups = kTile / kGroup;  // groups per tile
  constexpr int kWaves = kBlock / 64;
  constexpr int kTargets = 64 * TPL;       // targets per workgroup
  constexpr bool kAux = RES != O2D_VELONLY;
  constexpr int kComp = kAux ? 3 : 2;
  // the staged tiles, and after the last stage the waves' sums [component][wave][target]
  constexpr int kRedDoubles = kComp * kWaves * kTargets;
  constexpr int kTileDoubles = kStage * kTile * (int)(sizeof(float4) / sizeof(double));
  __shared__ __align__(16) double lds[kRedDoubles > kTileDoubles ? kRedDoubles : kTileDoubles];
  __shared__ GroupInfo tgrp[kStage * kGroups];
  float4* tile = reinterpret_cast<float4*>(lds);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t t0 = (int64_t)blockIdx.x * kTargets;

  double px[TPL], py[TPL], au[TPL], av[TPL], aw[TPL];
  float pw[TPL];
#pragma unroll
  for (int t = 0; t < TPL; ++t) {
    const int64_t i = t0 + lane + 64 * t;
    const int64_t ii = i < nt ? i : nt - 1;
    const double2 p = tq[ii];
    px[t] = p.x;
    py[t] = p.y;
    pw[t] = TR ? tw[ii] : 0.0f;
    au[t] = av[t] = aw[t] = 0.0;
  }
  // when every target of the workgroup has the same radius term w0, it is added to each
  // source's term once as the tile is staged, and the pair loop runs without the per-pair add
  // (the same fp32 sum, so the same result)
  float w0 = 0.0f;
  bool uni = false;
  if constexpr (TR) {
    w0 = tw[t0 < nt ? t0 : nt - 1];
    bool same = true;
#pragma unroll
    for (int t = 0; t < TPL; ++t) same = same && pw[t] == w0;
    uni = __syncthreads_and(same) != 0;
  }

  const int64_t tb = (int64_t)blockIdx.y * tiles_per_chunk;
  const int64_t te = tb + tiles_per_chunk < ntiles ? tb + tiles_per_chunk : ntiles;
  for (int64_t it = tb; it < te; it += kStage) {
    const int nst = te - it < kStage ? (int)(te - it) : kStage;  // tiles in this stage
    // loaded after the barrier, straight into LDS: sixteen staging registers held across it
    // would push the kernel past its register budget (the other resident workgroups of the CU
    // cover the load latency)
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kStage; ++k) {
      if (k < nst) {
        float4 q = src[(it + k) * kTile + tid];
        if (uni) q.w = wsum(q.w, w0);
        tile[k * kTile + tid] = q;
      }
    }
    if (tid < nst * kGroups) tgrp[tid] = grp[it * kGroups + tid];
    __syncthreads();
    const int ng = nst * kGroups;
#pragma unroll 1
    for (int g = wave; g < ng; g += kWaves) {
      const GroupInfo cg = tgrp[g];
      float ex[TPL], ey[TPL], fu[TPL], fv[TPL], fw[TPL];
#pragma unroll
      for (int t = 0; t < TPL; ++t) {
        ex[t] = (float)(px[t] - cg.cx);
        ey[t] = (float)(py[t] - cg.cy);
        fu[t] = 0.0f;
        fv[t] = 0.0f;
        fw[t] = 0.0f;
      }
      const float4* q0 = tile + g * kGroup;
      if constexpr (CORE == O2D_CORE_GAUSSIAN) {
        // every pair of the group is outside the compact core (d^2 > 16 w) when the target's
        // distance to the centroid exceeds the group radius plus 4 sqrt(wmax + w_target);
        // margins of 1e-5 absorb the fp32 rounding of the offsets and of this test
        bool far = true;
#pragma unroll
        for (int t = 0; t < TPL; ++t) {
          const float lim = cg.rad + 4.0f * __builtin_sqrtf(cg.wmax + pw[t]);
          const float e2 = __builtin_fmaf(ex[t], ex[t], ey[t] * ey[t]);
          far = far && (e2 * (1.0f - 1e-5f) > lim * lim * (1.0f + 1e-5f));
        }
        if (__all(far))
          group_pairs<CORE, RES, TR, TPL, 16, false>(q0, ex, ey, pw, fu, fv, fw);
        else if (uni)
          group_pairs<CORE, RES, false, TPL, 1, true>(q0, ex, ey, pw, fu, fv, fw);
        else
          group_pairs<CORE, RES, TR, TPL, 1, true>(q0, ex, ey, pw, fu, fv, fw);
      } else if (uni) {
        group_pairs<CORE, RES, false, TPL, UNROLL, true>(q0, ex, ey, pw, fu, fv, fw);
      } else {
        group_pairs<CORE, RES, TR, TPL, UNROLL, true>(q0, ex, ey, pw, fu, fv, fw);
      }
#pragma unroll
      for (int t = 0; t < TPL; ++t) {
        au[t] += (double)fu[t];
        av[t] += (double)fv[t];
        if (kAux) aw[t] += (double)fw[t];
      }
    }
  }

  // the waves' sums of each target, added in wave order
  __syncthreads();  // every wave is done with the last stage's tiles
  double* red = lds;
#pragma unroll
  for (int t = 0; t < TPL; ++t) {
    const int j = lane + 64 * t;
    red[(0 * kWaves + wave) * kTargets + j] = au[t];
    red[(1 * kWaves + wave) * kTargets + j] = av[t];
    if (kAux) red[(2 * kWaves + wave) * kTargets + j] = aw[t];
  }
  __syncthreads();
  for (int j = tid; j < kTargets; j += kBlock) {
    const int64_t i = t0 + j;
    if (i >= nt) continue;
    double su = red[j], sv = red[kWaves * kTargets + j], sw = 0.0;
    if (kAux) sw = red[2 * kWaves * kTargets + j];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) {
      su += red[w * kTargets + j];
      sv += red[(kWaves + w) * kTargets + j];
      if (kAux) sw += red[(2 * kWaves + w) * kTargets + j];
    }
    if (direct) {
      const int64_t o = torder ? (int64_t)torder[i] : i;
      ou[o] += su;
      ov[o] += sv;
      if (kAux) ow[o] += sw;
    } else {
      const int64_t o = (int64_t)blockIdx.y * nt + i;
      ou[o] = su;
      ov[o] = sv;
      if (kAux) ow[o] = sw;
    }
  }
}

// out[order[i]] += sum over chunks c (in order) of part[c*nt + i] (order null: the identity)
__global__ __launch_bounds__(kBlock) void reduce_chunks_kernel(int64_t nchunk, int64_t nt,
                                                               const uint32_t* __restrict__ order,
                                                               const double* __restrict__ pu,
                                                               const double* __restrict__ pv,
                                                               const double* __restrict__ pw,
                                                               double* __restrict__ ou,
                                                               double* __restrict__ ov,
                                                               double* __restrict__ ow) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= nt) return;
  double su = 0.0, sv = 0.0, sw = 0.0;
  for (int64_t c = 0; c < nchunk; ++c) {
    su += pu[c * nt + i];
    sv += pv[c * nt + i];
    if (pw) sw += pw[c * nt + i];
  }
  const int64_t o = order ? (int64_t)order[i] : i;
  ou[o] += su;
  ov[o] += sv;
  if (pw) ow[o] += sw;
}

// ---------------------------------------------------------------------------- Morton order

struct BBox {
  double x0, y0, inv_extent, pad;
};

__device__ __forceinline__ uint32_t spread16(uint32_t v) {
  v &= 0xFFFFu;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

__device__ __forceinline__ double warp_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double warp_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double warp_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ float warp_maxf(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// per-block min/max of x and y -> part[4*block + {0,1,2,3}] = xmin, ymin, xmax, ymax
__global__ __launch_bounds__(kBlock) void bbox_partial_kernel(int64_t n, const double* __restrict__ x,
                                                              const double* __restrict__ y,
                                                              double* __restrict__ part) {
  __shared__ double red[4][kBlock / 64];
  double xmn = INFINITY, ymn = INFINITY, xmx = -INFINITY, ymx = -INFINITY;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    xmn = fmin(xmn, x[i]);
    xmx = fmax(xmx, x[i]);
    ymn = fmin(ymn, y[i]);
    ymx = fmax(ymx, y[i]);
  }
  xmn = warp_min(xmn);
  ymn = warp_min(ymn);
  xmx = warp_max(xmx);
  ymx = warp_max(ymx);
  const int wv = threadIdx.x / 64, ln = threadIdx.x % 64;
  if (ln == 0) {
    red[0][wv] = xmn;
    red[1][wv] = ymn;
    red[2][wv] = xmx;
    red[3][wv] = ymx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / 64; ++w) {
      xmn = fmin(xmn, red[0][w]);
      ymn = fmin(ymn, red[1][w]);
      xmx = fmax(xmx, red[2][w]);
      ymx = fmax(ymx, red[3][w]);
    }
    part[4 * blockIdx.x + 0] = xmn;
    part[4 * blockIdx.x + 1] = ymn;
    part[4 * blockIdx.x + 2] = xmx;
    part[4 * blockIdx.x + 3] = ymx;
  }
}

// one workgroup of kBlock threads folds the per-block partials (min/max are exact, so the
// order does not matter)
__global__ __launch_bounds__(kBlock) void bbox_final_kernel(int nblk, const double* __restrict__ part,
                                                            BBox* __restrict__ out) {
  __shared__ double red[4][kBlock / 64];
  double xmn = INFINITY, ymn = INFINITY, xmx = -INFINITY, ymx = -INFINITY;
  for (int b = threadIdx.x; b < nblk; b += kBlock) {
    xmn = fmin(xmn, part[4 * b + 0]);
    ymn = fmin(ymn, part[4 * b + 1]);
    xmx = fmax(xmx, part[4 * b + 2]);
    ymx = fmax(ymx, part[4 * b + 3]);
  }
  xmn = warp_min(xmn);
  ymn = warp_min(ymn);
  xmx = warp_max(xmx);
  ymx = warp_max(ymx);
  const int wv = threadIdx.x / 64, ln = threadIdx.x % 64;
  if (ln == 0) {
    red[0][wv] = xmn;
    red[1][wv] = ymn;
    red[2][wv] = xmx;
    red[3][wv] = ymx;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int w = 1; w < kBlock / 64; ++w) {
    xmn = fmin(xmn, red[0][w]);
    ymn = fmin(ymn, red[1][w]);
    xmx = fmax(xmx, red[2][w]);
    ymx = fmax(ymx, red[3][w]);
  }
  const double ext = fmax(xmx - xmn, ymx - ymn);
  BBox b;
  b.x0 = xmn;
  b.y0 = ymn;
  b.inv_extent = ext > 0.0 ? 1.0 / ext : 0.0;
  b.pad = 0.0;
  *out = b;
}

// q = clamp(floor(((x - x0) * inv) * 65536), 0, 65535) with unfused IEEE operations
__device__ __forceinline__ uint32_t quant16(double v, double v0, double inv) {
#pragma clang fp contract(off)
  double q = floor(__dmul_rn(__dmul_rn(__dsub_rn(v, v0), inv), 65536.0));
  q = q < 0.0 ? 0.0 : (q > 65535.0 ? 65535.0 : q);
  return (uint32_t)q;
}

__global__ __launch_bounds__(kBlock) void morton_kernel(int64_t n, const double* __restrict__ x,
                                                        const double* __restrict__ y,
                                                        const BBox* __restrict__ bb,
                                                        double x0, double y0, double inv,
                                                        uint32_t* __restrict__ keys,
                                                        uint32_t* __restrict__ order) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  if (bb) {
    x0 = bb->x0;
    y0 = bb->y0;
    inv = bb->inv_extent;
  }
  keys[i] = spread16(quant16(x[i], x0, inv)) | (spread16(quant16(y[i], y0, inv)) << 1);
  order[i] = (uint32_t)i;
}

// One wave per group of 64 sources (in Morton order): the fp64 centroid (a wave sum in a fixed
// order), the fp32 offsets from it, the largest offset length and radius term, and the staged
// source record (x - cx, y - cy, strength, radius term). Padding past ns is a zero-strength
// record at the centroid with radius term 1; a group with no sources has centroid (0, 0) and
// zero radius and radius term.
template <int CORE>
__global__ __launch_bounds__(kBlock) void group_kernel(int64_t ns, int64_t ngroups,
                                                       const uint32_t* __restrict__ order,
                                                       const double* __restrict__ sx,
                                                       const double* __restrict__ sy,
                                                       const double* __restrict__ ss,
                                                       const double* __restrict__ sr,
                                                       GroupInfo* __restrict__ grp,
                                                       float4* __restrict__ out) {
  static_assert(kGroup == 64, "one wave per group");
  const int64_t g = (int64_t)blockIdx.x * (kBlock / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (g >= ngroups) return;  // the whole wave
  const int64_t j = g * kGroup + lane;
  const bool valid = j < ns;
  const int64_t left = ns - g * kGroup;
  const int cnt = left <= 0 ? 0 : (left < kGroup ? (int)left : kGroup);
  uint32_t o = 0;
  double x = 0.0, y = 0.0;
  if (valid) {
    o = order[j];
    x = sx[o];
    y = sy[o];
  }
  double cx = 0.0, cy = 0.0;
  float rad = 0.0f, wmax = 0.0f;
  float4 rec = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
  if (cnt > 0) {
    const double inv = 1.0 / (double)cnt;
    cx = warp_sum(x) * inv;
    cy = warp_sum(y) * inv;
    if (valid) {
      const float ox = (float)(x - cx), oy = (float)(y - cy);
      const float w = (float)rterm<CORE>(sr[o]);
      rec = make_float4(ox, oy, (float)ss[o], w);
      rad = __builtin_sqrtf(__builtin_fmaf(ox, ox, oy * oy));
      wmax = w;
    }
    rad = warp_maxf(rad);
    wmax = warp_maxf(wmax);
  }
  out[j] = rec;
  if (lane == 0) grp[g] = GroupInfo{cx, cy, rad, wmax, 0.0f, 0.0f};
}

// The fp64 form's staged sources: record i is input source order[i] (i when order is null)
// with its radius term.
template <int CORE>
__global__ __launch_bounds__(kBlock) void stage64_kernel(int64_t n, const uint32_t* __restrict__ order,
                                                         const double* __restrict__ sx,
                                                         const double* __restrict__ sy,
                                                         const double* __restrict__ ss,
                                                         const double* __restrict__ sr,
                                                         Src64* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int64_t o = order ? (int64_t)order[i] : i;
  Src64 q;
  q.x = sx[o];
  q.y = sy[o];
  q.s = ss[o];
  q.w = rterm<CORE>(sr[o]);
  out[i] = q;
}

// Targets in Morton order: position i is input target order[i] (i when order is null); w (when
// tr is given) its radius term, in the type the form's pair arithmetic adds it in.
template <int CORE, class W>
__global__ __launch_bounds__(kBlock) void gather_targets_kernel(int64_t n,
                                                                const uint32_t* __restrict__ order,
                                                                const double* __restrict__ tx,
                                                                const double* __restrict__ ty,
                                                                const double* __restrict__ tr,
                                                                double2* __restrict__ pos,
                                                                W* __restrict__ w) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int64_t o = order ? (int64_t)order[i] : i;
  pos[i] = make_double2(tx[o], ty[o]);
  if (tr) w[i] = (W)rterm<CORE>(tr[o]);
}

// ---------------------------------------------------------------------------- launch helpers

struct ChunkPlan {
  int64_t nchunk;
  int64_t per_chunk;  // sources (plain) or tiles (split) per chunk
};

// Splits the source tiles into chunks (grid.y) so that the launch has per_cu workgroups per CU.
// Beyond 8 per CU a chunk keeps at least kMinChunkTiles tiles, so that few targets do not turn
// into thousands of tiny chunks and partial sums. At most max_chunks chunks.
ChunkPlan plan_chunks(const Device* d, int64_t target_wgs, int64_t ntiles, int per_cu,
                      int64_t max_chunks) {
  const auto chunks_for = [&](int64_t wgs_per_cu) {
    const int64_t want = wgs_per_cu * (int64_t)d->num_cu;
    return target_wgs >= want ? (int64_t)1 : (want + target_wgs - 1) / target_wgs;
  };
  int64_t nchunk = chunks_for(per_cu);
  const int64_t most = std::max(chunks_for(8), (ntiles + kMinChunkTiles - 1) / kMinChunkTiles);
  if (nchunk > most) nchunk = most;
  if (nchunk > max_chunks) nchunk = max_chunks;
  if (nchunk > ntiles) nchunk = ntiles;
  if (nchunk < 1) nchunk = 1;
  const int64_t per = (ntiles + nchunk - 1) / nchunk;
  nchunk = (ntiles + per - 1) / per;
  return {nchunk, per};
}

template <int CORE, int RES, bool TR>
void launch64(dim3 grid, hipStream_t s, int64_t ns, const Src64* src, int64_t nt,
              const double2* tq, const double* tw, const uint32_t* torder, int64_t chunk,
              double* ou, double* ov, double* ow, int direct) {
  nbody64_kernel<CORE, RES, TR, kTpl64><<<grid, kBlock, 0, s>>>(ns, src, nt, tq, tw, torder, chunk,
                                                               ou, ov, ow, direct);
}

template <int CORE>
void dispatch64(int res, bool tr_present, dim3 grid, hipStream_t s, int64_t ns, const Src64* src,
                int64_t nt, const double2* tq, const double* tw, const uint32_t* torder,
                int64_t chunk, double* ou, double* ov, double* ow, int direct) {
#define O2D_L64(R, T) \
  launch64<CORE, R, T>(grid, s, ns, src, nt, tq, tw, torder, chunk, ou, ov, ow, direct)
  if (res == O2D_VELONLY) {
    if (tr_present) O2D_L64(O2D_VELONLY, true);
    else O2D_L64(O2D_VELONLY, false);
  } else if (res == O2D_VELANDVORT) {
    if (tr_present) O2D_L64(O2D_VELANDVORT, true);
    else O2D_L64(O2D_VELANDVORT, false);
  } else {
    if (tr_present) O2D_L64(O2D_VELANDSHEAR, true);
    else O2D_L64(O2D_VELANDSHEAR, false);
  }
#undef O2D_L64
}

template <int CORE, int RES>
void launch32(bool tr_present, dim3 grid, hipStream_t s, const float4* src, const GroupInfo* grp,
              int64_t ntiles, int64_t nt, const double2* tq, const float* tw,
              const uint32_t* torder, int64_t per, double* ou, double* ov, double* ow,
              int direct) {
  constexpr int kUnroll = 16;
  if (tr_present)
    nbody32_kernel<CORE, RES, true, kTpl32, kUnroll><<<grid, kBlock, 0, s>>>(
        src, grp, ntiles, nt, tq, tw, torder, per, ou, ov, ow, direct);
  else
    nbody32_kernel<CORE, RES, false, kTpl32, kUnroll><<<grid, kBlock, 0, s>>>(
        src, grp, ntiles, nt, tq, tw, torder, per, ou, ov, ow, direct);
}

template <int CORE>
void dispatch32(int res, bool tr_present, dim3 grid, hipStream_t s, const float4* src,
                const GroupInfo* grp, int64_t ntiles, int64_t nt, const double2* tq,
                const float* tw, const uint32_t* torder, int64_t per, double* ou, double* ov,
                double* ow, int direct) {
  if (res == O2D_VELONLY)
    launch32<CORE, O2D_VELONLY>(tr_present, grid, s, src, grp, ntiles, nt, tq, tw, torder, per,
                                ou, ov, ow, direct);
  else if (res == O2D_VELANDVORT)
    launch32<CORE, O2D_VELANDVORT>(tr_present, grid, s, src, grp, ntiles, nt, tq, tw, torder,
                                   per, ou, ov, ow, direct);
  else
    launch32<CORE, O2D_VELANDSHEAR>(tr_present, grid, s, src, grp, ntiles, nt, tq, tw, torder,
                                    per, ou, ov, ow, direct);
}

template <int CORE>
void launch_prep(int64_t ngroups, int64_t npad, hipStream_t s, int64_t ns, const uint32_t* order,
                 const double* sx, const double* sy, const double* ss, const double* sr,
                 GroupInfo* grp, float4* out) {
  (void)npad;  // ngroups * kGroup
  constexpr int kGroupsPerBlock = kBlock / 64;
  group_kernel<CORE><<<(unsigned)((ngroups + kGroupsPerBlock - 1) / kGroupsPerBlock), kBlock, 0,
                       s>>>(ns, ngroups, order, sx, sy, ss, sr, grp, out);
}

inline dim3 grid1(int64_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

template <int CORE>
void launch_stage64(hipStream_t s, int64_t n, const uint32_t* order, const double* sx,
                    const double* sy, const double* ss, const double* sr, Src64* out) {
  stage64_kernel<CORE><<<grid1(n), kBlock, 0, s>>>(n, order, sx, sy, ss, sr, out);
}

template <int CORE, class W>
void launch_gather(hipStream_t s, int64_t n, const uint32_t* order, const double* tx,
                   const double* ty, const double* tr, double2* pos, W* w) {
  gather_targets_kernel<CORE, W><<<grid1(n), kBlock, 0, s>>>(n, order, tx, ty, tr, pos, w);
}

// targets in the given order, radius terms of type W (when tr is given)
template <class W>
int gather_targets(hipStream_t s, int core, int64_t n, const uint32_t* order, const double* tx,
                   const double* ty, const double* tr, double2* pos, W* w) {
  switch (core) {
    case O2D_CORE_RM: launch_gather<O2D_CORE_RM, W>(s, n, order, tx, ty, tr, pos, w); break;
    case O2D_CORE_GAUSSIAN: launch_gather<O2D_CORE_GAUSSIAN, W>(s, n, order, tx, ty, tr, pos, w); break;
    case O2D_CORE_WL: launch_gather<O2D_CORE_WL, W>(s, n, order, tx, ty, tr, pos, w); break;
    case O2D_CORE_V2: launch_gather<O2D_CORE_V2, W>(s, n, order, tx, ty, tr, pos, w); break;
    default: launch_gather<O2D_CORE_V3, W>(s, n, order, tx, ty, tr, pos, w); break;
  }
  O2D_LAUNCH_CHECK();
  return 0;
}

}  // namespace

// Morton keys and the stable key order; bbox == true derives the box from the data on device.
int morton_order(Device* d, hipStream_t s, int64_t n, const double* x, const double* y,
                 bool bbox, double x0, double y0, double inv, uint32_t* keys_out,
                 uint32_t* order_out, Workspace* ws) {
  if (n <= 0) return 0;
  void *kA, *oA, *bbp = nullptr;
  O2D_TRY(take(d, ws, kKeysA, n * sizeof(uint32_t), s, &kA));
  O2D_TRY(take(d, ws, kOrderA, n * sizeof(uint32_t), s, &oA));
  BBox* bb = nullptr;
  if (bbox) {
    const int nblk = (int)std::min<int64_t>((n + kBlock - 1) / kBlock, 1024);
    O2D_TRY(take(d, ws, kBBox, sizeof(double) * (4 * nblk + 4), s, &bbp));
    double* part = static_cast<double*>(bbp);
    bb = reinterpret_cast<BBox*>(part + 4 * nblk);
    bbox_partial_kernel<<<nblk, kBlock, 0, s>>>(n, x, y, part);
    bbox_final_kernel<<<1, kBlock, 0, s>>>(nblk, part, bb);
    O2D_LAUNCH_CHECK();
  }
  morton_kernel<<<grid1(n), kBlock, 0, s>>>(n, x, y, bb, x0, y0, inv, static_cast<uint32_t*>(kA),
                                            static_cast<uint32_t*>(oA));
  O2D_LAUNCH_CHECK();
  size_t tbytes = 0;
  O2D_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, static_cast<uint32_t*>(kA),
                                             keys_out, static_cast<uint32_t*>(oA), order_out,
                                             (int)n, 0, 32, s));
  void* tmp;
  O2D_TRY(take(d, ws, kSortTemp, tbytes, s, &tmp));
  O2D_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tbytes, static_cast<uint32_t*>(kA), keys_out,
                                             static_cast<uint32_t*>(oA), order_out, (int)n, 0, 32,
                                             s));
  return 0;
}

// The most source chunks whose fp64 partials (nf per target) fit what is left of *ws; no limit
// when the temporaries come from the buffer pool.
int64_t max_chunks(const Workspace* ws, int nf, int64_t nt) {
  if (!ws) return INT64_MAX;
  const int64_t c = (int64_t)(ws->left() / (sizeof(double) * (size_t)nf * (size_t)nt));
  return c < 1 ? 1 : c;
}

// Targets for the kernels: their positions and radius terms gathered into *pos and *w, and
// *order mapping each gathered target to its input index. Targets that are the sources
// themselves (the same position arrays) reuse the sources' order src_order; other targets are
// put in Morton order when `sort`, else kept in input order (*order null).
template <class W>
int prep_targets(Device* d, hipStream_t s, Workspace* ws, int core, int64_t nt, const double* tx,
                 const double* ty, const double* tr, bool sort, const uint32_t* src_order,
                 const uint32_t** order, const double2** pos, const W** w) {
  const uint32_t* o = src_order;
  if (!o && sort) {
    void *kT, *oT;
    O2D_TRY(take(d, ws, kTrgKeys, nt * sizeof(uint32_t), s, &kT));
    O2D_TRY(take(d, ws, kTrgOrder, nt * sizeof(uint32_t), s, &oT));
    O2D_TRY(morton_order(d, s, nt, tx, ty, true, 0.0, 0.0, 0.0, static_cast<uint32_t*>(kT),
                         static_cast<uint32_t*>(oT), ws));
    o = static_cast<const uint32_t*>(oT);
  }
  void *pp, *wp = nullptr;
  O2D_TRY(take(d, ws, kTrgPos, nt * sizeof(double2), s, &pp));
  if (tr) O2D_TRY(take(d, ws, kTrgW, nt * sizeof(W), s, &wp));
  O2D_TRY(gather_targets<W>(s, core, nt, o, tx, ty, tr, static_cast<double2*>(pp),
                            static_cast<W*>(wp)));
  *order = o;
  *pos = static_cast<const double2*>(pp);
  *w = static_cast<const W*>(wp);
  return 0;
}

int nbody_dev(Device* d, hipStream_t s, int core, int prec, int res, int64_t ns, const double* sx,
              const double* sy, const double* ss, const double* sr, int64_t nt, const double* tx,
              const double* ty, const double* tr, double* tu, double* tv, double* tw,
              Workspace* ws) {
  if (nt == 0 || ns == 0) return 0;
  // the radix sort of the Morton keys takes an int count
  if (ns > 0x7fffffffLL || nt > 0x7fffffffLL)
    return fail(O2D_ERR_SIZE, "more than 2^31 - 1 sources or targets");
  const bool trp = tr != nullptr;
  // targets that are the sources themselves (a collection acting on itself) share their order
  const bool self = tx == sx && ty == sy && nt == ns;
  // sources in Morton order (both forms): the stable order of 32-bit Morton keys
  void *kB, *oB;
  O2D_TRY(take(d, ws, kKeysB, ns * sizeof(uint32_t), s, &kB));
  O2D_TRY(take(d, ws, kOrderB, ns * sizeof(uint32_t), s, &oB));

  if (prec == O2D_PLAIN_FP64) {
    // only the Gaussian's velocity sum uses locality (its far-field tiles); every other fp64
    // sum keeps the input order and skips the sort
    const bool sort = core == O2D_CORE_GAUSSIAN && res == O2D_VELONLY;
    void* sp;
    O2D_TRY(take(d, ws, kSrc64, ns * sizeof(Src64), s, &sp));
    const uint32_t* order = nullptr;
    if (sort) {
      O2D_TRY(morton_order(d, s, ns, sx, sy, true, 0.0, 0.0, 0.0, static_cast<uint32_t*>(kB),
                           static_cast<uint32_t*>(oB), ws));
      order = static_cast<const uint32_t*>(oB);
    }
    Src64* src = static_cast<Src64*>(sp);
    switch (core) {
      case O2D_CORE_RM: launch_stage64<O2D_CORE_RM>(s, ns, order, sx, sy, ss, sr, src); break;
      case O2D_CORE_GAUSSIAN: launch_stage64<O2D_CORE_GAUSSIAN>(s, ns, order, sx, sy, ss, sr, src); break;
      case O2D_CORE_WL: launch_stage64<O2D_CORE_WL>(s, ns, order, sx, sy, ss, sr, src); break;
      case O2D_CORE_V2: launch_stage64<O2D_CORE_V2>(s, ns, order, sx, sy, ss, sr, src); break;
      default: launch_stage64<O2D_CORE_V3>(s, ns, order, sx, sy, ss, sr, src); break;
    }
    O2D_LAUNCH_CHECK();
    const uint32_t* torder;
    const double2* tq;
    const double* twv;
    O2D_TRY(prep_targets<double>(d, s, ws, core, nt, tx, ty, tr, sort, self ? order : nullptr,
                                 &torder, &tq, &twv));

    const int64_t twgs = (nt + kBlock * kTpl64 - 1) / (kBlock * kTpl64);
    const int64_t ntiles = (ns + kTile - 1) / kTile;
    const int nf = res == O2D_VELONLY ? 2 : 3;
    ChunkPlan cp = plan_chunks(d, twgs, ntiles, O2D_WANT64, max_chunks(ws, nf, nt));
    const int64_t chunk = cp.per_chunk * kTile;
    const dim3 grid((unsigned)twgs, (unsigned)cp.nchunk);
    double *ou = tu, *ov = tv, *ow = tw;
    const int direct = cp.nchunk == 1;
    if (!direct) {
      void* p;
      O2D_TRY(take(d, ws, kPartials, sizeof(double) * nf * cp.nchunk * nt, s, &p));
      ou = static_cast<double*>(p);
      ov = ou + cp.nchunk * nt;
      ow = res == O2D_VELONLY ? nullptr : ov + cp.nchunk * nt;
    }
    KernelTimer kt(trp ? O2D_KT_NBODY_PLAIN : O2D_KT_NBODY_PLAIN_PT, s);
    switch (core) {
      case O2D_CORE_RM: dispatch64<O2D_CORE_RM>(res, trp, grid, s, ns, src, nt, tq, twv, torder, chunk, ou, ov, ow, direct); break;
      case O2D_CORE_GAUSSIAN: dispatch64<O2D_CORE_GAUSSIAN>(res, trp, grid, s, ns, src, nt, tq, twv, torder, chunk, ou, ov, ow, direct); break;
      case O2D_CORE_WL: dispatch64<O2D_CORE_WL>(res, trp, grid, s, ns, src, nt, tq, twv, torder, chunk, ou, ov, ow, direct); break;
      case O2D_CORE_V2: dispatch64<O2D_CORE_V2>(res, trp, grid, s, ns, src, nt, tq, twv, torder, chunk, ou, ov, ow, direct); break;
      default: dispatch64<O2D_CORE_V3>(res, trp, grid, s, ns, src, nt, tq, twv, torder, chunk, ou, ov, ow, direct); break;
    }
    kt.stop();
    O2D_LAUNCH_CHECK();
    if (!direct) {
      reduce_chunks_kernel<<<grid1(nt), kBlock, 0, s>>>(cp.nchunk, nt, torder, ou, ov, ow, tu, tv,
                                                        tw);
      O2D_LAUNCH_CHECK();
    }
    return 0;
  }

  // split form
  const int64_t ntiles = (ns + kTile - 1) / kTile;
  const int64_t npad = ntiles * kTile;
  const int64_t ngroups = npad / kGroup;
  void *tsrc, *tcent;
  O2D_TRY(take(d, ws, kTileSrc, npad * sizeof(float4), s, &tsrc));
  O2D_TRY(take(d, ws, kTileCent, ngroups * sizeof(GroupInfo), s, &tcent));
  KernelTimer kprep(O2D_KT_SPLIT_PREP, s);
  O2D_TRY(morton_order(d, s, ns, sx, sy, true, 0.0, 0.0, 0.0, static_cast<uint32_t*>(kB),
                       static_cast<uint32_t*>(oB), ws));
  const uint32_t* order = static_cast<const uint32_t*>(oB);
  GroupInfo* cent = static_cast<GroupInfo*>(tcent);
  float4* src = static_cast<float4*>(tsrc);
  switch (core) {
    case O2D_CORE_RM: launch_prep<O2D_CORE_RM>(ngroups, npad, s, ns, order, sx, sy, ss, sr, cent, src); break;
    case O2D_CORE_GAUSSIAN: launch_prep<O2D_CORE_GAUSSIAN>(ngroups, npad, s, ns, order, sx, sy, ss, sr, cent, src); break;
    case O2D_CORE_WL: launch_prep<O2D_CORE_WL>(ngroups, npad, s, ns, order, sx, sy, ss, sr, cent, src); break;
    case O2D_CORE_V2: launch_prep<O2D_CORE_V2>(ngroups, npad, s, ns, order, sx, sy, ss, sr, cent, src); break;
    default: launch_prep<O2D_CORE_V3>(ngroups, npad, s, ns, order, sx, sy, ss, sr, cent, src); break;
  }
  O2D_LAUNCH_CHECK();
  const uint32_t* torder;
  const double2* tq;
  const float* twv;
  // the target order matters to the Gaussian's wave-uniform far-field test only
  O2D_TRY(prep_targets<float>(d, s, ws, core, nt, tx, ty, tr, core == O2D_CORE_GAUSSIAN,
                              self ? order : nullptr, &torder, &tq, &twv));
  kprep.stop();

  const int64_t twgs = (nt + 64 * kTpl32 - 1) / (64 * kTpl32);  // one target set per workgroup
  const int nf = res == O2D_VELONLY ? 2 : 3;
  ChunkPlan cp = plan_chunks(d, twgs, ntiles, O2D_WANT32, max_chunks(ws, nf, nt));
  const dim3 grid((unsigned)twgs, (unsigned)cp.nchunk);
  double *ou = tu, *ov = tv, *ow = tw;
  const int direct = cp.nchunk == 1;
  if (!direct) {
    void* p;
    O2D_TRY(take(d, ws, kPartials, sizeof(double) * nf * cp.nchunk * nt, s, &p));
    ou = static_cast<double*>(p);
    ov = ou + cp.nchunk * nt;
    ow = res == O2D_VELONLY ? nullptr : ov + cp.nchunk * nt;
  }
  KernelTimer kt(trp ? O2D_KT_NBODY_SPLIT : O2D_KT_NBODY_SPLIT_PT, s);
  switch (core) {
    case O2D_CORE_RM: dispatch32<O2D_CORE_RM>(res, trp, grid, s, src, cent, ntiles, nt, tq, twv, torder, cp.per_chunk, ou, ov, ow, direct); break;
    case O2D_CORE_GAUSSIAN: dispatch32<O2D_CORE_GAUSSIAN>(res, trp, grid, s, src, cent, ntiles, nt, tq, twv, torder, cp.per_chunk, ou, ov, ow, direct); break;
    case O2D_CORE_WL: dispatch32<O2D_CORE_WL>(res, trp, grid, s, src, cent, ntiles, nt, tq, twv, torder, cp.per_chunk, ou, ov, ow, direct); break;
    case O2D_CORE_V2: dispatch32<O2D_CORE_V2>(res, trp, grid, s, src, cent, ntiles, nt, tq, twv, torder, cp.per_chunk, ou, ov, ow, direct); break;
    default: dispatch32<O2D_CORE_V3>(res, trp, grid, s, src, cent, ntiles, nt, tq, twv, torder, cp.per_chunk, ou, ov, ow, direct); break;
  }
  kt.stop();
  O2D_LAUNCH_CHECK();
  if (!direct) {
    reduce_chunks_kernel<<<grid1(nt), kBlock, 0, s>>>(cp.nchunk, nt, torder, ou, ov, ow, tu, tv,
                                                      tw);
    O2D_LAUNCH_CHECK();
  }
  return 0;
}

}  // namespace o2d
