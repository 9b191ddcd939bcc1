// panels.hip - constant-strength straight panels acting on points, and point vortices acting on
// panels (the BEM right-hand side). fp64 throughout. From 128 panels both sums take far
// clusters of 32 panels by an expansion about the cluster's centre (below).
//
// panels_on_points (src/Influence.h:422-614, kernels src/Kernels.h:314-405): one lane per
// target point, 256-panel tiles staged in LDS, fp64 sums.
//
// points_on_panels (src/Influence.h:633-771): targets are the few panels, sources the many
// particles. A workgroup owns kPanTile panels (held in scalar registers) and a chunk of
// particles; each lane walks its share of particles and accumulates every panel's integral,
// then the 256 lanes are reduced with wavefront shuffles and one LDS step. Chunk partials are
// summed in a fixed order by a second kernel, which applies the -1/L factor.
//
// The subtended angle u* = atan2(dx1,dy1) - atan2(dx0,dy0) wrapped to [-pi, pi]
// (src/Kernels.h:262-268) is evaluated as the single atan2 of the cross and dot products of the
// two endpoint vectors, which is the same angle already wrapped (one atan2 instead of two).
// Targets far from a panel use the convergent series of the same function (panel_math.h).
#include <algorithm>
#include <cfloat>

#include "nbody.h"
#include "o2d_math.h"
#include "o2d_runtime.h"
#include "panel_math.h"

// workgroups per CU the particle chunks of points_on_panels aim for (32: 0.990 -> 0.975 ms on
// 200k particles x 2,048 panels; 128 took 1.073 ms)
#ifndef O2D_WANT_RHS
#define O2D_WANT_RHS 32
#endif

namespace o2d {

namespace {

constexpr int kBlock = 256;
constexpr int kPTile = 256;  // panels per LDS stage (panels on points)
#ifndef O2D_PANTILE
#define O2D_PANTILE 4
#endif
constexpr int kPanTile = O2D_PANTILE;  // panels per workgroup (points on panels)


// Both panel kernels are built for five waves per SIMD (96 VGPRs; panels on points spills two
// values around its tile staging, points on panels reloads eight per particle from scratch):
// C3 0.96 -> 0.94 ms (panels on points), 0.97 -> 0.94 ms (points on panels); six waves per
// SIMD spilled more and were no faster.
template <bool SRC>
__global__ __launch_bounds__(kBlock, 5) void panels_on_points_kernel(
    int64_t np, const double* __restrict__ nx, const double* __restrict__ ny,
    const uint32_t* __restrict__ idx, const double* __restrict__ vs, const double* __restrict__ ss,
    int64_t nt, const double* __restrict__ tx, const double* __restrict__ ty, int64_t chunk,
    double* __restrict__ ou, double* __restrict__ ov, int direct) {
  __shared__ PanelPre tile[kPTile];
  const int tid = threadIdx.x;
  const int64_t i = (int64_t)blockIdx.x * kBlock + tid;
  const int64_t ii = i < nt ? i : nt - 1;
  const double px = tx[ii], py = ty[ii];
  double au = 0.0, av = 0.0;
  const int64_t jb = (int64_t)blockIdx.y * chunk;
  const int64_t je = jb + chunk < np ? jb + chunk : np;
  for (int64_t j0 = jb; j0 < je; j0 += kPTile) {
    const int64_t j = j0 + tid;
    PanelPre p{};
    if (j < je) {
      const uint32_t a = idx[2 * j], b = idx[2 * j + 1];
      p = panel_pre(nx[a], ny[a], nx[b], ny[b], vs[j], SRC ? ss[j] : 0.0);
    }
    __syncthreads();
    tile[tid] = p;
    __syncthreads();
    const int nk = (int)(je - j0 < kPTile ? je - j0 : kPTile);
    for (int k = 0; k < nk; ++k) {
      const PanelPre& q = tile[k];
      double lr, li;
      panel_log(q, px, py, lr, li);
      // v + i u = W L
      au += __builtin_fma(q.wr, li, q.wi * lr);
      av += __builtin_fma(q.wr, lr, -q.wi * li);
    }
  }
  if (i >= nt) return;
  if (direct) {
    ou[i] += au;
    ov[i] += av;
  } else {
    ou[(int64_t)blockIdx.y * nt + i] = au;
    ov[(int64_t)blockIdx.y * nt + i] = av;
  }
}

// ---- clustered far field of panels on points. The panels are cut into clusters of kClus
// consecutive panels (contiguous along a body). About a cluster's centre c, with W_k the
// complex factor of panel k (panel_pre: v + i u = W L) and a_k = z0_k - c, b_k = z1_k - c,
//   sum_k W_k L_k(z) = sum_{n>=1} A_n (z - c)^-n,   A_n = sum_k W_k (b_k^n - a_k^n) / n,
// for |z - c| > R = max |end node - c|. b^n - a^n = (z1 - z0) S_n with S_1 = 1 and
// S_{n+1} = b S_n + a^n (S_n = sum_{j<n} b^j a^(n-1-j)), free of the cancellation of two close
// powers. |S_n| <= n R^(n-1), so the terms past kTerms sum to at most
// (R / |z - c|)^kTerms / (1 - R / |z - c|) of sum_k |W_k| |z1_k - z0_k| / |z - c|, the scale of
// the first term: below 1e-17 of it at |z - c| >= kFarRatio R, where a target takes the
// expansion (the per-panel series stops at the same relative size). Nearer targets sum the
// cluster's panels one by one as the direct kernel does.
constexpr int kClus = 32;
constexpr int kTerms = 22;
constexpr double kFarRatio = 6.0;
constexpr int kHead = 4;  // per cluster: centre x, y, (kFarRatio R)^2, unused

// head[kHead c ...] and mom[2 kTerms c + 2 (n - 1) + {0, 1}] = A_n of cluster c: one workgroup
// of 64 threads per cluster, a thread per panel, the moments summed over panels in order (with
// mom null, the heads alone)
__global__ __launch_bounds__(64) void cluster_kernel(int64_t np, const double* __restrict__ nx,
                                                     const double* __restrict__ ny,
                                                     const uint32_t* __restrict__ idx,
                                                     const double* __restrict__ vs,
                                                     const double* __restrict__ ss,
                                                     double* __restrict__ head,
                                                     double* __restrict__ mom) {
  __shared__ double ends[kClus][4];
  __shared__ double part[kClus][kTerms][2];
  __shared__ double cen[2];
  const int t = threadIdx.x;
  const int64_t c = blockIdx.x, k0 = c * kClus;
  const int nk = (int)(np - k0 < kClus ? np - k0 : kClus);
  PanelPre q{};
  if (t < nk) {
    const uint32_t a = idx[2 * (k0 + t)], b = idx[2 * (k0 + t) + 1];
    q = panel_pre(nx[a], ny[a], nx[b], ny[b], vs ? vs[k0 + t] : 1.0, ss ? ss[k0 + t] : 0.0);
    ends[t][0] = q.x0;
    ends[t][1] = q.y0;
    ends[t][2] = q.x1;
    ends[t][3] = q.y1;
  }
  __syncthreads();
  if (t == 0) {  // the centre of the end nodes' box, and R^2
    double x0 = ends[0][0], x1 = x0, y0 = ends[0][1], y1 = y0;
    for (int k = 0; k < nk; ++k) {
      x0 = fmin(x0, fmin(ends[k][0], ends[k][2]));
      x1 = fmax(x1, fmax(ends[k][0], ends[k][2]));
      y0 = fmin(y0, fmin(ends[k][1], ends[k][3]));
      y1 = fmax(y1, fmax(ends[k][1], ends[k][3]));
    }
    const double cx = 0.5 * (x0 + x1), cy = 0.5 * (y0 + y1);
    double r2 = 0.0;
    for (int k = 0; k < nk; ++k) {
      for (int e = 0; e < 2; ++e) {
        const double dx = ends[k][2 * e] - cx, dy = ends[k][2 * e + 1] - cy;
        r2 = fmax(r2, dx * dx + dy * dy);
      }
    }
    cen[0] = cx;
    cen[1] = cy;
    head[kHead * c] = cx;
    head[kHead * c + 1] = cy;
    // (kFarRatio R)^2, rounded up by far more than its rounding error
    head[kHead * c + 2] = kFarRatio * kFarRatio * r2 * (1.0 + 0x1p-30);
    head[kHead * c + 3] = 0.0;
  }
  __syncthreads();
  if (!mom) return;  // the heads alone (points on panels)
  if (t < kClus) {
    if (t < nk) {
      const double ax = q.x0 - cen[0], ay = q.y0 - cen[1];
      const double bx = q.x1 - cen[0], by = q.y1 - cen[1];
      // W (z1 - z0), then S_n and a^n
      const double dzx = q.x1 - q.x0, dzy = q.y1 - q.y0;
      const double fx = q.wr * dzx - q.wi * dzy, fy = q.wr * dzy + q.wi * dzx;
      double sx = 1.0, sy = 0.0, px = ax, py = ay;  // S_1, a^1
      for (int n = 1; n <= kTerms; ++n) {
        const double inv = 1.0 / (double)n;
        part[t][n - 1][0] = (fx * sx - fy * sy) * inv;
        part[t][n - 1][1] = (fx * sy + fy * sx) * inv;
        const double nsx = bx * sx - by * sy + px, nsy = bx * sy + by * sx + py;  // b S + a^n
        sx = nsx;
        sy = nsy;
        const double npx = px * ax - py * ay;  // a^(n+1)
        py = px * ay + py * ax;
        px = npx;
      }
    } else {
      for (int n = 0; n < kTerms; ++n) part[t][n][0] = part[t][n][1] = 0.0;
    }
  }
  __syncthreads();
  if (t < kTerms) {
    double ar = 0.0, ai = 0.0;
    for (int k = 0; k < kClus; ++k) {
      ar += part[k][t][0];
      ai += part[k][t][1];
    }
    mom[2 * (kTerms * c + t)] = ar;
    mom[2 * (kTerms * c + t) + 1] = ai;
  }
}

// panels on points with the clustered far field: one lane per target (targets taken in the
// order torder, so that a wave is compact and its lanes agree on near and far), the clusters
// [blockIdx.y cchunk, (blockIdx.y + 1) cchunk); with several chunks, chunk y's sums go to
// pu/pv[y nt + position in torder] for reduce_sorted_kernel, else they are added to ou/ov
template <bool SRC>
__global__ __launch_bounds__(kBlock, 5) void panels_clustered_kernel(
    int64_t np, const double* __restrict__ nx, const double* __restrict__ ny,
    const uint32_t* __restrict__ idx, const double* __restrict__ vs, const double* __restrict__ ss,
    int64_t nclus, const double* __restrict__ head, const double* __restrict__ mom, int64_t nt,
    const double* __restrict__ tx, const double* __restrict__ ty,
    const uint32_t* __restrict__ torder, int64_t cchunk, double* __restrict__ ou,
    double* __restrict__ ov, int direct) {
  const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t gg = g < nt ? g : nt - 1;
  const int64_t i = torder ? (int64_t)torder[gg] : gg;
  const double px = tx[i], py = ty[i];
  double au = 0.0, av = 0.0;
  const int64_t cb = (int64_t)blockIdx.y * cchunk;
  const int64_t ce = cb + cchunk < nclus ? cb + cchunk : nclus;
  for (int64_t c = cb; c < ce; ++c) {
    const double* h = head + kHead * c;
    const double zx = px - h[0], zy = py - h[1];
    const double z2 = __builtin_fma(zx, zx, zy * zy);
    if (z2 >= h[2] && z2 > 0.0) {  // far (a point at the centre of a cluster of no extent is near)
      // w = 1 / (z - c); sum_n A_n w^n = w (A_1 + w (A_2 + ... + w A_kTerms))
      const double iz = rcp64(z2);
      const double wx = zx * iz, wy = -zy * iz;
      const double* A = mom + 2 * kTerms * c;
      double sx = A[2 * (kTerms - 1)], sy = A[2 * (kTerms - 1) + 1];
#pragma unroll
      for (int n = kTerms - 2; n >= 0; --n) {
        const double nx_ = __builtin_fma(sx, wx, __builtin_fma(-sy, wy, A[2 * n]));
        sy = __builtin_fma(sx, wy, __builtin_fma(sy, wx, A[2 * n + 1]));
        sx = nx_;
      }
      av += __builtin_fma(sx, wx, -sy * wy);  // v + i u = W L
      au += __builtin_fma(sx, wy, sy * wx);
    } else {
      const int64_t k1 = (c + 1) * kClus < np ? (c + 1) * kClus : np;
      for (int64_t k = c * kClus; k < k1; ++k) {
        const uint32_t a = idx[2 * k], b = idx[2 * k + 1];
        const PanelPre q = panel_pre(nx[a], ny[a], nx[b], ny[b], vs[k], SRC ? ss[k] : 0.0);
        double lr, li;
        panel_log(q, px, py, lr, li);
        au += __builtin_fma(q.wr, li, q.wi * lr);
        av += __builtin_fma(q.wr, lr, -q.wi * li);
      }
    }
  }
  if (g >= nt) return;
  if (direct) {
    ou[i] += au;
    ov[i] += av;
  } else {
    ou[(int64_t)blockIdx.y * nt + g] = au;
    ov[(int64_t)blockIdx.y * nt + g] = av;
  }
}

// ou[torder[g]] += sum over chunks of pu[c nt + g], in chunk order (the same for ov)
__global__ __launch_bounds__(kBlock) void reduce_sorted_kernel(int64_t nchunk, int64_t nt,
                                                               const uint32_t* __restrict__ torder,
                                                               const double* __restrict__ pu,
                                                               const double* __restrict__ pv,
                                                               double* __restrict__ ou,
                                                               double* __restrict__ ov) {
  const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (g >= nt) return;
  double su = 0.0, sv = 0.0;
  for (int64_t c = 0; c < nchunk; ++c) {
    su += pu[c * nt + g];
    sv += pv[c * nt + g];
  }
  const int64_t i = torder ? (int64_t)torder[g] : g;
  ou[i] += su;
  ov[i] += sv;
}

__global__ __launch_bounds__(kBlock) void reduce_chunks2_kernel(int64_t nchunk, int64_t n,
                                                                const double* __restrict__ pu,
                                                                const double* __restrict__ pv,
                                                                double* __restrict__ ou,
                                                                double* __restrict__ ov) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double su = 0.0, sv = 0.0;
  for (int64_t c = 0; c < nchunk; ++c) {
    su += pu[c * n + i];
    sv += pv[c * n + i];
  }
  ou[i] += su;
  ov[i] += sv;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// part[(c*2 + {0,1}) * np + p] = chunk c's (u, v) integral sum for panel p
__global__ __launch_bounds__(kBlock, 5) void points_on_panels_kernel(
    int64_t ns, const double* __restrict__ sx, const double* __restrict__ sy,
    const double* __restrict__ ss, int64_t np, const double* __restrict__ nx,
    const double* __restrict__ ny, const uint32_t* __restrict__ idx, int64_t chunk,
    double* __restrict__ part, const double* __restrict__ head,
    const double* __restrict__ gbox) {
  // with head, only the particles nearer than kFarRatio radii to the centre of the cluster of
  // this workgroup's panels (rhs_far_kernel takes the others). The particles then come in
  // Morton order with the box of every 64 (gbox), chunks are whole groups, and a wave skips a
  // group whose box lies beyond those radii (by a margin far above the rounding of both tests,
  // so that every particle it skips is one the far kernel takes)
  __shared__ double red[kBlock / 64][2 * kPanTile];
  __shared__ PanelPre pp[kPanTile];
  const int tid = threadIdx.x;
  const int64_t p0 = (int64_t)blockIdx.x * kPanTile;
  if (tid < kPanTile) {
    const int64_t p = p0 + tid < np ? p0 + tid : np - 1;
    const uint32_t a = idx[2 * p], b = idx[2 * p + 1];
    pp[tid] = panel_pre(nx[a], ny[a], nx[b], ny[b], 1.0, 0.0);  // W = conj(e)
  }
  __syncthreads();
  // sum over particles of s_j L_j per panel; conj(e) is applied once at the end
  double ar[kPanTile], ai[kPanTile];
#pragma unroll
  for (int k = 0; k < kPanTile; ++k) ar[k] = ai[k] = 0.0;
  const int64_t jb = (int64_t)blockIdx.y * chunk;
  const int64_t je = jb + chunk < ns ? jb + chunk : ns;
  static_assert(kClus % kPanTile == 0, "a panel tile lies in one cluster");
  double hx = 0.0, hy = 0.0, hr2 = 0.0;
  if (head) {
    const double* h = head + kHead * (p0 / kClus);
    hx = h[0];
    hy = h[1];
    hr2 = h[2];
  }
  // with head, the whole chunk is skipped when the box of its groups lies beyond those radii
  bool skip = false;
  if (head) {
    __shared__ double cb[4][kBlock];
    double x0 = DBL_MAX, x1 = -DBL_MAX, y0 = DBL_MAX, y1 = -DBL_MAX;
    for (int64_t q = jb / 64 + tid; q < (je + 63) / 64; q += kBlock) {
      const double* b = gbox + 4 * q;
      x0 = fmin(x0, b[0]);
      x1 = fmax(x1, b[1]);
      y0 = fmin(y0, b[2]);
      y1 = fmax(y1, b[3]);
    }
    cb[0][tid] = x0;
    cb[1][tid] = x1;
    cb[2][tid] = y0;
    cb[3][tid] = y1;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
      if (tid < w) {
        cb[0][tid] = fmin(cb[0][tid], cb[0][tid + w]);
        cb[1][tid] = fmax(cb[1][tid], cb[1][tid + w]);
        cb[2][tid] = fmin(cb[2][tid], cb[2][tid + w]);
        cb[3][tid] = fmax(cb[3][tid], cb[3][tid + w]);
      }
      __syncthreads();
    }
    const double bx = fmax(0.0, fmax(cb[0][0] - hx, hx - cb[1][0]));
    const double by = fmax(0.0, fmax(cb[2][0] - hy, hy - cb[3][0]));
    skip = (bx * bx + by * by) * (1.0 - 0x1p-40) >= hr2;
  }
  for (int64_t j = jb + tid; j < je && !skip; j += kBlock) {
    if (head) {
      const double* b = gbox + 4 * (j / 64);  // the same for the whole wave
      const double bx = fmax(0.0, fmax(b[0] - hx, hx - b[1]));
      const double by = fmax(0.0, fmax(b[2] - hy, hy - b[3]));
      if ((bx * bx + by * by) * (1.0 - 0x1p-40) >= hr2) continue;
    }
    const double qx = sx[j], qy = sy[j], qs = ss[j];
    if (head) {
      const double dx = qx - hx, dy = qy - hy;
      const double z2 = __builtin_fma(dx, dx, dy * dy);
      if (z2 >= hr2 && z2 > 0.0) continue;  // far: rhs_far_kernel's test
    }
#pragma unroll
    for (int k = 0; k < kPanTile; ++k) {
      // read the panel from LDS at each use rather than keeping all eight in registers
      asm volatile("" ::: "memory");
      const PanelPre q = pp[k];
      double lr, li;
      panel_log(q, qx, qy, lr, li);
      ar[k] = __builtin_fma(qs, lr, ar[k]);
      ai[k] = __builtin_fma(qs, li, ai[k]);
    }
  }
  const int wv = tid / 64, ln = tid % 64;
#pragma unroll
  for (int k = 0; k < kPanTile; ++k) {
    // v + i u = conj(e) sum_j s_j L_j
    const double su = wave_sum(__builtin_fma(pp[k].wr, ai[k], pp[k].wi * ar[k]));
    const double sv = wave_sum(__builtin_fma(pp[k].wr, ar[k], -pp[k].wi * ai[k]));
    if (ln == 0) {
      red[wv][2 * k] = su;
      red[wv][2 * k + 1] = sv;
    }
  }
  __syncthreads();
  if (tid < 2 * kPanTile) {
    const int k = tid / 2, c = tid % 2;
    if (p0 + k < np) {
      double s = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) s += red[w][tid];
      part[((int64_t)blockIdx.y * 2 + c) * np + p0 + k] = s;
    }
  }
}

// the particles in the order sorder (xs, ys, qs) and the box of every 64 of them, gbox[4 g ...] =
// (x min, x max, y min, y max); one thread per sorted position, a group per wave
__global__ __launch_bounds__(kBlock) void gather_groups_kernel(int64_t ns,
                                                               const double* __restrict__ sx,
                                                               const double* __restrict__ sy,
                                                               const double* __restrict__ ss,
                                                               const uint32_t* __restrict__ sorder,
                                                               double* __restrict__ xs,
                                                               double* __restrict__ ys,
                                                               double* __restrict__ qs,
                                                               double* __restrict__ gbox) {
  const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double x0 = DBL_MAX, x1 = -DBL_MAX, y0 = DBL_MAX, y1 = -DBL_MAX;
  if (g < ns) {
    const int64_t j = sorder ? (int64_t)sorder[g] : g;
    const double x = sx[j], y = sy[j];
    xs[g] = x;
    ys[g] = y;
    qs[g] = ss[j];
    x0 = x1 = x;
    y0 = y1 = y;
    if (!(x == x) || !(y == y)) {  // a NaN position: the group is never skipped, so the NaN
      x0 = y0 = -INFINITY;           // reaches the sums as in the direct kernel
      x1 = y1 = INFINITY;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    x0 = fmin(x0, __shfl_xor(x0, o, 64));
    x1 = fmax(x1, __shfl_xor(x1, o, 64));
    y0 = fmin(y0, __shfl_xor(y0, o, 64));
    y1 = fmax(y1, __shfl_xor(y1, o, 64));
  }
  if (g % 64 == 0 && g < ns) {
    gbox[4 * (g / 64)] = x0;
    gbox[4 * (g / 64) + 1] = x1;
    gbox[4 * (g / 64) + 2] = y0;
    gbox[4 * (g / 64) + 3] = y1;
  }
}

// The far particles of each cluster (points on panels): per cluster c and particle chunk y,
// G_n = sum_j s_j (z_j - c)^-n over the particles at kFarRatio radii or more, n = 1..kTerms,
// to gpart[((c nchunk + y) kTerms + n - 1) 2 + {0, 1}]. A panel's sum over those particles is
// then sum_j s_j L_p(z_j) = sum_n G_n (b_p^n - a_p^n) / n (rhs_combine_kernel), the expansion of
// the panels-on-points moments read the other way, with the same remainder bound.
__global__ __launch_bounds__(kBlock) void rhs_far_kernel(int64_t ns, const double* __restrict__ sx,
                                                         const double* __restrict__ sy,
                                                         const double* __restrict__ ss,
                                                         const double* __restrict__ head,
                                                         int64_t chunk, int64_t nchunk,
                                                         double* __restrict__ gpart) {
  __shared__ double red[kBlock / 64][2 * kTerms];
  const int tid = threadIdx.x;
  const int64_t c = blockIdx.x;
  const double hx = head[kHead * c], hy = head[kHead * c + 1], hr2 = head[kHead * c + 2];
  double gr[kTerms], gi[kTerms];
#pragma unroll
  for (int n = 0; n < kTerms; ++n) gr[n] = gi[n] = 0.0;
  const int64_t jb = (int64_t)blockIdx.y * chunk;
  const int64_t je = jb + chunk < ns ? jb + chunk : ns;
  for (int64_t j = jb + tid; j < je; j += kBlock) {
    const double dx = sx[j] - hx, dy = sy[j] - hy;
    const double z2 = __builtin_fma(dx, dx, dy * dy);
    if (z2 >= hr2 && z2 > 0.0) {  // far: the complement of points_on_panels_kernel's test
      const double iz = rcp64(z2);
      const double wx = dx * iz, wy = -dy * iz;  // w = 1 / (z - c)
      double pr = ss[j] * wx, pi = ss[j] * wy;   // s w^n
#pragma unroll
      for (int n = 0; n < kTerms; ++n) {
        gr[n] += pr;
        gi[n] += pi;
        const double npr = __builtin_fma(pr, wx, -pi * wy);
        pi = __builtin_fma(pr, wy, pi * wx);
        pr = npr;
      }
    }
  }
  const int wv = tid / 64, ln = tid % 64;
#pragma unroll
  for (int n = 0; n < kTerms; ++n) {
    const double a = wave_sum(gr[n]), b = wave_sum(gi[n]);
    if (ln == 0) {
      red[wv][2 * n] = a;
      red[wv][2 * n + 1] = b;
    }
  }
  __syncthreads();
  if (tid < 2 * kTerms) {
    double acc = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) acc += red[w][tid];
    gpart[((c * nchunk + blockIdx.y) * kTerms) * 2 + tid] = acc;
  }
}

// gsum[(c kTerms + n) 2 + {0, 1}] = sum over the chunks of gpart, in chunk order
__global__ __launch_bounds__(kBlock) void rhs_gsum_kernel(int64_t nclus, int64_t nchunk,
                                                          const double* __restrict__ gpart,
                                                          double* __restrict__ gsum) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;  // (c kTerms + n) 2 + part
  if (e >= nclus * kTerms * 2) return;
  const int64_t c = e / (2 * kTerms), r = e % (2 * kTerms);
  double acc = 0.0;
  for (int64_t y = 0; y < nchunk; ++y) acc += gpart[(c * nchunk + y) * kTerms * 2 + r];
  gsum[e] = acc;
}

// the far particles' share of each panel's sum, as one more chunk row of part: row[{0,1} np + p]
// (the same form as points_on_panels_kernel's rows: Im and Re of conj(e) sum_j s_j L_p(z_j))
__global__ __launch_bounds__(kBlock) void rhs_combine_kernel(int64_t np,
                                                             const double* __restrict__ nx,
                                                             const double* __restrict__ ny,
                                                             const uint32_t* __restrict__ idx,
                                                             const double* __restrict__ head,
                                                             const double* __restrict__ gsum,
                                                             double* __restrict__ row) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= np) return;
  const int64_t c = p / kClus;
  const uint32_t a = idx[2 * p], b = idx[2 * p + 1];
  const PanelPre q = panel_pre(nx[a], ny[a], nx[b], ny[b], 1.0, 0.0);  // W = conj(e)
  const double cx = head[kHead * c], cy = head[kHead * c + 1];
  const double ax = q.x0 - cx, ay = q.y0 - cy, bx = q.x1 - cx, by = q.y1 - cy;
  const double dzx = q.x1 - q.x0, dzy = q.y1 - q.y0;
  double tr = 0.0, ti = 0.0;
  double sx = 1.0, sy = 0.0, px = ax, py = ay;  // S_1, a^1
  for (int n = 1; n <= kTerms; ++n) {
    const double gr = gsum[(c * kTerms + n - 1) * 2], gi = gsum[(c * kTerms + n - 1) * 2 + 1];
    // coefficient (z1 - z0) S_n / n
    const double inv = 1.0 / (double)n;
    const double kr = (dzx * sx - dzy * sy) * inv, ki = (dzx * sy + dzy * sx) * inv;
    tr += kr * gr - ki * gi;
    ti += kr * gi + ki * gr;
    const double nsx = bx * sx - by * sy + px, nsy = bx * sy + by * sx + py;  // b S + a^n
    sx = nsx;
    sy = nsy;
    const double npx = px * ax - py * ay;
    py = px * ay + py * ax;
    px = npx;
  }
  row[p] = __builtin_fma(q.wr, ti, q.wi * tr);
  row[np + p] = __builtin_fma(q.wr, tr, -q.wi * ti);
}

// pu[p] -= (1/L_p) * sum over chunks (src/Influence.h:726-759)
__global__ __launch_bounds__(kBlock) void points_on_panels_finish_kernel(
    int64_t nchunk, int64_t np, const double* __restrict__ nx, const double* __restrict__ ny,
    const uint32_t* __restrict__ idx, const double* __restrict__ part, double* __restrict__ pu,
    double* __restrict__ pv) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= np) return;
  double su = 0.0, sv = 0.0;
  for (int64_t c = 0; c < nchunk; ++c) {
    su += part[(c * 2) * np + p];
    sv += part[(c * 2 + 1) * np + p];
  }
  const uint32_t a = idx[2 * p], b = idx[2 * p + 1];
  const double ex = nx[b] - nx[a], ey = ny[b] - ny[a];
  const double plen = 1.0 / sqrt(ex * ex + ey * ey);
  pu[p] -= plen * su;
  pv[p] -= plen * sv;
}

inline dim3 grid1(int64_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }
constexpr int64_t kClusMin = 4 * kClus;  // panels from which the panel sums take clusters
constexpr int64_t kClusMinSources = 1024;  // particles from which points on panels takes them

}  // namespace

int panels_on_points_dev(Device* d, hipStream_t s, int64_t nnodes, const double* nx,
                         const double* ny, int64_t np, const uint32_t* idx, const double* vs,
                         const double* ss, int64_t nt, const double* tx, const double* ty,
                         double* tu, double* tv) {
  (void)nnodes;
  if (np == 0 || nt == 0) return 0;
  if (np >= kClusMin) {
    // the clustered far field, with the targets in Morton order
    const int64_t nclus = (np + kClus - 1) / kClus;
    void* hp;
    O2D_TRY(slot(d, kPanClus, sizeof(double) * (kHead + 2 * kTerms) * nclus, s, &hp));
    double* head = static_cast<double*>(hp);
    double* mom = head + kHead * nclus;
    cluster_kernel<<<(unsigned)nclus, 64, 0, s>>>(np, nx, ny, idx, vs, ss, head, mom);
    O2D_LAUNCH_CHECK();
    const uint32_t* torder = nullptr;
    if (nt > kBlock) O2D_TRY(morton_order16(d, s, nt, tx, ty, &torder));
    // chunks of clusters for 16 workgroups per CU (C3: 782 target tiles, 6 chunks)
    const int64_t twgs = (nt + kBlock - 1) / kBlock;
    const int64_t want = 16 * (int64_t)d->num_cu;
    int64_t nchunk = twgs >= want ? 1 : std::min<int64_t>((want + twgs - 1) / twgs, nclus);
    const int64_t cchunk = (nclus + nchunk - 1) / nchunk;
    nchunk = (nclus + cchunk - 1) / cchunk;
    double *ou = tu, *ov = tv;
    const int direct = nchunk == 1;
    if (!direct) {
      void* p;
      O2D_TRY(slot(d, kPartials, sizeof(double) * 2 * nchunk * nt, s, &p));
      ou = static_cast<double*>(p);
      ov = ou + nchunk * nt;
    }
    const dim3 grid((unsigned)twgs, (unsigned)nchunk);
    KernelTimer kt(O2D_KT_PANELS, s);
    if (ss)
      panels_clustered_kernel<true><<<grid, kBlock, 0, s>>>(
          np, nx, ny, idx, vs, ss, nclus, head, mom, nt, tx, ty, torder, cchunk, ou, ov, direct);
    else
      panels_clustered_kernel<false><<<grid, kBlock, 0, s>>>(
          np, nx, ny, idx, vs, ss, nclus, head, mom, nt, tx, ty, torder, cchunk, ou, ov, direct);
    kt.stop();
    O2D_LAUNCH_CHECK();
    if (!direct) {
      reduce_sorted_kernel<<<grid1(nt), kBlock, 0, s>>>(nchunk, nt, torder, ou, ov, tu, tv);
      O2D_LAUNCH_CHECK();
    }
    return 0;
  }
  const int64_t twgs = (nt + kBlock - 1) / kBlock;
  // 32 workgroups per CU: at C3 size (200,000 targets, 2,048 panels) that is 11 panel chunks,
  // and 0.96 ms against 1.07 ms with 8 per CU
  const int64_t want = 32 * (int64_t)d->num_cu;
  int64_t nchunk = twgs >= want ? 1 : std::min<int64_t>((want + twgs - 1) / twgs, np);
  const int64_t chunk = (np + nchunk - 1) / nchunk;
  nchunk = (np + chunk - 1) / chunk;
  double *ou = tu, *ov = tv;
  const int direct = nchunk == 1;
  if (!direct) {
    void* p;
    O2D_TRY(slot(d, kPartials, sizeof(double) * 2 * nchunk * nt, s, &p));
    ou = static_cast<double*>(p);
    ov = ou + nchunk * nt;
  }
  const dim3 grid((unsigned)twgs, (unsigned)nchunk);
  KernelTimer kt(O2D_KT_PANELS, s);
  if (ss)
    panels_on_points_kernel<true><<<grid, kBlock, 0, s>>>(np, nx, ny, idx, vs, ss, nt, tx, ty, chunk,
                                                          ou, ov, direct);
  else
    panels_on_points_kernel<false><<<grid, kBlock, 0, s>>>(np, nx, ny, idx, vs, ss, nt, tx, ty,
                                                           chunk, ou, ov, direct);
  kt.stop();
  O2D_LAUNCH_CHECK();
  if (!direct) {
    reduce_chunks2_kernel<<<grid1(nt), kBlock, 0, s>>>(nchunk, nt, ou, ov, tu, tv);
    O2D_LAUNCH_CHECK();
  }
  return 0;
}

int points_on_panels_dev(Device* d, hipStream_t s, int64_t ns, const double* sx, const double* sy,
                         const double* ss, int64_t nnodes, const double* nx, const double* ny,
                         int64_t np, const uint32_t* idx, double* pu, double* pv) {
  (void)nnodes;
  if (np == 0 || ns == 0) return 0;
  const int64_t pwgs = (np + kPanTile - 1) / kPanTile;
  const int64_t want = O2D_WANT_RHS * (int64_t)d->num_cu;
  int64_t nchunk = pwgs >= want ? 1 : (want + pwgs - 1) / pwgs;
  const int64_t maxchunk = (ns + kBlock - 1) / kBlock;  // at least one particle per lane
  if (nchunk > maxchunk) nchunk = maxchunk;
  int64_t chunk = (ns + nchunk - 1) / nchunk;
  // clusters when there are enough panels and particles: the near particles of each cluster
  // (in Morton order, so that a wave's lanes agree) by the direct kernel, the far ones by the
  // cluster expansion, as one more row of partial sums
  const bool clus = np >= kClusMin && ns >= kClusMinSources;
  const int64_t nclus = (np + kClus - 1) / kClus;
  double* head = nullptr;
  double* gbox = nullptr;
  const double *nsx = sx, *nsy = sy, *nss = ss;  // the near kernel's particles
  if (clus) {
    void* hp;
    O2D_TRY(slot(d, kPanClus, sizeof(double) * kHead * nclus, s, &hp));
    head = static_cast<double*>(hp);
    cluster_kernel<<<(unsigned)nclus, 64, 0, s>>>(np, nx, ny, idx, nullptr, nullptr, head,
                                                  nullptr);
    O2D_LAUNCH_CHECK();
    const uint32_t* sorder = nullptr;
    O2D_TRY(morton_order16(d, s, ns, sx, sy, &sorder));
    const int64_t ngroups = (ns + 63) / 64;
    void* sp;
    O2D_TRY(slot(d, kPanClusS, sizeof(double) * (3 * ns + 4 * ngroups), s, &sp));
    double* xs = static_cast<double*>(sp);
    gbox = xs + 3 * ns;
    gather_groups_kernel<<<grid1(ns), kBlock, 0, s>>>(ns, sx, sy, ss, sorder, xs, xs + ns,
                                                      xs + 2 * ns, gbox);
    O2D_LAUNCH_CHECK();
    nsx = xs;
    nsy = xs + ns;
    nss = xs + 2 * ns;
    chunk = (chunk + kBlock - 1) / kBlock * kBlock;  // whole groups of 64 (and of the block)
  }
  nchunk = (ns + chunk - 1) / chunk;
  const int64_t rows = nchunk + (clus ? 1 : 0);
  void* p;
  O2D_TRY(slot(d, kPartials, sizeof(double) * 2 * rows * np, s, &p));
  double* part = static_cast<double*>(p);
  // far chunks for 16 workgroups per CU
  int64_t fchunk = 0, nfchunk = 0;
  double* gpart = nullptr;
  if (clus) {
    const int64_t fwant = 16 * (int64_t)d->num_cu;
    nfchunk = nclus >= fwant ? 1 : (fwant + nclus - 1) / nclus;
    if (nfchunk > maxchunk) nfchunk = maxchunk;
    fchunk = (ns + nfchunk - 1) / nfchunk;
    nfchunk = (ns + fchunk - 1) / fchunk;
    void* gp;
    O2D_TRY(slot(d, kPanClusG, sizeof(double) * 2 * kTerms * nclus * (nfchunk + 1), s, &gp));
    gpart = static_cast<double*>(gp);
  }
  KernelTimer kt(O2D_KT_PANEL_RHS, s);
  points_on_panels_kernel<<<dim3((unsigned)pwgs, (unsigned)nchunk), kBlock, 0, s>>>(
      ns, nsx, nsy, nss, np, nx, ny, idx, chunk, part, head, gbox);
  O2D_LAUNCH_CHECK();
  if (clus) {
    rhs_far_kernel<<<dim3((unsigned)nclus, (unsigned)nfchunk), kBlock, 0, s>>>(
        ns, sx, sy, ss, head, fchunk, nfchunk, gpart);
    O2D_LAUNCH_CHECK();
    double* gsum = gpart + 2 * kTerms * nclus * nfchunk;
    rhs_gsum_kernel<<<grid1(2 * kTerms * nclus), kBlock, 0, s>>>(nclus, nfchunk, gpart, gsum);
    O2D_LAUNCH_CHECK();
    rhs_combine_kernel<<<grid1(np), kBlock, 0, s>>>(np, nx, ny, idx, head, gsum,
                                                    part + 2 * nchunk * np);
    O2D_LAUNCH_CHECK();
  }
  kt.stop();
  points_on_panels_finish_kernel<<<grid1(np), kBlock, 0, s>>>(rows, np, nx, ny, idx, part, pu, pv);
  O2D_LAUNCH_CHECK();
  return 0;
}

// shared validation of panel geometry arguments
static int check_panels(int64_t nnodes, const void* nx, const void* ny, int64_t np,
                        const uint32_t* idx_host) {
  if (nnodes < 0 || np < 0) return fail(O2D_ERR_SIZE, "negative count");
  if (np > 0x7fffffffLL || nnodes > 0xffffffffLL) return fail(O2D_ERR_SIZE, "too many panels");
  if (np > 0 && (!nx || !ny || !idx_host)) return fail(O2D_ERR_ARG, "NULL panel array");
  return 0;
}

static int check_idx_host(int64_t nnodes, int64_t np, const uint32_t* idx) {
  for (int64_t j = 0; j < 2 * np; ++j)
    if ((int64_t)idx[j] >= nnodes)
      return fail(O2D_ERR_SIZE, "panel node index %u at %lld >= nnodes %lld", idx[j],
                  (long long)j, (long long)nnodes);
  return 0;
}

}  // namespace o2d

using namespace o2d;

extern "C" {

int o2d_hip_panels_on_points_dev_d(const o2d_env* env, int64_t nnodes, const double* nx,
                                   const double* ny, int64_t np, const uint32_t* idx,
                                   const double* vstr, const double* sstr, int64_t nt,
                                   const double* tx, const double* ty, double* tu, double* tv) {
  clear_error();
  O2D_TRY(check_env(env));
  if (env->results != O2D_VELONLY)
    return fail(O2D_ERR_RESULTS, "panels compute velocity only (results=%d)", env->results);
  O2D_TRY(check_panels(nnodes, nx, ny, np, idx));
  if (np > 0 && !vstr) return fail(O2D_ERR_ARG, "NULL vstr");
  if (nt < 0 || nt > 0x7fffffffLL) return fail(O2D_ERR_SIZE, "bad target count");
  if (nt > 0 && (!tx || !ty || !tu || !tv)) return fail(O2D_ERR_ARG, "NULL target array");
  Device* d;
  O2D_TRY(device(env->device, &d));
  return panels_on_points_dev(d, dev_stream(env), nnodes, nx, ny, np, idx, vstr, sstr, nt, tx,
                              ty, tu, tv);
}

int o2d_hip_panels_on_points_d(const o2d_env* env, int64_t nnodes, const double* nx,
                               const double* ny, int64_t np, const uint32_t* idx,
                               const double* vstr, const double* sstr, int64_t nt,
                               const double* tx, const double* ty, double* tu, double* tv) {
  clear_error();
  O2D_TRY(check_env(env));
  if (env->results != O2D_VELONLY)
    return fail(O2D_ERR_RESULTS, "panels compute velocity only (results=%d)", env->results);
  O2D_TRY(check_panels(nnodes, nx, ny, np, idx));
  if (np > 0 && !vstr) return fail(O2D_ERR_ARG, "NULL vstr");
  if (nt < 0 || nt > 0x7fffffffLL) return fail(O2D_ERR_SIZE, "bad target count");
  if (nt > 0 && (!tx || !ty || !tu || !tv)) return fail(O2D_ERR_ARG, "NULL target array");
  O2D_TRY(check_idx_host(nnodes, np, idx));
  if (np == 0 || nt == 0) return 0;
  Device* d;
  O2D_TRY(device(env->device, &d));
  hipStream_t s = stream_of(env, d);
  double *dnx, *dny, *dvs, *dss = nullptr, *dtx, *dty, *du, *dv;
  uint32_t* didx;
  O2D_TRY(upload<double>(d, kPanX, nx, nnodes, s, &dnx));
  O2D_TRY(upload<double>(d, kPanY, ny, nnodes, s, &dny));
  O2D_TRY(upload<uint32_t>(d, kPanIdx, idx, 2 * np, s, &didx));
  O2D_TRY(upload<double>(d, kPanVs, vstr, np, s, &dvs));
  if (sstr) O2D_TRY(upload<double>(d, kPanSs, sstr, np, s, &dss));
  O2D_TRY(upload<double>(d, kTrgX, tx, nt, s, &dtx));
  O2D_TRY(upload<double>(d, kTrgY, ty, nt, s, &dty));
  O2D_TRY(upload<double>(d, kOutU, tu, nt, s, &du));
  O2D_TRY(upload<double>(d, kOutV, tv, nt, s, &dv));
  O2D_TRY(panels_on_points_dev(d, s, nnodes, dnx, dny, np, didx, dvs, dss, nt, dtx, dty, du, dv));
  O2D_TRY(download<double>(tu, du, nt, s));
  O2D_TRY(download<double>(tv, dv, nt, s));
  O2D_HIP(hipStreamSynchronize(s));
  return 0;
}

int o2d_hip_points_on_panels_dev_d(const o2d_env* env, int64_t ns, const double* sx,
                                   const double* sy, const double* sstr, int64_t nnodes,
                                   const double* nx, const double* ny, int64_t np,
                                   const uint32_t* idx, double* pu, double* pv) {
  clear_error();
  O2D_TRY(check_env(env));
  if (env->results != O2D_VELONLY)
    return fail(O2D_ERR_RESULTS, "panels compute velocity only (results=%d)", env->results);
  O2D_TRY(check_panels(nnodes, nx, ny, np, idx));
  if (ns < 0) return fail(O2D_ERR_SIZE, "negative source count");
  if (ns > 0 && (!sx || !sy || !sstr)) return fail(O2D_ERR_ARG, "NULL source array");
  if (np > 0 && (!pu || !pv)) return fail(O2D_ERR_ARG, "NULL output");
  Device* d;
  O2D_TRY(device(env->device, &d));
  return points_on_panels_dev(d, dev_stream(env), ns, sx, sy, sstr, nnodes, nx, ny, np, idx, pu,
                              pv);
}

int o2d_hip_points_on_panels_d(const o2d_env* env, int64_t ns, const double* sx,
                               const double* sy, const double* sstr, int64_t nnodes,
                               const double* nx, const double* ny, int64_t np,
                               const uint32_t* idx, double* pu, double* pv) {
  clear_error();
  O2D_TRY(check_env(env));
  if (env->results != O2D_VELONLY)
    return fail(O2D_ERR_RESULTS, "panels compute velocity only (results=%d)", env->results);
  O2D_TRY(check_panels(nnodes, nx, ny, np, idx));
  if (ns < 0) return fail(O2D_ERR_SIZE, "negative source count");
  if (ns > 0 && (!sx || !sy || !sstr)) return fail(O2D_ERR_ARG, "NULL source array");
  if (np > 0 && (!pu || !pv)) return fail(O2D_ERR_ARG, "NULL output");
  O2D_TRY(check_idx_host(nnodes, np, idx));
  if (np == 0 || ns == 0) return 0;
  Device* d;
  O2D_TRY(device(env->device, &d));
  hipStream_t s = stream_of(env, d);
  double *dsx, *dsy, *dss, *dnx, *dny, *du, *dv;
  uint32_t* didx;
  O2D_TRY(upload<double>(d, kSrcX, sx, ns, s, &dsx));
  O2D_TRY(upload<double>(d, kSrcY, sy, ns, s, &dsy));
  O2D_TRY(upload<double>(d, kSrcS, sstr, ns, s, &dss));
  O2D_TRY(upload<double>(d, kPanX, nx, nnodes, s, &dnx));
  O2D_TRY(upload<double>(d, kPanY, ny, nnodes, s, &dny));
  O2D_TRY(upload<uint32_t>(d, kPanIdx, idx, 2 * np, s, &didx));
  O2D_TRY(upload<double>(d, kOutU, pu, np, s, &du));
  O2D_TRY(upload<double>(d, kOutV, pv, np, s, &dv));
  O2D_TRY(points_on_panels_dev(d, s, ns, dsx, dsy, dss, nnodes, dnx, dny, np, didx, du, dv));
  O2D_TRY(download<double>(pu, du, np, s));
  O2D_TRY(download<double>(pv, dv, np, s));
  O2D_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // extern "C"
