// treecode.hip - O(N log N) particles-on-points velocity, the summation the reference reserves
// as summation_t barneshut (src/ExecEnv.h:26-31) behind the same operator as the direct sum.
//
// Tree: sources in Morton order, cut into leaves of kLeaf consecutive sources; each coarser
// level groups kFan consecutive cells, so every cell is one contiguous Morton range and the tree
// needs no pointers. Per cell: bounding-box centre c, radius r (covers every source), largest
// radius term, and the complex moments M_k = sum_j s_j (z_j - c)^k, k < p (fp64). Parents get
// their moments by the binomial shift of their children's.
//
// Evaluation: one wave per 64 Morton-ordered targets walks the tree from the root with a stack
// in LDS. A cell is used through its moments when, for every lane of the wave,
//   r <= theta |z - c|                      (truncation of the Laurent series)
//   the core equals 1/d^2 for every pair within core_tol, from d_min = |z - c| - r: exactly
//   beyond the Gaussian's compact cutoff (d^2 > 16 w, src/CoreFunc.h:120-122); for the other
//   cores through the bound on |K d^2 - 1| (RM w/d^2, WL w^2/d^4, V2 w/(2 d^4), V3 w/(3 d^6)).
// Otherwise the cell is opened, and a leaf is summed directly with the exact core in fp64.
// With S = sum_k M_k / (z - c)^(k+1), the unscaled velocity is u - i v = -i S, i.e.
// u = Im S, v = Re S.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "nbody.h"
#include "o2d_math.h"
#include "o2d_runtime.h"

namespace o2d {

namespace {

constexpr int kLeaf = 64;     // sources per leaf
constexpr int kFan = 4;       // children per cell
constexpr int kMaxP = 32;     // largest expansion order
constexpr int kMaxLevels = 15;  // 2^31 sources make 14 levels
constexpr int kTBlock = 256;  // 4 waves, each with its own 64 targets
constexpr int kStack = 96;    // per-wave stack entries (depth <= (kFan - 1) * levels + 1)
constexpr int kBlock = 256;

struct Cell {
  double cx, cy, r, wmax;
  double x0, y0, x1, y1;  // bounding box of the cell's sources
};

struct Levels {
  int n;                      // number of levels (level 0 = leaves, n - 1 = root)
  int64_t off[kMaxLevels];    // first cell of each level in the cell array
  int64_t cnt[kMaxLevels];    // cells per level
};

__global__ __launch_bounds__(kBlock) void tree_gather_kernel(
    int64_t n, const uint32_t* __restrict__ order, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ s, const double* __restrict__ r,
    int core, double* __restrict__ ox, double* __restrict__ oy, double* __restrict__ os,
    double* __restrict__ ow) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = order[i];
  ox[i] = x[j];
  oy[i] = y[j];
  if (os) os[i] = s[j];
  if (ow) {
    const double rr = r ? r[j] : 0.0;
    ow[i] = core == O2D_CORE_V2   ? rterm<O2D_CORE_V2>(rr)
            : core == O2D_CORE_V3 ? rterm<O2D_CORE_V3>(rr)
                                  : rterm<O2D_CORE_RM>(rr);
  }
}

// wave-wide reductions (every lane receives the result)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmin(v, __shfl_xor(v, m));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m));
  return v;
}

// leaves, one wave each (lane j holds source j of the leaf): bounding box, centre, radius,
// largest radius term and the moments M_k = sum_j s_j (z_j - c)^k of <= kLeaf sources, each
// moment a wave sum; lane k keeps M_k
__global__ __launch_bounds__(kBlock) void tree_leaf_kernel(
    int64_t ns, int64_t nleaf, int p, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ s, const double* __restrict__ w, Cell* __restrict__ cell,
    double2* __restrict__ mom) {
  const int lane = threadIdx.x & 63;
  const int64_t c = (int64_t)blockIdx.x * (kBlock / 64) + threadIdx.x / 64;
  if (c >= nleaf) return;  // whole waves leave together
  const int64_t j = c * kLeaf + lane;
  const bool in = j < ns;
  const double xj = in ? x[j] : 0.0, yj = in ? y[j] : 0.0;
  const double x0 = wave_min(in ? xj : INFINITY), x1 = wave_max(in ? xj : -INFINITY);
  const double y0 = wave_min(in ? yj : INFINITY), y1 = wave_max(in ? yj : -INFINITY);
  const double wm = wave_max(in ? w[j] : 0.0);
  const double cx = 0.5 * (x0 + x1), cy = 0.5 * (y0 + y1);
  const double dx = xj - cx, dy = yj - cy;
  const double r2 = wave_max(in ? dx * dx + dy * dy : 0.0);
  double pr = in ? s[j] : 0.0, pi = 0.0;  // s (z - c)^k
  double mr = 0.0, mi = 0.0;
  for (int k = 0; k < p; ++k) {
    const double sr = wave_sum(pr), si = wave_sum(pi);
    if (lane == k) {
      mr = sr;
      mi = si;
    }
    const double nr = pr * dx - pi * dy;
    pi = pr * dy + pi * dx;
    pr = nr;
  }
  // rounding of the offsets: a relative margin keeps r an upper bound
  if (lane == 0) cell[c] = Cell{cx, cy, sqrt(r2) * (1.0 + 1e-12), wm, x0, y0, x1, y1};
  if (lane < p) mom[c * p + lane] = make_double2(mr, mi);
}

// one coarser level, one wave per cell: bounding-box union, covering radius, and moments
// shifted to the new centre, lane k forming M'_k = sum_children sum_{l <= k} C(k, l) M_l d^(k-l)
// with d = c_child - c (the children's moments staged in LDS)
__global__ __launch_bounds__(kBlock) void tree_up_kernel(int64_t nchild, int64_t ncell, int p,
                                                         const Cell* __restrict__ child,
                                                         const double2* __restrict__ cmom,
                                                         Cell* __restrict__ cell,
                                                         double2* __restrict__ mom) {
  __shared__ double2 sm[kBlock / 64][kFan * kMaxP];
  const int lane = threadIdx.x & 63, wv = threadIdx.x / 64;
  const int64_t c = (int64_t)blockIdx.x * (kBlock / 64) + wv;
  if (c >= ncell) return;  // whole waves leave together; no block-wide barrier below
  const int64_t k0 = c * kFan;
  const int nch = (int)((k0 + kFan < nchild) ? kFan : nchild - k0);
  for (int e = lane; e < nch * p; e += 64) sm[wv][e] = cmom[k0 * p + e];
  double x0 = child[k0].x0, y0 = child[k0].y0, x1 = child[k0].x1, y1 = child[k0].y1, wm = 0.0;
  for (int k = 0; k < nch; ++k) {
    x0 = fmin(x0, child[k0 + k].x0);
    y0 = fmin(y0, child[k0 + k].y0);
    x1 = fmax(x1, child[k0 + k].x1);
    y1 = fmax(y1, child[k0 + k].y1);
    wm = fmax(wm, child[k0 + k].wmax);
  }
  const double cx = 0.5 * (x0 + x1), cy = 0.5 * (y0 + y1);
  // the staged moments are read by other lanes of the same wave
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  double r = 0.0, ar = 0.0, ai = 0.0;
  for (int ch = 0; ch < nch; ++ch) {
    const double dx = child[k0 + ch].cx - cx, dy = child[k0 + ch].cy - cy;
    r = fmax(r, sqrt(dx * dx + dy * dy) + child[k0 + ch].r);
    if (lane < p) {
      const int k = lane;
      // l from k down to 0: C(k, l) M_l d^(k - l), with d^(k-l) and C(k, l) by recurrence
      double dpr = 1.0, dpi = 0.0, binom = 1.0;
      for (int l = k; l >= 0; --l) {
        const double2 ml = sm[wv][ch * p + l];
        ar += binom * (ml.x * dpr - ml.y * dpi);
        ai += binom * (ml.x * dpi + ml.y * dpr);
        const double nr = dpr * dx - dpi * dy;
        dpi = dpr * dy + dpi * dx;
        dpr = nr;
        binom = binom * (double)l / (double)(k - l + 1);
      }
    }
  }
  if (lane == 0) cell[c] = Cell{cx, cy, r * (1.0 + 1e-12), wm, x0, y0, x1, y1};
  if (lane < p) mom[c * p + lane] = make_double2(ar, ai);
}

// may the moments of a cell stand for its sources at a target d_min away?
template <int CORE>
__device__ __forceinline__ bool core_far(double dm2, double w, double tol) {
  if constexpr (CORE == O2D_CORE_GAUSSIAN) {
    return dm2 > 16.0 * w * (1.0 + 1e-12);
  } else if constexpr (CORE == O2D_CORE_RM) {
    return w <= tol * dm2;
  } else if constexpr (CORE == O2D_CORE_WL) {
    return w * w <= tol * dm2 * dm2;
  } else if constexpr (CORE == O2D_CORE_V2) {
    return w <= 2.0 * tol * dm2 * dm2;
  } else {
    return w <= 3.0 * tol * dm2 * dm2 * dm2;
  }
}

struct TreeArgs {
  Levels lv;
  const Cell* cell;
  const double2* mom;
  const double* sx;
  const double* sy;
  const double* ss;
  const double* sw;
  int64_t ns;
  int p;
  double theta;
  double tol;
  // the walk may be split over ksplit waves per target group: wave k starts from the cells of
  // level slevel whose index is k mod ksplit and writes its partial sums to part
  int ksplit;
  int slevel;
  double* part;  // [2][ksplit][nt]: u then v partials (ksplit > 1)
};

// makes an in-wave LDS write visible to the other lanes of the wave before they read it, and
// keeps later writes behind earlier reads
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct alignas(16) LeafSrc {
  double x, y, s, w;
};

// Every lane of a wave walks the same cells. The popped stack entry is made wave-uniform
// (readfirstlane), so cell records come in through scalar loads; a leaf's sources and a cell's
// moments are loaded once per wave, one element per lane (coalesced), into the wave's LDS and
// read back as broadcasts, instead of one dependent load per pair or per coefficient.
template <int CORE>
__global__ __launch_bounds__(kTBlock) void tree_eval_kernel(TreeArgs a, int64_t nt,
                                                            const uint32_t* __restrict__ torder,
                                                            const double* __restrict__ tx,
                                                            const double* __restrict__ ty,
                                                            const double* __restrict__ tr,
                                                            double* __restrict__ tu,
                                                            double* __restrict__ tv) {
  constexpr int kWaves = kTBlock / 64;
  __shared__ unsigned stack[kWaves][kStack];
  __shared__ LeafSrc lsrc[kWaves][kLeaf];
  __shared__ double2 lmom[kWaves][kMaxP];
  const int wave = threadIdx.x / 64, lane = threadIdx.x & 63;
  unsigned* st = stack[wave];
  LeafSrc* ls = lsrc[wave];
  double2* lm = lmom[wave];
  const int64_t i = (int64_t)blockIdx.x * kTBlock + threadIdx.x;
  const bool valid = i < nt;
  const uint32_t ti = valid ? torder[i] : 0u;
  const double x = valid ? tx[ti] : 0.0, y = valid ? ty[ti] : 0.0;
  const double wt = (valid && tr) ? rterm<CORE>(tr[ti]) : 0.0;
  double u = 0.0, v = 0.0;     // direct part
  double sr = 0.0, si = 0.0;   // S, the far part
  const double th2 = a.theta * a.theta;
  const int p = a.p;
  // a wave with no valid lane (the tail of the grid) has nothing to do
  if (__any(valid)) {
    int sp = 0;
    if (a.ksplit == 1) {
      const int root = a.lv.n - 1;
      if (lane == 0) st[0] = (unsigned)root << 27;  // (level << 27) | index; the root is index 0
      sp = 1;
    } else {
      // cells k, k + ksplit, ... of level slevel, the smallest index on top
      const int k = blockIdx.y;
      const int ck = (int)((a.lv.cnt[a.slevel] - k + a.ksplit - 1) / a.ksplit);
      if (lane < ck)
        st[lane] = ((unsigned)a.slevel << 27) | (unsigned)(k + (ck - 1 - lane) * a.ksplit);
      sp = ck;
    }
    wave_lds_sync();
    while (sp > 0) {
      const unsigned e = __builtin_amdgcn_readfirstlane(st[--sp]);
      const int lev = (int)(e >> 27), idx = (int)(e & ((1u << 27) - 1));
      const int64_t ci = a.lv.off[lev] + idx;
      const Cell c = a.cell[ci];
      const double dx = x - c.cx, dy = y - c.cy;
      const double d2 = dx * dx + dy * dy;
      const double dm = sqrt(d2) - c.r;  // distance to the nearest possible source of the cell
      // every pair with this cell is beyond the core's cutoff (or within core_tol of 1/d^2)
      const bool cfar = !valid || (dm > 0.0 && core_far<CORE>(dm * dm, c.wmax + wt, a.tol));
      const bool acc = cfar && (!valid || c.r * c.r <= th2 * d2);
      if (__all(acc)) {
        if (lane < p) lm[lane] = a.mom[ci * p + lane];
        wave_lds_sync();
        if (valid) {
          // t = 1/(z - c); S += sum_k M_k t^(k+1) by Horner
          const double q = 1.0 / d2;
          const double tre = dx * q, tim = -dy * q;
          double hr = lm[p - 1].x, hi = lm[p - 1].y;
#pragma unroll 4
          for (int k = p - 2; k >= 0; --k) {
            const double2 mk = lm[k];
            const double nr = hr * tre - hi * tim + mk.x;
            hi = hr * tim + hi * tre + mk.y;
            hr = nr;
          }
          sr += hr * tre - hi * tim;
          si += hr * tim + hi * tre;
        }
        wave_lds_sync();
      } else if (lev == 0) {
        const int64_t j0 = (int64_t)idx * kLeaf;
        const int nj = (int)((j0 + kLeaf < a.ns) ? kLeaf : a.ns - j0);
        if (lane < nj) {
          const int64_t j = j0 + lane;
          ls[lane] = LeafSrc{a.sx[j], a.sy[j], a.ss[j], a.sw[j]};
        }
        wave_lds_sync();
        if (CORE == O2D_CORE_GAUSSIAN && __all(cfar)) {
          // the whole leaf is beyond the compact cutoff for every lane: K = 1/d^2
          if (valid) {
#pragma unroll 8
            for (int k = 0; k < nj; ++k) {
              const LeafSrc q = ls[k];
              const double ex = x - q.x, ey = y - q.y;
              const double g = q.s * rcp64(__builtin_fma(ex, ex, ey * ey));
              u = __builtin_fma(-g, ey, u);
              v = __builtin_fma(g, ex, v);
            }
          }
        } else if (valid) {
#pragma unroll 4
          for (int k = 0; k < nj; ++k) {
            const LeafSrc q = ls[k];
            const double ex = x - q.x, ey = y - q.y;
            const double e2 = __builtin_fma(ex, ex, ey * ey);
            const double g = q.s * core_k64<CORE>(ex, ey, e2, wsum(q.w, wt));
            u = __builtin_fma(-g, ey, u);
            v = __builtin_fma(g, ex, v);
          }
        }
        wave_lds_sync();
      } else {
        const int64_t first = (int64_t)idx * kFan;
        const int64_t cnt = a.lv.cnt[lev - 1];
        // lane k writes child kFan-1-k, so the first child ends on top of the stack
        int nc = 0;
        for (int k = 0; k < kFan; ++k) nc += (first + k < cnt) ? 1 : 0;
        if (lane < nc)
          st[sp + lane] = ((unsigned)(lev - 1) << 27) | (unsigned)(first + nc - 1 - lane);
        sp += nc;
        wave_lds_sync();
      }
    }
  }
  if (valid) {
    if (a.ksplit == 1) {
      tu[ti] += u + si;
      tv[ti] += v + sr;
    } else {
      a.part[(int64_t)blockIdx.y * nt + ti] = u + si;
      a.part[(int64_t)(a.ksplit + blockIdx.y) * nt + ti] = v + sr;
    }
  }
}

// tu[i] += sum over k (in order) of the split walk's partials
__global__ __launch_bounds__(kBlock) void tree_reduce_kernel(int64_t nt, int ksplit,
                                                             const double* __restrict__ part,
                                                             double* __restrict__ tu,
                                                             double* __restrict__ tv) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= nt) return;
  double su = 0.0, sv = 0.0;
  for (int k = 0; k < ksplit; ++k) {
    su += part[(int64_t)k * nt + i];
    sv += part[(int64_t)(ksplit + k) * nt + i];
  }
  tu[i] += su;
  tv[i] += sv;
}

}  // namespace

int tree_dev(Device* d, hipStream_t s, int core, double theta, int p, double tol, int64_t ns,
             const double* sx, const double* sy, const double* ss, const double* sr, int64_t nt,
             const double* tx, const double* ty, const double* tr, double* tu, double* tv) {
  if (ns == 0 || nt == 0) return 0;
  // levels
  Levels lv{};
  int64_t cnt = (ns + kLeaf - 1) / kLeaf, total = 0;
  for (;;) {
    if (lv.n == kMaxLevels) return fail(O2D_ERR_SIZE, "too many sources for the tree");
    lv.off[lv.n] = total;
    lv.cnt[lv.n] = cnt;
    total += cnt;
    ++lv.n;
    if (cnt == 1) break;
    cnt = (cnt + kFan - 1) / kFan;
  }
  // a split walk starts from up to kFan cells of one level
  if ((kFan - 1) * lv.n + kFan > kStack) return fail(O2D_ERR_SIZE, "tree too deep");
  // buffers
  void *ko, *to, *tk, *sbuf, *cb, *mb;
  O2D_TRY(slot(d, kTreeKeys, sizeof(uint32_t) * (ns > nt ? ns : nt), s, &tk));
  O2D_TRY(slot(d, kTreeOrder, sizeof(uint32_t) * ns, s, &ko));
  O2D_TRY(slot(d, kTreeTOrder, sizeof(uint32_t) * nt, s, &to));
  O2D_TRY(slot(d, kTreeSrc, sizeof(double) * 4 * ns, s, &sbuf));
  O2D_TRY(slot(d, kTreeCells, sizeof(Cell) * total, s, &cb));
  O2D_TRY(slot(d, kTreeMom, sizeof(double2) * p * total, s, &mb));
  uint32_t* sorder = static_cast<uint32_t*>(ko);
  uint32_t* torder = static_cast<uint32_t*>(to);
  double* xs = static_cast<double*>(sbuf);
  double *ys = xs + ns, *ssorted = ys + ns, *ws = ssorted + ns;
  Cell* cells = static_cast<Cell*>(cb);
  double2* mom = static_cast<double2*>(mb);
  {
    KernelTimer kt(O2D_KT_TREE_BUILD, s);
    O2D_TRY(morton_order(d, s, ns, sx, sy, true, 0.0, 0.0, 0.0, static_cast<uint32_t*>(tk),
                         sorder));
    tree_gather_kernel<<<(unsigned)((ns + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        ns, sorder, sx, sy, ss, sr, core, xs, ys, ssorted, ws);
    O2D_LAUNCH_CHECK();
    constexpr int kCellsPerBlock = kBlock / 64;  // one wave per cell
    tree_leaf_kernel<<<(unsigned)((lv.cnt[0] + kCellsPerBlock - 1) / kCellsPerBlock), kBlock, 0,
                       s>>>(ns, lv.cnt[0], p, xs, ys, ssorted, ws, cells, mom);
    O2D_LAUNCH_CHECK();
    for (int l = 1; l < lv.n; ++l) {
      tree_up_kernel<<<(unsigned)((lv.cnt[l] + kCellsPerBlock - 1) / kCellsPerBlock), kBlock, 0, s>>>(
          lv.cnt[l - 1], lv.cnt[l], p, cells + lv.off[l - 1], mom + lv.off[l - 1] * p,
          cells + lv.off[l], mom + lv.off[l] * p);
      O2D_LAUNCH_CHECK();
    }
    O2D_TRY(morton_order(d, s, nt, tx, ty, true, 0.0, 0.0, 0.0, static_cast<uint32_t*>(tk),
                         torder));
    kt.stop();
  }
  // each wave's walk is a chain of dependent loads, so the walk wants many waves per SIMD, and
  // the groups' costs vary widely: below 256 waves per CU, split the walk of each group over up
  // to 64 waves, each from its share of one level's cells (measured on the 1M blob workload: the
  // tracers' 1,024 groups over 64 waves each and the particles' 16,384 groups over 4 each took
  // 6.1 ms, against 7.5 ms for 4 and 1)
  const int64_t waves = (nt + 63) / 64, want = 256 * (int64_t)d->num_cu;
  int ksplit = 1, slevel = lv.n - 1;
  if (waves < want && lv.n > 1) {
    ksplit = (int)std::min<int64_t>(64, (want + waves - 1) / waves);
    for (int l = lv.n - 1; l >= 0; --l)
      if (lv.cnt[l] >= ksplit) {
        slevel = l;
        break;
      }
    if (lv.cnt[slevel] < ksplit) ksplit = (int)lv.cnt[slevel];
  }
  double* part = nullptr;
  if (ksplit > 1) {
    void* pp;
    O2D_TRY(slot(d, kPartials, sizeof(double) * 2 * ksplit * nt, s, &pp));
    part = static_cast<double*>(pp);
  }
  TreeArgs a{lv, cells, mom, xs, ys, ssorted, ws, ns, p, theta, tol, ksplit, slevel, part};
  const dim3 grid((unsigned)((nt + kTBlock - 1) / kTBlock), (unsigned)ksplit);
  KernelTimer kt(O2D_KT_TREE_EVAL, s);
  switch (core) {
    case O2D_CORE_RM:
      tree_eval_kernel<O2D_CORE_RM><<<grid, kTBlock, 0, s>>>(a, nt, torder, tx, ty, tr, tu, tv);
      break;
    case O2D_CORE_GAUSSIAN:
      tree_eval_kernel<O2D_CORE_GAUSSIAN><<<grid, kTBlock, 0, s>>>(a, nt, torder, tx, ty, tr, tu,
                                                                   tv);
      break;
    case O2D_CORE_WL:
      tree_eval_kernel<O2D_CORE_WL><<<grid, kTBlock, 0, s>>>(a, nt, torder, tx, ty, tr, tu, tv);
      break;
    case O2D_CORE_V2:
      tree_eval_kernel<O2D_CORE_V2><<<grid, kTBlock, 0, s>>>(a, nt, torder, tx, ty, tr, tu, tv);
      break;
    default:
      tree_eval_kernel<O2D_CORE_V3><<<grid, kTBlock, 0, s>>>(a, nt, torder, tx, ty, tr, tu, tv);
      break;
  }
  kt.stop();
  O2D_LAUNCH_CHECK();
  if (ksplit > 1) {
    tree_reduce_kernel<<<(unsigned)((nt + kBlock - 1) / kBlock), kBlock, 0, s>>>(nt, ksplit, part,
                                                                               tu, tv);
    O2D_LAUNCH_CHECK();
  }
  return 0;
}

}  // namespace o2d

using namespace o2d;

extern "C" {

static int check_tree(const o2d_env* env, const o2d_tree_opts* opt, int64_t ns, const void* sx,
                      const void* sy, const void* ss, const void* sr, int64_t nt,
                      const void* tx, const void* ty, const void* tu, const void* tv) {
  O2D_TRY(check_env(env));
  if (env->results != O2D_VELONLY)
    return fail(O2D_ERR_RESULTS, "the treecode computes velocity only (results=%d)",
                env->results);
  if (!opt) return fail(O2D_ERR_ARG, "tree options are NULL");
  if (!(opt->theta > 0.0 && opt->theta < 1.0))
    return fail(O2D_ERR_ARG, "theta must lie in (0, 1) (got %g)", opt->theta);
  if (opt->order < 1 || opt->order > kMaxP)
    return fail(O2D_ERR_ARG, "expansion order must be 1..%d (got %d)", kMaxP, (int)opt->order);
  if (!(opt->core_tol >= 0.0)) return fail(O2D_ERR_ARG, "core_tol must be >= 0");
  if (ns < 0 || nt < 0) return fail(O2D_ERR_SIZE, "negative count");
  if (ns > 0x7fffffffLL || nt > 0x7fffffffLL) return fail(O2D_ERR_SIZE, "count too large");
  if (ns > 0 && (!sx || !sy || !ss || !sr)) return fail(O2D_ERR_ARG, "NULL source array");
  if (nt > 0 && (!tx || !ty || !tu || !tv)) return fail(O2D_ERR_ARG, "NULL target array");
  return 0;
}

int o2d_hip_points_on_points_tree_dev_d(const o2d_env* env, const o2d_tree_opts* opt,
                                        int64_t ns, const double* sx, const double* sy,
                                        const double* sstr, const double* srad, int64_t nt,
                                        const double* tx, const double* ty, const double* trad,
                                        double* tu, double* tv) {
  clear_error();
  O2D_TRY(check_tree(env, opt, ns, sx, sy, sstr, srad, nt, tx, ty, tu, tv));
  if (ns == 0 || nt == 0) return 0;
  Device* d;
  O2D_TRY(device(env->device, &d));
  return tree_dev(d, dev_stream(env), env->core, opt->theta, opt->order, opt->core_tol, ns, sx,
                  sy, sstr, srad, nt, tx, ty, trad, tu, tv);
}

int o2d_hip_points_on_points_tree_d(const o2d_env* env, const o2d_tree_opts* opt, int64_t ns,
                                    const double* sx, const double* sy, const double* sstr,
                                    const double* srad, int64_t nt, const double* tx,
                                    const double* ty, const double* trad, double* tu,
                                    double* tv) {
  clear_error();
  O2D_TRY(check_tree(env, opt, ns, sx, sy, sstr, srad, nt, tx, ty, tu, tv));
  if (ns == 0 || nt == 0) return 0;
  Device* d;
  O2D_TRY(device(env->device, &d));
  hipStream_t s = stream_of(env, d);
  double *dsx, *dsy, *dss, *dsr, *dtx, *dty, *dtr = nullptr, *du, *dv;
  O2D_TRY(upload<double>(d, kSrcX, sx, ns, s, &dsx));
  O2D_TRY(upload<double>(d, kSrcY, sy, ns, s, &dsy));
  O2D_TRY(upload<double>(d, kSrcS, sstr, ns, s, &dss));
  O2D_TRY(upload<double>(d, kSrcR, srad, ns, s, &dsr));
  O2D_TRY(upload<double>(d, kTrgX, tx, nt, s, &dtx));
  O2D_TRY(upload<double>(d, kTrgY, ty, nt, s, &dty));
  if (trad) O2D_TRY(upload<double>(d, kTrgR, trad, nt, s, &dtr));
  O2D_TRY(upload<double>(d, kOutU, tu, nt, s, &du));
  O2D_TRY(upload<double>(d, kOutV, tv, nt, s, &dv));
  O2D_TRY(tree_dev(d, s, env->core, opt->theta, opt->order, opt->core_tol, ns, dsx, dsy, dss,
                   dsr, nt, dtx, dty, dtr, du, dv));
  O2D_TRY(download<double>(tu, du, nt, s));
  O2D_TRY(download<double>(tv, dv, nt, s));
  O2D_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // extern "C"
