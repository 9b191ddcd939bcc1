// reflect.hip - particle/panel proximity on the device (src/Reflect.h): the distance from each
// target to the nearest panel (get_nearest_distances, :792-826), reflection of targets that lie
// behind their nearest panel (reflect_panp2, :167-311) and the near-wall clearing pass
// (clear_inner_panp2, :487-727; method 1 pushes targets out to the cutoff layer, method 0 trims
// strength with the cut tables of :333-432).
//
// Every target takes the panels in index order, as the reference does, because the set of
// nearest elements depends on that order: a distance below nextafter(min, 0) replaces the set,
// one below nextafter(min, +max) joins it. Groups of 64 targets (in Morton order) split the
// panels over four waves that keep exact summaries, and skip tiles of panels that lie beyond a
// bound on every group target's nearest distance (both below, with the argument that neither
// changes a result). All arithmetic is IEEE double with no fused multiply-add, in the reference's
// operation order, so results are bit-identical to a plain C loop. Two shortcuts keep the
// common case free of divisions without changing any result:
//   * t = dot/|b|^2 lies in (0,1) exactly when 0 < dot < |b|^2, unless the quotient could
//     underflow (dot < |b|^2 * 2^-1000), where the division is done;
//   * a panel-interior distance c^2/|b|^2 is divided exactly only when c^2 * (1/|b|^2) does not
//     already exceed the current acceptance bound by a margin far above its rounding error.
// Node normals are the sums of the normals of the panels that share the node, added in panel
// order (end 0 before end 1), from a stable sort of the panel ends by node.
#include <hipcub/hipcub.hpp>

#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>

#include "nbody.h"
#include "o2d_runtime.h"

#pragma clang fp contract(off)

namespace o2d {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxPasses = 64;    // clearing passes per target before it counts as unsettled
constexpr int kCutCap = 64;       // rows of the cut table
enum Mode { kNearest = 0, kReflect = 1, kClear = 2 };

// one panel as the scan reads it
struct PanelRec {
  double x0, y0, x1, y1;        // end nodes
  double bx, by, blensq, rbl;   // segment vector, |b|^2 and 1/|b|^2 (0 when not finite)
  double pnx, pny;              // panel normal
  double n0x, n0y, n1x, n1y;    // node normals at the two ends
};
constexpr int kRecDoubles = sizeof(PanelRec) / sizeof(double);
static_assert(sizeof(PanelRec) == 14 * sizeof(double), "PanelRec layout");

__device__ inline double next_down(double m) {  // nextafter(m, 0) for m >= 0
  return m > 0.0 ? __longlong_as_double(__double_as_longlong(m) - 1) : m;
}
__device__ inline double next_up(double m) {    // nextafter(m, DBL_MAX) for 0 <= m <= DBL_MAX
  return m < DBL_MAX ? __longlong_as_double(__double_as_longlong(m) + 1) : m;
}

// key/value pairs (node, panel end) for the node-normal sort; out-of-range nodes are clamped
// here and counted by panel_rec_kernel
__global__ __launch_bounds__(kBlock) void end_keys_kernel(int64_t nends, int64_t nnodes,
                                                          const uint32_t* __restrict__ idx,
                                                          uint32_t* __restrict__ key,
                                                          uint32_t* __restrict__ val) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= nends) return;
  const uint32_t k = idx[e];
  key[e] = (int64_t)k < nnodes ? k : (uint32_t)(nnodes - 1);
  val[e] = (uint32_t)e;
}

// node normal = sum over the node's panel ends in end order; the first end of each run sums it
__global__ __launch_bounds__(kBlock) void node_normal_kernel(int64_t nends,
                                                             const uint32_t* __restrict__ key,
                                                             const uint32_t* __restrict__ val,
                                                             const double* __restrict__ pnx,
                                                             const double* __restrict__ pny,
                                                             double* __restrict__ nnx,
                                                             double* __restrict__ nny) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= nends) return;
  const uint32_t k = key[p];
  if (p > 0 && key[p - 1] == k) return;
  double sx = 0.0, sy = 0.0;
  for (int64_t q = p; q < nends && key[q] == k; ++q) {
    const uint32_t j = val[q] >> 1;
    sx = sx + pnx[j];
    sy = sy + pny[j];
  }
  nnx[k] = sx;
  nny[k] = sy;
}

__global__ void zero_counter_kernel(int* c) {
  if (threadIdx.x == 0) c[0] = 0;
}

// panel records; counts panels with an out-of-range node or zero length into *bad
__global__ __launch_bounds__(kBlock) void panel_rec_kernel(
    int64_t np, int64_t nnodes, const double* __restrict__ nx, const double* __restrict__ ny,
    const uint32_t* __restrict__ idx, const double* __restrict__ pnx,
    const double* __restrict__ pny, const double* __restrict__ nnx,
    const double* __restrict__ nny, PanelRec* __restrict__ rec, int* __restrict__ bad) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= np) return;
  uint32_t a = idx[2 * j], b = idx[2 * j + 1];
  bool isbad = false;
  if ((int64_t)a >= nnodes) { a = (uint32_t)(nnodes - 1); isbad = true; }
  if ((int64_t)b >= nnodes) { b = (uint32_t)(nnodes - 1); isbad = true; }
  PanelRec r;
  r.x0 = nx[a];
  r.y0 = ny[a];
  r.x1 = nx[b];
  r.y1 = ny[b];
  r.bx = r.x1 - r.x0;
  r.by = r.y1 - r.y0;
  r.blensq = r.bx * r.bx + r.by * r.by;
  const double rb = 1.0 / r.blensq;
  r.rbl = (rb < DBL_MAX) ? rb : 0.0;  // 0 disables the division shortcut
  if (!(r.blensq != 0.0)) isbad = true;
  r.pnx = pnx ? pnx[j] : 0.0;
  r.pny = pny ? pny[j] : 0.0;
  r.n0x = nnx ? nnx[a] : 0.0;
  r.n0y = nny ? nny[a] : 0.0;
  r.n1x = nnx ? nnx[b] : 0.0;
  r.n1y = nny ? nny[b] : 0.0;
  rec[j] = r;
  if (isbad) atomicAdd(bad, 1);
}

struct CutTable {
  int rows;
  const double* pos;
  const double* sm;
  const double* ds;
};

struct ClearParams {
  int method;
  int max_passes;
  double cutoff;   // cutoff_mult * ips
  double ips;
  double push_tol; // -0.001 * ips
  CutTable ct;
};

// get_cut_entry (src/Reflect.h:400-432)
__device__ void cut_entry(const CutTable& ct, double p, double* smult, double* dshift) {
  *smult = 1.0;
  *dshift = 0.0;
  if (p < ct.pos[0]) {
    *smult = 0.0;
    *dshift = 1.0 - p;
  } else if (p > ct.pos[ct.rows - 1]) {
    *smult = 1.0;
    *dshift = 0.0;
  } else {
    for (int i = 1; i < ct.rows; ++i) {
      if (p < ct.pos[i]) {
        const double frac = (p - ct.pos[i - 1]) / (ct.pos[i] - ct.pos[i - 1]);
        *smult = frac * ct.sm[i] + (1.0 - frac) * ct.sm[i - 1];
        *dshift = frac * ct.ds[i] + (1.0 - frac) * ct.ds[i - 1];
        break;
      }
    }
  }
}

// running state of one target's scan
struct Scan {
  double m, lo, hi, thr;   // current minimum, nextafter(m, 0), nextafter(m, max), filter bound
  int64_t nh;              // hits in the current set
  double snx, sny, scx, scy;
  double h0d, h0nx, h0ny, h0cx, h0cy;  // first hit: distance, normal, contact point

  __device__ void start(double m0) {
    set_min(m0);
    nh = 0;
    snx = sny = scx = scy = 0.0;
    h0d = h0nx = h0ny = h0cx = h0cy = 0.0;
  }
  __device__ void set_min(double v) {
    m = v;
    lo = next_down(v);
    hi = next_up(v);
    thr = hi * (1.0 + 0x1p-40) + 0x1p-1000;
  }
};

template <int MODE>
__device__ __forceinline__ void scan_panel(const PanelRec& p, double x, double y, Scan& st) {
  const double ax = x - p.x0, ay = y - p.y0;
  const double dot = ax * p.bx + ay * p.by;
  bool in = (dot > 0.0) && (dot < p.blensq);
  if (in && dot < p.blensq * 0x1p-1000) {
    const double t = dot / p.blensq;
    in = (t > 0.0) && (t < 1.0);
  }
  double d;
  int kind;  // 0 panel, 1 node 0, 2 node 1
  if (in) {
    const double c = ay * p.bx - ax * p.by;
    const double sq = c * c;
    if (sq * p.rbl > st.thr) return;  // cannot reach the acceptance bound
    d = sq / p.blensq;
    kind = 0;
  } else {
    const double rij2 = ax * ax + ay * ay;
    const double ex = x - p.x1, ey = y - p.y1;
    const double rij12 = ex * ex + ey * ey;
    if (rij2 < rij12) {
      d = rij2;
      kind = 1;
    } else {
      d = rij12;
      kind = 2;
    }
  }
  if (MODE == kNearest) {
    if (d < st.lo) st.set_min(d);
    return;
  }
  if (d < st.lo) {
    st.set_min(d);
    st.nh = 0;
    st.snx = st.sny = st.scx = st.scy = 0.0;
  } else if (!(d < st.hi)) {
    return;
  }
  double hnx, hny, hcx, hcy;
  if (kind == 0) {
    const double t = dot / p.blensq;
    hnx = p.pnx;
    hny = p.pny;
    hcx = (1.0 - t) * p.x0 + t * p.x1;
    hcy = (1.0 - t) * p.y0 + t * p.y1;
  } else if (kind == 1) {
    hnx = p.n0x;
    hny = p.n0y;
    hcx = p.x0;
    hcy = p.y0;
  } else {
    hnx = p.n1x;
    hny = p.n1y;
    hcx = p.x1;
    hcy = p.y1;
  }
  if (st.nh == 0) {
    st.h0d = d;
    st.h0nx = hnx;
    st.h0ny = hny;
    st.h0cx = hcx;
    st.h0cy = hcy;
  }
  st.snx = st.snx + hnx;
  st.sny = st.sny + hny;
  st.scx = st.scx + hcx;
  st.scy = st.scy + hcy;
  st.nh += 1;
}

// a block's (count0, count1, value) sums
struct Partial {
  long long c0, c1;
  double v;
  double pad;
};

// fixed-order block sums of (count0, count1, value) into partial[blockIdx.x] when only wave 0
// holds non-zero values: a tree over wave 0's lanes (lane t adds lane t + w for w = 32, 16, ...,
// 1), in registers. It pairs the values as a tree over all the block's threads would, since the
// other threads' zeros leave every partial sum of that tree unchanged (up to the sign of a
// zero sum).
__device__ void wave0_partial(long long c0, long long c1, double v, Partial* out) {
  if (threadIdx.x >= 64) return;
  for (int w = 32; w > 0; w >>= 1) {
    c0 += __shfl_down(c0, w, 64);
    c1 += __shfl_down(c1, w, 64);
    v = v + __shfl_down(v, w, 64);
  }
  if (threadIdx.x == 0) out[blockIdx.x] = Partial{c0, c1, v, 0.0};
}

// ---- the split scan. A workgroup takes kGroupTargets targets (one per lane) and its kWaves waves
// split the panels: wave w scans the w-th of kWaves consecutive ranges, staged through its own
// LDS tile. The sequential scan is a state machine whose result depends on panel order, so each
// wave keeps an exact summary of its share instead. Let M be the smallest distance over all
// panels and T > M a value at which no panel lies. Every panel before the first one below T is
// at least next_up(T), so that first one replaces whatever set the scan held; from then on the
// bound is below T and no panel at or above T can replace or join. So the sequential scan ends
// exactly as a scan of the panels below T alone, in index order, ends. The summary keeps what
// that needs for T = next_up(M) or T = next_up(next_up(M)): the wave's smallest distance m, its
// panels at m and at next_up(m) (up to kHits of each, in index order) and whether one lies at
// next_up(next_up(m)). T is next_up(M) when no panel lies there, else next_up(next_up(M)). The
// waves' lists of panels below T are merged by index and replayed with the sequential code. A
// target with panels at all three of M, next_up(M) and next_up(next_up(M)), or more than kHits
// panels of one wave below T, is rescanned sequentially over all panels.
// Results are bit-identical to the sequential scan either way.
constexpr int kWaves = 4;
constexpr int kNearBlock = 64 * kWaves;
constexpr int kGroupTargets = 64;
constexpr int kHits = 4;
constexpr int kWaveTile = 32;                           // panels per wave tile
constexpr int kTileLoads = kWaveTile * kRecDoubles / 64;  // doubles per lane per tile
static_assert(kTileLoads * 64 == kWaveTile * kRecDoubles, "whole tile loads");

// LDS writes of this wave visible to its own later reads, and no compiler reordering across
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Distances are non-negative doubles, so their bit patterns order them as integers and
// next_up(d) is the pattern plus one: one 64-bit difference of patterns classifies a distance
// against the running minimum.
struct Chunk {
  long long bm;  // bit pattern of the smallest distance m
  double thr;    // a bound above which an interior distance cannot matter (above m + 2 ulps)
  int n0, n1;    // panels at m and at next_up(m)
  int f2;        // a panel at next_up(next_up(m)) was seen
  int sel;       // which hit list holds the panels at m
  int k0, k1;    // the first panel at m and at next_up(m); later ones go to the LDS lists
};

__device__ __forceinline__ void chunk_start(Chunk& c, double m0) {
  c.bm = __double_as_longlong(m0);
  c.thr = __longlong_as_double(c.bm + 2) * (1.0 + 0x1p-40) + 0x1p-1000;
  c.n0 = c.n1 = c.f2 = c.sel = 0;
  c.k0 = c.k1 = 0;
}

// one panel of a wave's share: the distance exactly as scan_panel computes it, then the
// summary. hits[(list * kHits + j) * kGroupTargets] is entry j (j >= 1) of hit list 0 or 1.
template <int MODE>
__device__ __forceinline__ void chunk_panel(const PanelRec& p, double x, double y, int k,
                                            Chunk& c, int* __restrict__ hits) {
  const double ax = x - p.x0, ay = y - p.y0;
  const double dot = ax * p.bx + ay * p.by;
  bool in = (dot > 0.0) && (dot < p.blensq);
  if (in && dot < p.blensq * 0x1p-1000) {
    const double t = dot / p.blensq;
    in = (t > 0.0) && (t < 1.0);
  }
  double d;
  if (in) {
    const double cc = ay * p.bx - ax * p.by;
    const double sq = cc * cc;
    if (sq * p.rbl > c.thr) return;  // more than two ulps above m: it cannot matter
    d = sq / p.blensq;
  } else {
    const double rij2 = ax * ax + ay * ay;
    const double ex = x - p.x1, ey = y - p.y1;
    const double rij12 = ex * ex + ey * ey;
    d = rij2 < rij12 ? rij2 : rij12;
  }
  const long long bd = __double_as_longlong(d);
  const long long dl = bd - c.bm;
  if (dl > 2) return;
  if (dl < 0) {
    c.bm = bd;
    c.thr = __longlong_as_double(bd + 2) * (1.0 + 0x1p-40) + 0x1p-1000;
    if (dl == -1) {  // the panels at the old minimum are now at next_up(m)
      c.f2 = c.n1 > 0;
      c.n1 = c.n0;
      c.k1 = c.k0;
      c.sel ^= 1;
    } else {
      c.f2 = dl == -2;
      c.n1 = 0;
    }
    c.n0 = 1;
    c.k0 = k;
  } else if (dl == 0) {
    if (c.n0 == 0) c.k0 = k;
    else if (c.n0 < kHits) hits[(c.sel * kHits + c.n0) * kGroupTargets] = k;
    ++c.n0;
  } else if (dl == 1) {
    if (c.n1 == 0) c.k1 = k;
    else if (c.n1 < kHits) hits[((c.sel ^ 1) * kHits + c.n1) * kGroupTargets] = k;
    ++c.n1;
  } else {
    c.f2 = 1;
  }
}

// ---- tile pruning. The scan takes panels in tiles of kWaveTile consecutive panels, each with
// the bounding box of its end nodes. Before a pass every active target gets an upper bound U
// on its nearest distance (squared): the distance to end node 0 of every kProbeStride-th panel,
// computed as the scan computes a node distance, which is at least the distance the scan
// computes for that panel up to a few ulps. A tile whose box lies farther (squared) from every
// active target than next_up(next_up(U)) (1 + 2^-40) of that target holds no panel within two
// ulps of any of those targets' smallest distance M (the relative margin covers the rounding of
// both bounds); such panels never join a nearest set and never enter the split scan's lists,
// so the tile is skipped. The box of the group's targets and the largest U sift the tiles
// first, one tile per lane. The kept tiles are dealt to the waves in turn, so each
// wave's panels still ascend, as the replay of its lists needs.
constexpr int kProbeStride = 16;

// box[4 t ...] = (x min, x max, y min, y max) of the end nodes of tile t. fmin/fmax leave NaN
// coordinates out, which is safe: the scan's distance to a panel with a NaN node is NaN or
// the distance to its other node, which the box holds. One thread per
// panel; a tile is 32 consecutive threads of a wave, reduced with shuffles.
static_assert(kWaveTile == 32, "a tile is half a wave");
__global__ __launch_bounds__(kBlock) void tile_box_kernel(int64_t np,
                                                          const PanelRec* __restrict__ pan,
                                                          int64_t ntile, double* __restrict__ box) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;  // k < 32 ntile (the grid)
  double x0 = DBL_MAX, x1 = -DBL_MAX, y0 = DBL_MAX, y1 = -DBL_MAX;
  if (k < np) {
    const PanelRec& p = pan[k];
    x0 = fmin(p.x0, p.x1);
    x1 = fmax(p.x0, p.x1);
    y0 = fmin(p.y0, p.y1);
    y1 = fmax(p.y0, p.y1);
  }
  for (int o = 16; o > 0; o >>= 1) {
    x0 = fmin(x0, __shfl_xor(x0, o, 64));
    x1 = fmax(x1, __shfl_xor(x1, o, 64));
    y0 = fmin(y0, __shfl_xor(y0, o, 64));
    y1 = fmax(y1, __shfl_xor(y1, o, 64));
  }
  const int64_t t = k / kWaveTile;
  if (k % kWaveTile == 0 && t < ntile) {
    box[4 * t] = x0;
    box[4 * t + 1] = x1;
    box[4 * t + 2] = y0;
    box[4 * t + 3] = y1;
  }
}

__device__ __forceinline__ double wave_min_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_max_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

template <int MODE>
__global__ __launch_bounds__(kNearBlock, 5) void near_kernel(int64_t np, const PanelRec* __restrict__ pan,
                                                      int64_t nt, double* __restrict__ tx,
                                                      double* __restrict__ ty,
                                                      double* __restrict__ ts,
                                                      const double* __restrict__ tr,
                                                      double* __restrict__ dist, ClearParams cp,
                                                      Partial* __restrict__ part,
                                                      const uint32_t* __restrict__ torder,
                                                      const double* __restrict__ tbox,
                                                      int64_t ntile) {
  __shared__ double wtile[kWaves][kWaveTile * kRecDoubles];  // each wave's panel tile
  __shared__ double gx[kGroupTargets], gy[kGroupTargets];   // the targets of this pass
  __shared__ int gact[kGroupTargets];
  __shared__ long long cm[kWaves][kGroupTargets];            // the waves' summaries
  __shared__ int c0n[kWaves][kGroupTargets], c1n[kWaves][kGroupTargets];
  __shared__ int cfs[kWaves][kGroupTargets];                 // f2 | sel << 1
  __shared__ int ch[kWaves][2][kHits][kGroupTargets];
  __shared__ double ub[kWaves][kGroupTargets];               // the waves' probe bounds
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // the group's targets are entries gi of the order torder (input order when null)
  const int64_t gi = (int64_t)blockIdx.x * kGroupTargets + lane;
  double* mytile = wtile[wave];
  const PanelRec* lp = reinterpret_cast<const PanelRec*>(mytile);
  // the target's state lives in wave 0
  const bool valid = wave == 0 && gi < nt;
  const int64_t i = valid ? (torder ? (int64_t)torder[gi] : gi) : 0;
  double x = valid ? tx[i] : 0.0, y = valid ? ty[i] : 0.0;
  double s = (valid && ts && tr) ? ts[i] : 0.0;
  const double rad = (valid && tr) ? tr[i] : cp.ips;
  bool untested = valid;
  bool moved = false;
  int passes = 0;
  long long c0 = 0;
  double removed = 0.0;
  const int maxp = (MODE == kClear) ? cp.max_passes : 1;
  const double m0 = (MODE == kClear) ? 0.1 * DBL_MAX : DBL_MAX;
  for (;;) {
    const bool act = untested && passes < maxp;
    if (wave == 0) {
      gx[lane] = x;
      gy[lane] = y;
      gact[lane] = act;
    }
    if (!__syncthreads_or(act)) break;
    const double xw = gx[lane], yw = gy[lane];
    const bool aw = gact[lane] != 0;
    Chunk c;
    chunk_start(c, m0);
    int* hits = &ch[wave][0][0][lane];
    // the bound U from this wave's share of the probe panels, and the box of the active targets
    double u = DBL_MAX;
    if (aw) {
      for (int64_t k = (int64_t)wave * kProbeStride; k < np; k += kWaves * kProbeStride) {
        const double ax = xw - pan[k].x0, ay = yw - pan[k].y0;
        const double r2 = ax * ax + ay * ay;
        u = r2 < u ? r2 : u;
      }
    }
    ub[wave][lane] = u;
    const double bx0 = wave_min_d(aw ? xw : DBL_MAX), bx1 = wave_max_d(aw ? xw : -DBL_MAX);
    const double by0 = wave_min_d(aw ? yw : DBL_MAX), by1 = wave_max_d(aw ? yw : -DBL_MAX);
    __syncthreads();
#pragma unroll
    for (int w = 0; w < kWaves; ++w) u = ub[w][lane] < u ? ub[w][lane] : u;
    const double thr_l = next_up(next_up(u)) * (1.0 + 0x1p-40) + 0x1p-1000;  // this target's
    const double umax = wave_max_d(aw ? u : 0.0);
    const double thr_u = next_up(next_up(umax)) * (1.0 + 0x1p-40) + 0x1p-1000;  // the group's
    // the kept tiles of 64 at base (called with the whole wave active): first against the box
    // of the active targets and the group's bound, one tile per lane; then each of those
    // against every active target and its own bound
    auto chunk_mask = [&](int64_t base) -> unsigned long long {
      const int64_t t = base + lane;
      bool keep = false;
      if (t < ntile) {
        const double* b = tbox + 4 * t;
        const double dx = fmax(0.0, fmax(b[0] - bx1, bx0 - b[1]));
        const double dy = fmax(0.0, fmax(b[2] - by1, by0 - b[3]));
        keep = !((dx * dx + dy * dy) * (1.0 - 0x1p-40) > thr_u);
      }
      unsigned long long out = 0ull;
      for (unsigned long long r = __ballot(keep); r != 0ull; r &= r - 1ull) {
        const int j = __builtin_ctzll(r);
        const double* b = tbox + 4 * (base + j);
        const double dx = fmax(0.0, fmax(b[0] - xw, xw - b[1]));
        const double dy = fmax(0.0, fmax(b[2] - yw, yw - b[3]));
        const bool need = aw && !((dx * dx + dy * dy) * (1.0 - 0x1p-40) > thr_l);
        if (__ballot(need) != 0ull) out |= 1ull << j;
      }
      return out;
    };
    // this wave's next kept tile (kept tiles in index order, dealt to the waves in turn), or -1
    int64_t cbase = 0, rank = 0;
    unsigned long long cmask = ntile > 0 ? chunk_mask(0) : 0ull;
    auto next_tile = [&]() -> int64_t {
      for (;;) {
        while (cmask == 0ull) {
          cbase += 64;
          if (cbase >= ntile) return -1;
          cmask = chunk_mask(cbase);
        }
        const int b = __builtin_ctzll(cmask);
        cmask &= cmask - 1ull;
        const int64_t r = rank++;
        if ((int)(r % kWaves) == wave) return cbase + b;
      }
    };
    const int64_t tfirst = next_tile();
    if (__ballot(aw) != 0 && tfirst >= 0) {
      // the wave's tiles through its own LDS tile; the next tile's loads are in flight while
      // the current one is scanned
      const double* src = reinterpret_cast<const double*>(pan);
      double reg[kTileLoads];
      int64_t t0 = tfirst * kWaveTile;
      int n = (int)(np - t0 < kWaveTile ? np - t0 : kWaveTile);
#pragma unroll
      for (int j = 0; j < kTileLoads; ++j) {
        const int e = lane + 64 * j;
        reg[j] = e < n * kRecDoubles ? src[t0 * kRecDoubles + e] : 0.0;
      }
      for (;;) {
        wave_sync();  // the previous tile's reads are done
#pragma unroll
        for (int j = 0; j < kTileLoads; ++j) mytile[lane + 64 * j] = reg[j];
        wave_sync();
        const int64_t tn = next_tile();
        const int64_t t1 = tn >= 0 ? tn * kWaveTile : np;
        const int n1 = tn >= 0 ? (int)(np - t1 < kWaveTile ? np - t1 : kWaveTile) : 0;
        if (n1 > 0) {
#pragma unroll
          for (int j = 0; j < kTileLoads; ++j) {
            const int e = lane + 64 * j;
            reg[j] = e < n1 * kRecDoubles ? src[t1 * kRecDoubles + e] : 0.0;
          }
        }
        if (aw) {
#pragma unroll 4
          for (int k = 0; k < n; ++k) chunk_panel<MODE>(lp[k], xw, yw, (int)t0 + k, c, hits);
        }
        if (n1 <= 0) break;
        t0 = t1;
        n = n1;
      }
    }
    if (c.n0 > 0) hits[(c.sel * kHits) * kGroupTargets] = c.k0;
    if (c.n1 > 0) hits[((c.sel ^ 1) * kHits) * kGroupTargets] = c.k1;
    cm[wave][lane] = c.bm;
    c0n[wave][lane] = c.n0;
    c1n[wave][lane] = c.n1;
    cfs[wave][lane] = c.f2 | (c.sel << 1);
    __syncthreads();
    if (!act) continue;  // only wave 0 lanes with an active target go on
    ++passes;
    long long M = cm[0][lane];  // bit patterns: M, M + 1 = next_up(M), M + 2
#pragma unroll
    for (int w = 1; w < kWaves; ++w) M = cm[w][lane] < M ? cm[w][lane] : M;
    const long long M1 = M + 1, M2 = M + 2;
    // the panels below T = M2 are the waves' lists at M and M1; a panel at M2 or an overflowed
    // list sends the target to the sequential rescan (a kNearest scan ends at the distance of
    // the first panel below T, which every list keeps)
    // T = M1 when no panel lies at M1 (the lists at M alone), else T = M2 when none lies at M2
    // (the lists at M and M1); T must stay below the starting bound m0
    bool at1 = false, at2 = false;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      const long long mw = cm[w][lane];
      const int n1 = c1n[w][lane], f2 = cfs[w][lane] & 1;
      if (mw == M) {
        at1 = at1 || n1 > 0;
        at2 = at2 || f2;
      } else if (mw == M1) {
        at1 = true;
        at2 = at2 || n1 > 0;
      } else if (mw == M2) {
        at2 = true;
      }
    }
    const bool two = at1;  // T = M2
    bool exact = (two ? !at2 && M2 < __double_as_longlong(m0) : M1 < __double_as_longlong(m0));
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {  // c0n, c1n become the lengths of lists 0 and 1 below T
      const long long mw = cm[w][lane];
      const int n0 = c0n[w][lane], n1 = c1n[w][lane];
      const int sel = cfs[w][lane] >> 1;
      int l0 = 0, l1 = 0;  // entries of the lists at m and at next_up(m) that lie below T
      if (mw == M) {
        if (MODE != kNearest && (n0 > kHits || (two && n1 > kHits))) exact = false;
        l0 = n0 < kHits ? n0 : kHits;
        l1 = two ? (n1 < kHits ? n1 : kHits) : 0;
      } else if (two && mw == M1) {
        if (MODE != kNearest && n0 > kHits) exact = false;
        l0 = n0 < kHits ? n0 : kHits;
      }
      c0n[w][lane] = sel ? l1 : l0;
      c1n[w][lane] = sel ? l0 : l1;
    }
    Scan st;
    st.start(m0);
    if (exact) {
      // the listed panels in index order (each list ascends), replayed with the sequential code
      int last = -1;
      for (;;) {
        int bk = INT_MAX;
        for (int w = 0; w < kWaves; ++w) {
          const int la = c0n[w][lane], lb = c1n[w][lane];
          for (int j = 0; j < la; ++j) {
            const int k = ch[w][0][j][lane];
            if (k > last && k < bk) bk = k;
          }
          for (int j = 0; j < lb; ++j) {
            const int k = ch[w][1][j][lane];
            if (k > last && k < bk) bk = k;
          }
        }
        if (bk == INT_MAX) break;
        scan_panel<MODE>(pan[bk], x, y, st);
        last = bk;
        if (MODE == kNearest) break;  // the first panel below T fixes the distance
      }
    } else {
      for (int64_t k = 0; k < np; ++k) scan_panel<MODE>(pan[k], x, y, st);
    }
    if (MODE == kNearest) {
      dist[i] = sqrt(st.m);
      untested = false;
      continue;
    }
    if (st.nh == 0) {  // no nearest element (reference: assert / retry)
      if (MODE == kReflect) untested = false;
      continue;
    }
    double normx = st.snx, normy = st.sny, cpx = st.scx, cpy = st.scy;
    if (MODE == kReflect && normx * normx + normy * normy == 0.0) {
      normx = normx + st.h0nx;
      normy = normy + st.h0ny;
      cpx = st.h0cx;
      cpy = st.h0cy;
    }
    const double normilen = 1.0 / sqrt(normx * normx + normy * normy);
    normx = normx * normilen;
    normy = normy * normilen;
    cpx = cpx / (double)st.nh;
    cpy = cpy / (double)st.nh;
    if (MODE == kReflect) {
      const double dotp = normx * (x - cpx) + normy * (y - cpy);
      if (dotp < 0.0) {
        const double dd = sqrt(st.h0d);
        x = cpx + dd * normx;
        y = cpy + dd * normy;
        moved = true;
        ++c0;
      }
      untested = false;
      continue;
    }
    const double dotp = normx * (x - cpx) + normy * (y - cpy) - cp.cutoff;
    if (cp.method == 0) {
      if (dotp < rad) {
        if (tr) {
          double sm, dsh;
          cut_entry(cp.ct, dotp / rad, &sm, &dsh);
          removed = removed + s * (1.0 - sm);
          s = s * sm;
          x = x + dsh * rad * normx;
          y = y + dsh * rad * normy;
          moved = true;
        }
        ++c0;
      } else {
        untested = false;
      }
    } else {
      if (dotp < cp.push_tol) {
        x = x - dotp * normx;
        y = y - dotp * normy;
        moved = true;
        ++c0;
      } else {
        untested = false;
      }
    }
  }
  if (MODE == kNearest) return;
  if (valid && moved) {
    tx[i] = x;
    ty[i] = y;
    if (MODE == kClear && ts && tr) ts[i] = s;
  }
  wave0_partial(c0, (valid && untested) ? 1 : 0, removed, part);  // targets live in wave 0
}

// fixed-order sum of the block partials: counts[0..1], *value; counts[2] = bad panels
__global__ __launch_bounds__(kBlock) void finish_kernel(int nblk, const Partial* __restrict__ part,
                                                        const int* __restrict__ bad,
                                                        double* __restrict__ value,
                                                        int64_t* __restrict__ counts) {
  __shared__ long long s0[kBlock], s1[kBlock];
  __shared__ double sv[kBlock];
  const int t = threadIdx.x;
  long long a = 0, b = 0;
  double v = 0.0;
  for (int k = t; k < nblk; k += kBlock) {
    a += part[k].c0;
    b += part[k].c1;
    v = v + part[k].v;
  }
  s0[t] = a;
  s1[t] = b;
  sv[t] = v;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (t < w) {
      s0[t] += s0[t + w];
      s1[t] += s1[t + w];
      sv[t] = sv[t] + sv[t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    if (counts) {
      counts[0] = s0[0];
      counts[1] = s1[0];
      counts[2] = bad ? bad[0] : 0;
    }
    if (value) value[0] = sv[0];
  }
}

inline unsigned blocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }
inline unsigned groups(int64_t n) {
  return (unsigned)((n + kGroupTargets - 1) / kGroupTargets);
}

// init_cut_tables<double>(0.1) (src/Reflect.h:333-394), computed once on the host in IEEE double
// without fused operations, in the reference's order
struct HostCutTable {
  int rows = 0;
  double v[3 * kCutCap];
  HostCutTable() {
    const double dxin = 0.1, max_rad = 3.0;
    const int nx = (int)(0.5 + max_rad / dxin);
    const double dx = max_rad / (double)(float)nx;
    double* pos = v;
    double* sm = v + kCutCap;
    double* ds = v + 2 * kCutCap;
    int r = 0;
    pos[r] = (double)(-nx - 0.5) * dx;
    sm[r] = 0.0;
    ds[r] = 0.0;
    ++r;
    double twgt = 0.0, tmom = 0.0;
    for (int i = -nx; i < nx + 1 && r < kCutCap - 1; ++i) {
      const double xi = (double)i * dx;
      const double distxs = xi * xi;
      double rwgt = 0.0;
      for (int j = -nx; j < nx + 1; ++j) {
        const double yj = (double)j * dx;
        const double distxys = distxs + yj * yj;
        rwgt += std::exp(-distxys);
      }
      rwgt *= dx * dx;
      rwgt *= 1.0 / M_PI;
      twgt += rwgt;
      tmom += ((double)i * dx) * rwgt;
      pos[r] = (double)(i + 0.5) * dx;
      sm[r] = twgt;
      ds[r] = -tmom / twgt;
      ++r;
    }
    pos[r] = (double)(nx + 1.5) * dx;
    sm[r] = 1.0;
    ds[r] = 0.0;
    rows = r + 1;
  }
};

const HostCutTable& host_cut_table() {
  static const HostCutTable t;
  return t;
}

}  // namespace

// Shared driver of the three operations (device pointers, enqueued on s). dist: nearest
// distances (mode kNearest); value: removed circulation (kClear); counts[3]: moved target
// events, unsettled targets, bad panels.
int near_dev(Device* d, hipStream_t s, int mode, int method, double cutoff_mult, double ips,
             int64_t nnodes, const double* nx, const double* ny, int64_t np,
             const uint32_t* idx, const double* pnx, const double* pny, int64_t nt, double* tx,
             double* ty, double* ts, const double* tr, double* dist, double* value,
             int64_t* counts) {
  const unsigned nblk = groups(nt);
  void* pb;
  O2D_TRY(slot(d, kNearPart, sizeof(Partial) * (nblk > 0 ? nblk : 1) + 64, s, &pb));
  Partial* part = static_cast<Partial*>(pb);
  int* bad = reinterpret_cast<int*>(part + (nblk > 0 ? nblk : 1));
  zero_counter_kernel<<<1, 64, 0, s>>>(bad);
  O2D_LAUNCH_CHECK();
  PanelRec* rec = nullptr;
  double* tbox = nullptr;
  int64_t ntile = 0;
  const uint32_t* torder = nullptr;
  if (np > 0) {
    double *nnx = nullptr, *nny = nullptr;
    if (mode != kNearest) {
      const int64_t nends = 2 * np;
      void *kA, *vA, *kB, *vB, *nn;
      O2D_TRY(slot(d, kKeysA, nends * sizeof(uint32_t), s, &kA));
      O2D_TRY(slot(d, kOrderA, nends * sizeof(uint32_t), s, &vA));
      O2D_TRY(slot(d, kKeysB, nends * sizeof(uint32_t), s, &kB));
      O2D_TRY(slot(d, kOrderB, nends * sizeof(uint32_t), s, &vB));
      O2D_TRY(slot(d, kNodeNorm, 2 * sizeof(double) * (size_t)nnodes, s, &nn));
      nnx = static_cast<double*>(nn);
      nny = nnx + nnodes;
      end_keys_kernel<<<blocks(nends), kBlock, 0, s>>>(nends, nnodes, idx,
                                                       static_cast<uint32_t*>(kA),
                                                       static_cast<uint32_t*>(vA));
      O2D_LAUNCH_CHECK();
      int bits = 1;
      while (bits < 32 && ((uint64_t)1 << bits) < (uint64_t)nnodes) ++bits;
      size_t tbytes = 0;
      O2D_HIP(hipcub::DeviceRadixSort::SortPairs(
          nullptr, tbytes, static_cast<uint32_t*>(kA), static_cast<uint32_t*>(kB),
          static_cast<uint32_t*>(vA), static_cast<uint32_t*>(vB), (int)nends, 0, bits, s));
      void* tmp;
      O2D_TRY(slot(d, kSortTemp, tbytes, s, &tmp));
      O2D_HIP(hipcub::DeviceRadixSort::SortPairs(
          tmp, tbytes, static_cast<uint32_t*>(kA), static_cast<uint32_t*>(kB),
          static_cast<uint32_t*>(vA), static_cast<uint32_t*>(vB), (int)nends, 0, bits, s));
      node_normal_kernel<<<blocks(nends), kBlock, 0, s>>>(
          nends, static_cast<uint32_t*>(kB), static_cast<uint32_t*>(vB), pnx, pny, nnx, nny);
      O2D_LAUNCH_CHECK();
    }
    void* rp;
    O2D_TRY(slot(d, kPanRec, sizeof(PanelRec) * (size_t)np, s, &rp));
    rec = static_cast<PanelRec*>(rp);
    panel_rec_kernel<<<blocks(np), kBlock, 0, s>>>(np, nnodes, nx, ny, idx,
                                                   mode != kNearest ? pnx : nullptr,
                                                   mode != kNearest ? pny : nullptr, nnx, nny,
                                                   rec, bad);
    O2D_LAUNCH_CHECK();
    ntile = (np + kWaveTile - 1) / kWaveTile;
    void* tb;
    O2D_TRY(slot(d, kNearTiles, sizeof(double) * 4 * (size_t)ntile, s, &tb));
    tbox = static_cast<double*>(tb);
    tile_box_kernel<<<blocks(ntile * kWaveTile), kBlock, 0, s>>>(np, rec, ntile, tbox);
    O2D_LAUNCH_CHECK();
    if (nt > kGroupTargets && ntile > 1) {
      // the targets in Morton order, so that a group of 64 is compact and its box excludes
      // most tiles (the order changes no target's result)
      O2D_TRY(morton_order16(d, s, nt, tx, ty, &torder));
    }
  }
  ClearParams cp{};
  cp.method = method;
  cp.max_passes = kMaxPasses;
  cp.cutoff = cutoff_mult * ips;
  cp.ips = ips;
  cp.push_tol = -0.001 * ips;
  if (mode == kClear && method == 0) {
    const HostCutTable& h = host_cut_table();
    void* ctp;
    O2D_TRY(slot(d, kCutTable, sizeof(h.v), s, &ctp));
    O2D_HIP(hipMemcpyAsync(ctp, h.v, sizeof(h.v), hipMemcpyHostToDevice, s));
    const double* dv = static_cast<const double*>(ctp);
    cp.ct = CutTable{h.rows, dv, dv + kCutCap, dv + 2 * kCutCap};
  }
  if (nt > 0) {
    // clear and reflect with no panels leave every target where it is
    const bool run = (mode == kNearest) || np > 0;
    if (run) {
      KernelTimer kt(O2D_KT_NEAR, s);
      if (mode == kNearest)
        near_kernel<kNearest><<<nblk, kNearBlock, 0, s>>>(np, rec, nt, tx, ty, ts, tr, dist, cp, part, torder, tbox, ntile);
      else if (mode == kReflect)
        near_kernel<kReflect><<<nblk, kNearBlock, 0, s>>>(np, rec, nt, tx, ty, ts, tr, dist, cp, part, torder, tbox, ntile);
      else
        near_kernel<kClear><<<nblk, kNearBlock, 0, s>>>(np, rec, nt, tx, ty, ts, tr, dist, cp, part, torder, tbox, ntile);
      kt.stop();
      O2D_LAUNCH_CHECK();
    }
    if (mode != kNearest && (value || counts)) {
      if (run) {
        finish_kernel<<<1, kBlock, 0, s>>>((int)nblk, part, bad, value, counts);
      } else {
        finish_kernel<<<1, kBlock, 0, s>>>(0, part, bad, value, counts);
      }
      O2D_LAUNCH_CHECK();
    }
  } else if (mode != kNearest && (value || counts)) {
    finish_kernel<<<1, kBlock, 0, s>>>(0, part, bad, value, counts);
    O2D_LAUNCH_CHECK();
  }
  return 0;
}

}  // namespace o2d
