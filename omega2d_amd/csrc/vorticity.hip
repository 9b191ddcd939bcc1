// vorticity.hip - Gaussian vorticity of particles at targets, points_affect_points_vorticity
// (src/InfluenceVort.h:47-126, the x86 loop :103-117):
//   tw_i += sum_j 2 s_j exp(-d_ij^2 / r_j^2) / r_j^2      (unscaled; finalize_vels adds 1/2pi)
// in fp64. Sources are staged through LDS as (x, y, 1/r^2, 2 s/r^2); each lane owns two
// targets. A pair with d^2/r^2 > 746 contributes exactly +-0 (exp underflows to 0 in double), so
// its exponential is skipped without changing the sum. Few targets split the source range into
// chunks whose fp64 partials are summed in chunk order. With many sources and targets both go
// in Morton order: a tile of 256 sources carries the box of its positions and its smallest
// 1/r^2, each group of 256 targets (and each wave's 64) has a compact box, and a tile whose
// every pair with a group lies beyond the underflow bound (by a margin far above the rounding
// of both tests) is skipped for that group; it adds only +-0 terms. The exponential is a
// table-driven exp(-x) on [0, 746). The Morton order changes the order of the fp64 additions, so sums move by
// rounding only. A NaN position or radius adds NaN, as in the reference (its tile and its
// targets' group are never skipped); a non-finite strength or a zero radius beyond the bound
// adds nothing, where the reference's inf*0 gives NaN.
#include <cfloat>

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
constexpr int64_t kPruneSources = 4096, kPruneTargets = 1024;  // from here the sort pays
// pruned launch: source tiles per chunk, and the cap on the chunk partials' bytes
#ifndef O2D_VORT_PRUNED_TILES
#define O2D_VORT_PRUNED_TILES 2
#endif
constexpr int64_t kPartialBytes = int64_t(1) << 30;

struct VSrc {
  double x, y, rr, c;  // position, 1/r^2, 2 s / r^2
};

// 2^(j/64), j = 0..63, each the double nearest to the exact value
__constant__ double kExp2Tab[64] = {
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0, 0x1.0874518759bc8p+0,
    0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0, 0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0,
    0x1.172b83c7d517bp+0, 0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0, 0x1.2d285a6e4030bp+0,
    0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0, 0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0,
    0x1.3dea64c123422p+0, 0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0, 0x1.56f4736b527dap+0,
    0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0, 0x1.6247eb03a5585p+0, 0x1.6623882552225p+0,
    0x1.6a09e667f3bcdp+0, 0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0, 0x1.868d99b4492edp+0,
    0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0, 0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0,
    0x1.9c49182a3f090p+0, 0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0, 0x1.bcc1e904bc1d2p+0,
    0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0, 0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0,
    0x1.d5818dcfba487p+0, 0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0, 0x1.fa7c1819e90d8p+0,
};

// exp(-x) for 0 <= x < 746 (NaN for NaN), by Tang's table-driven method (P. T. P. Tang,
// ACM TOMS 15(2), 1989): -x = n ln2/64 + r with n = rint(-64 x/ln2) and |r| <= ln2/128, taken
// as r = (-x - n L1) - n L2 with ln2/64 = L1 + L2 and L1 of 30 bits, so n L1 and the first
// difference are exact; then exp(-x) = 2^m 2^(j/64) (1 + q(r)), n = 64 m + j, with q the
// degree-6 Taylor polynomial of e^r - 1 (remainder below 2^-60 relative). tab holds kExp2Tab.
// Checked against long double expl on 40 million points of the range: within 1.01 ulp, and
// within 1 unit of 2^-1074 where the result is subnormal.
__device__ __forceinline__ double exp_neg(double x, const double* tab) {
  const double y = -x;
  const double n = rint(y * 0x1.71547652b82fep+6);
  double r = fma(-n, 0x1.62e42ffp-7, y);
  r = fma(-n, -0x1.718432a1b0e26p-41, r);
  const int ni = (int)n;
  const int j = ni & 63;
  const int m = ni >> 6;  // floor(n / 64): n = 64 m + j
  double q = fma(r, 1.0 / 720.0, 1.0 / 120.0);
  q = fma(r, q, 1.0 / 24.0);
  q = fma(r, q, 1.0 / 6.0);
  q = fma(r, q, 0.5);
  q = fma(r, q, 1.0);
  q *= r;
  const double t = tab[j];
  return ldexp(fma(t, q, t), m);
}

// the staged record of a source (as vort_kernel stages it)
__device__ __forceinline__ VSrc vsrc_of(double x, double y, double s, double r) {
  const double r2 = r * r;
  const double rr = 1.0 / r2;
  return VSrc{x, y, rr, 2.0 * s * rr};
}

// the sources in the order sorder as staged records, and per tile of kTile the box of their
// positions and their smallest 1/r^2: tbox[5 t ...] = (x min, x max, y min, y max, rr min).
// One workgroup per tile. A NaN position or radius makes the tile's box unbounded.
__global__ __launch_bounds__(kBlock) void vort_gather_kernel(int64_t ns, const double* __restrict__ sx,
                                                             const double* __restrict__ sy,
                                                             const double* __restrict__ ss,
                                                             const double* __restrict__ sr,
                                                             const uint32_t* __restrict__ sorder,
                                                             VSrc* __restrict__ vs,
                                                             double* __restrict__ tbox) {
  __shared__ double sm[5][kBlock];
  const int t = threadIdx.x;
  const int64_t g = (int64_t)blockIdx.x * kBlock + t;
  double x0 = DBL_MAX, x1 = -DBL_MAX, y0 = DBL_MAX, y1 = -DBL_MAX, r0 = DBL_MAX;
  if (g < ns) {
    const int64_t j = sorder ? (int64_t)sorder[g] : g;
    const VSrc q = vsrc_of(sx[j], sy[j], ss[j], sr[j]);
    vs[g] = q;
    x0 = x1 = q.x;
    y0 = y1 = q.y;
    r0 = q.rr;
    if (!(q.x == q.x) || !(q.y == q.y) || !(q.rr == q.rr)) {
      x0 = y0 = -INFINITY;
      x1 = y1 = INFINITY;
      r0 = 0.0;
    }
  }
  sm[0][t] = x0;
  sm[1][t] = x1;
  sm[2][t] = y0;
  sm[3][t] = y1;
  sm[4][t] = r0;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (t < w) {
      sm[0][t] = fmin(sm[0][t], sm[0][t + w]);
      sm[1][t] = fmax(sm[1][t], sm[1][t + w]);
      sm[2][t] = fmin(sm[2][t], sm[2][t + w]);
      sm[3][t] = fmax(sm[3][t], sm[3][t + w]);
      sm[4][t] = fmin(sm[4][t], sm[4][t + w]);
    }
    __syncthreads();
  }
  if (t < 5) tbox[5 * blockIdx.x + t] = sm[t][0];
}

// the box of each kBlock consecutive targets in the order torder (the first or the second
// targets of one vort_kernel workgroup's lanes): gbox[4 g ...] = (x min, x max, y min, y max).
// One workgroup per group; a NaN position makes the group's box unbounded.
__global__ __launch_bounds__(kBlock) void vort_target_box_kernel(int64_t nt,
                                                                 const double* __restrict__ tx,
                                                                 const double* __restrict__ ty,
                                                                 const uint32_t* __restrict__ torder,
                                                                 double* __restrict__ gbox) {
  __shared__ double sm[4][kBlock];
  const int t = threadIdx.x;
  const int64_t p = (int64_t)blockIdx.x * kBlock + t;
  double x0 = DBL_MAX, x1 = -DBL_MAX, y0 = DBL_MAX, y1 = -DBL_MAX;
  if (p < nt) {
    const int64_t i = torder[p];
    x0 = x1 = tx[i];
    y0 = y1 = ty[i];
    if (!(x0 == x0) || !(y0 == y0)) {
      x0 = y0 = -INFINITY;
      x1 = y1 = INFINITY;
    }
  }
  sm[0][t] = x0;
  sm[1][t] = x1;
  sm[2][t] = y0;
  sm[3][t] = y1;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (t < w) {
      sm[0][t] = fmin(sm[0][t], sm[0][t + w]);
      sm[1][t] = fmax(sm[1][t], sm[1][t + w]);
      sm[2][t] = fmin(sm[2][t], sm[2][t + w]);
      sm[3][t] = fmax(sm[3][t], sm[3][t + w]);
    }
    __syncthreads();
  }
  if (t < 4) gbox[4 * blockIdx.x + t] = sm[t][0];
}

// true unless every pair of source tile box b (with its smallest 1/r^2) and target group box g
// lies beyond the underflow bound, by a relative margin of 2^-40
__device__ __forceinline__ bool tile_near(const double* b, const double* g) {
  const double dx = fmax(0.0, fmax(b[0] - g[1], g[0] - b[1]));
  const double dy = fmax(0.0, fmax(b[2] - g[3], g[2] - b[3]));
  return !((dx * dx + dy * dy) * b[4] * (1.0 - 0x1p-40) > kUnderflow);
}

// vs/tbox null: the sources in input order from sx..sr; else the staged records in Morton
// order with their tile boxes, gbox the boxes of the target groups of kBlock (a workgroup holds
// two: its lanes' first and second targets) and torder the targets' order: the target at
// position p is torder[p] and chunk partials are kept by position (vort_reduce_kernel maps
// them). A tile is skipped for a group it cannot reach, for both groups at once when neither,
// and, after staging, by each wave for its own 64 first or second targets that it cannot reach;
// a workgroup whose chunk reaches neither group writes no partials and ends. The grid is one
// dimensional: block b takes (target workgroup, chunk) number (b * perm) % (twgs * nchunk), row
// major by chunk, so that the few (workgroup, chunk) pairs near the Morton diagonal that hold
// work are spread over the whole dispatch. flag[workgroup * nchunk8 + chunk] says whether its
// partials were written (nchunk8: the chunk count rounded up to 8; the padding is never read).
__global__ __launch_bounds__(kBlock) void vort_kernel(int64_t ns, const double* __restrict__ sx,
                                                      const double* __restrict__ sy,
                                                      const double* __restrict__ ss,
                                                      const double* __restrict__ sr, int64_t nt,
                                                      const double* __restrict__ tx,
                                                      const double* __restrict__ ty,
                                                      int64_t chunk, double* __restrict__ out,
                                                      uint8_t* __restrict__ flag, int64_t twgs,
                                                      int64_t nchunk8,
                                                      int64_t perm, int direct,
                                                      const VSrc* __restrict__ vs,
                                                      const double* __restrict__ tbox,
                                                      const double* __restrict__ gbox,
                                                      const uint32_t* __restrict__ torder) {
  static_assert(kTpl == 2, "a workgroup holds two target groups");
  __shared__ VSrc tile[kTile];
  __shared__ double e2t[64];  // kExp2Tab, visible after the first tile's barrier
  if (threadIdx.x < 64) e2t[threadIdx.x] = kExp2Tab[threadIdx.x];
  const int64_t cell = ((int64_t)blockIdx.x * perm) % ((int64_t)gridDim.x);
  const int64_t bx = cell % twgs, by = cell / twgs;
  const int64_t j0 = by * chunk;
  const int64_t j1 = (j0 + chunk < ns) ? j0 + chunk : ns;
  double px[kTpl], py[kTpl], acc[kTpl];
  int64_t ti[kTpl];
#pragma unroll
  for (int t = 0; t < kTpl; ++t) ti[t] = (bx * kTpl + t) * kBlock + threadIdx.x;
  // the workgroup's two target groups (the second may hold no target)
  const double* g0 = gbox ? gbox + 4 * (2 * bx) : nullptr;
  const double* g1 = (gbox && (2 * bx + 1) * kBlock < nt) ? g0 + 4 : nullptr;
  if (tbox) {  // uniform over the workgroup: chunks are whole tiles
    bool any = false;
    for (int64_t base = j0; base < j1 && !any; base += kTile) {
      const double* b = tbox + 5 * (base / kTile);
      any = tile_near(b, g0) || (g1 && tile_near(b, g1));
    }
    if (!any) {
      if (!direct && threadIdx.x == 0) flag[bx * nchunk8 + by] = 0;
      return;
    }
  }
  if (!direct && threadIdx.x == 0) flag[bx * nchunk8 + by] = 1;
#pragma unroll
  for (int t = 0; t < kTpl; ++t) {
    const bool ok = ti[t] < nt;
    const int64_t i = ok ? (torder ? (int64_t)torder[ti[t]] : ti[t]) : 0;
    px[t] = ok ? tx[i] : 0.0;
    py[t] = ok ? ty[i] : 0.0;
    acc[t] = 0.0;
  }
  // each wave's box of its 64 first and 64 second targets (a NaN position: unbounded)
  double wb[kTpl][4];
#pragma unroll
  for (int t = 0; t < kTpl; ++t) {
    const bool ok = ti[t] < nt;
    const bool bad = ok && (!(px[t] == px[t]) || !(py[t] == py[t]));
    double x0 = ok ? px[t] : DBL_MAX, x1 = ok ? px[t] : -DBL_MAX;
    double y0 = ok ? py[t] : DBL_MAX, y1 = ok ? py[t] : -DBL_MAX;
    if (bad) {
      x0 = y0 = -INFINITY;
      x1 = y1 = INFINITY;
    }
    for (int o = 32; o > 0; o >>= 1) {
      x0 = fmin(x0, __shfl_xor(x0, o));
      x1 = fmax(x1, __shfl_xor(x1, o));
      y0 = fmin(y0, __shfl_xor(y0, o));
      y1 = fmax(y1, __shfl_xor(y1, o));
    }
    wb[t][0] = x0;
    wb[t][1] = x1;
    wb[t][2] = y0;
    wb[t][3] = y1;
  }
  for (int64_t base = j0; base < j1; base += kTile) {
    const int n = (int)((j1 - base) < kTile ? (j1 - base) : kTile);
    bool k0 = true, k1 = true;
    if (tbox) {
      const double* b = tbox + 5 * (base / kTile);
      k0 = tile_near(b, g0);
      k1 = g1 && tile_near(b, g1);
      if (!k0 && !k1) continue;
    }
    __syncthreads();
    if ((int)threadIdx.x < n) {
      const int64_t j = base + threadIdx.x;
      tile[threadIdx.x] = vs ? vs[j] : vsrc_of(sx[j], sy[j], ss[j], sr[j]);
    }
    __syncthreads();
    // the same test for this wave's targets: uniform over the wave, no barrier below
    if (tbox) {
      const double* b = tbox + 5 * (base / kTile);
      k0 = k0 && tile_near(b, wb[0]);
      k1 = k1 && tile_near(b, wb[1]);
    }
    if (k0 && k1) {
      for (int k = 0; k < n; ++k) {
        const VSrc q = tile[k];
#pragma unroll
        for (int t = 0; t < kTpl; ++t) {
          const double dx = px[t] - q.x, dy = py[t] - q.y;
          const double x = (dx * dx + dy * dy) * q.rr;
          if (!(x >= kUnderflow)) acc[t] += q.c * exp_neg(x, e2t);  // a NaN exponent adds NaN
        }
      }
    } else if (k0 || k1) {  // one group: the other's pairs with this tile all add nothing
      const double qx = k0 ? px[0] : px[1], qy = k0 ? py[0] : py[1];
      double a = k0 ? acc[0] : acc[1];
      for (int k = 0; k < n; ++k) {
        const VSrc q = tile[k];
        const double dx = qx - q.x, dy = qy - q.y;
        const double x = (dx * dx + dy * dy) * q.rr;
        if (!(x >= kUnderflow)) a += q.c * exp_neg(x, e2t);
      }
      if (k0)
        acc[0] = a;
      else
        acc[1] = a;
    }
  }
#pragma unroll
  for (int t = 0; t < kTpl; ++t) {
    if (ti[t] >= nt) continue;
    if (direct)
      out[torder ? (int64_t)torder[ti[t]] : ti[t]] += acc[t];
    else
      out[by * nt + ti[t]] = acc[t];
  }
}

__global__ __launch_bounds__(kBlock) void vort_reduce_kernel(int64_t nchunk, int64_t nt,
                                                             const double* __restrict__ part,
                                                             const uint8_t* __restrict__ flag,
                                                             int64_t nchunk8,
                                                             const uint32_t* __restrict__ torder,
                                                             double* __restrict__ tw) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= nt) return;
  const int64_t wg = p / (kBlock * kTpl);
  // the workgroup's flags eight at a time (the row starts 8-byte aligned)
  const uint64_t* fw = reinterpret_cast<const uint64_t*>(flag + wg * nchunk8);
  double s = 0.0;
  for (int64_t w = 0; w < nchunk8 / 8; ++w) {
    const uint64_t f = fw[w];
    if (!f) continue;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int64_t c = 8 * w + b;
      if (c < nchunk && ((f >> (8 * b)) & 0xff)) s += part[c * nt + p];
    }
  }
  tw[torder ? (int64_t)torder[p] : p] += s;
}

}  // namespace

int vorticity_dev(Device* d, hipStream_t s, int64_t ns, const double* sx, const double* sy,
                  const double* ss, const double* sr, int64_t nt, const double* tx,
                  const double* ty, double* tw) {
  if (ns == 0 || nt == 0) return 0;
  const int64_t twgs = (nt + kBlock * kTpl - 1) / (kBlock * kTpl);
  const int64_t ntiles = (ns + kTile - 1) / kTile;
  // with many sources and targets, both in Morton order and source tiles skipped whole
  const bool prune = ns >= kPruneSources && nt >= kPruneTargets;
  int64_t nchunk;
  if (prune) {  // short chunks: the work sits in few (workgroup, chunk) pairs
    nchunk = (ntiles + O2D_VORT_PRUNED_TILES - 1) / O2D_VORT_PRUNED_TILES;
    const int64_t cap = kPartialBytes / ((int64_t)sizeof(double) * nt);
    if (nchunk > cap) nchunk = cap > 1 ? cap : 1;
  } else {  // enough workgroups to keep every CU busy
    const int64_t want = O2D_WANT_VORT * (int64_t)d->num_cu;
    nchunk = twgs >= want ? 1 : (want + twgs - 1) / twgs;
    if (nchunk > ntiles) nchunk = ntiles;
  }
  int64_t per = (ntiles + nchunk - 1) / nchunk;
  while (per < ntiles && twgs * ((ntiles + per - 1) / per) > INT32_MAX) per *= 2;  // 1-D grid
  nchunk = (ntiles + per - 1) / per;
  const int64_t chunk = per * kTile;
  const int64_t cells = twgs * nchunk;
  // a multiplier coprime to the cell count spreads the cells over the dispatch order
  int64_t perm = 1;
  if (prune && cells > 1) {
    perm = (int64_t)(0.6180339887 * (double)cells) | 1;
    auto gcd = [](int64_t a, int64_t b) {
      while (b) {
        const int64_t t = a % b;
        a = b;
        b = t;
      }
      return a;
    };
    while (gcd(perm, cells) != 1) perm += 2;
    perm %= cells;
  }
  double* out = tw;
  uint8_t* flag = nullptr;
  const int direct = nchunk == 1;
  const int64_t nchunk8 = (nchunk + 7) & ~int64_t(7);
  if (!direct) {
    void* p;
    O2D_TRY(slot(d, kPartials, sizeof(double) * nchunk * nt + twgs * nchunk8, s, &p));
    out = static_cast<double*>(p);
    flag = reinterpret_cast<uint8_t*>(out + nchunk * nt);
  }
  VSrc* vs = nullptr;
  double* tbox = nullptr;
  double* gbox = nullptr;
  const uint32_t* torder = nullptr;
  if (prune) {
    const uint32_t* sorder = nullptr;
    O2D_TRY(morton_order16(d, s, ns, sx, sy, &sorder));
    void* vp;
    const int64_t ngrp = (nt + kBlock - 1) / kBlock;
    O2D_TRY(slot(d, kVortSrc, sizeof(VSrc) * ns + sizeof(double) * (5 * ntiles + 4 * ngrp), s,
                 &vp));
    vs = static_cast<VSrc*>(vp);
    tbox = reinterpret_cast<double*>(vs + ns);
    gbox = tbox + 5 * ntiles;
    vort_gather_kernel<<<(unsigned)ntiles, kBlock, 0, s>>>(ns, sx, sy, ss, sr, sorder, vs, tbox);
    O2D_LAUNCH_CHECK();
    O2D_TRY(morton_order16(d, s, nt, tx, ty, &torder));  // reuses the order slot: vs holds the sources
    vort_target_box_kernel<<<(unsigned)ngrp, kBlock, 0, s>>>(nt, tx, ty, torder, gbox);
    O2D_LAUNCH_CHECK();
  }
  {
    KernelTimer kt(O2D_KT_VORT, s);
    vort_kernel<<<(unsigned)cells, kBlock, 0, s>>>(ns, sx, sy, ss, sr, nt, tx, ty, chunk, out,
                                                   flag, twgs, nchunk8, perm, direct, vs, tbox,
                                                   gbox, torder);
    kt.stop();
    O2D_LAUNCH_CHECK();
  }
  if (!direct) {
    vort_reduce_kernel<<<(unsigned)((nt + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        nchunk, nt, out, flag, nchunk8, torder, tw);
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
