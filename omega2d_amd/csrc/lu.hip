// lu.hip - LU factorisation with partial pivoting of the BEM matrix (LAPACK dgetrf semantics:
// P A = L U, unit lower L, 1-based row interchanges ipiv, info = first zero pivot), for the
// solve of src/BEM.h:169-234.
//
// A library dgetrf on a matrix of a few thousand rows spends its time in three kernel launches
// per column (pivot search, singularity check, scale and update), microseconds each. Here the
// factorisation is right-looking in blocks of kNB columns, with three launches per block:
//   lu_panel_kernel   one workgroup factors the block's panel (rows P..n-1, kNB columns) in
//                     leaves of kLeaf columns, left-looking: a leaf's columns are loaded into
//                     registers, take the updates of the block's earlier leaves (the U rows by
//                     a wavefront forward substitution, one wave per column, the rows below as
//                     one streamed update), and are factored with one workgroup barrier per
//                     column (pivot searches reduce on the DPP paths). The kernel avoids
//                     branches where it can: a row slot's existence is known at compile time
//                     for all but the last slot, and picks and loads use selects. Row
//                     interchanges are not applied in memory: the workgroup keeps the map from
//                     logical to physical row in LDS, reads and writes through it, and lists
//                     the block's interchanges at the end as (destination, source) row pairs;
//   lu_swap_solve_kernel  those row pairs applied to every column of the matrix (all loads of a
//                     column before its stores), and on the trailing columns the block's top
//                     rows solved, L11^-1 A12 (one wave per column, the forward substitution as
//                     a wavefront over its lanes);
//   rocBLAS dgemm     the trailing update on the fp64 matrix cores; a large one is split so
//                     that the next block's panel runs beside most of it (look-ahead).
// Every entry takes its updates in the same order and with the same fused multiply-adds as a
// right-looking factorisation that applies each leaf's rank-kLeaf update as soon as the leaf is
// factored, so the factors do not depend on the loop order. Pivots are chosen as LAPACK's idamax
// chooses them (largest magnitude, first index on ties), so the factors match a library dgetrf
// up to rounding.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <climits>

#include "nbody.h"
#include "o2d_runtime.h"

namespace o2d {

int blas_of(Device* d, hipStream_t s, rocblas_handle* out);  // bem.hip

namespace {

constexpr int kNB = 64;                  // columns per block (trailing update by dgemm)
constexpr int kLeaf = 8;                 // columns per leaf
constexpr int kMaxOrder = 6144;          // panel rows held in registers
constexpr int kPT = 512;                 // threads of the panel kernel (two waves per SIMD)
constexpr int kPW = kPT / 64;            // its waves
static_assert(kPW <= 8, "the waves' pivot candidates are picked within eight lanes");
static_assert(kPW >= kLeaf, "one wave per leaf column in the U-row solve");
static_assert(kLeaf * (kNB - kLeaf) <= kPT, "one earlier L entry of a leaf's rows per thread");
static_assert(kNB <= kPT, "rows past the first kPT lie below every diagonal row of the block");
constexpr int kMaxQ = (kMaxOrder + kPT - 1) / kPT;
constexpr int kMaxPairs = 2 * kNB;
constexpr int kLookAhead = 2048;  // trailing columns beside the next panel, at least

// (value, row) of the larger magnitude; the smaller row on ties (idamax's first index)
__device__ __forceinline__ void pick(double& v, int& i, double ov, int oi) {
  const bool take = (ov > v) | ((ov == v) & (oi < i));  // selects, no short-circuit branch
  v = take ? ov : v;
  i = take ? oi : i;
}

// ---- wave-wide reductions on the DPP paths (no LDS crossbar): xor 1, xor 2, half-row mirror
// and row mirror leave every lane with its 16-lane row's result; four lane reads combine the
// rows, so every lane returns the wave's result
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int x) {
  return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = dpp_i32<CTRL>((int)b);
  const int hi = dpp_i32<CTRL>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double lane_f64(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// the largest x over the wave (values are magnitudes or -1; no NaN reaches a pick)
__device__ __forceinline__ double wave_max(double x) {
  x = fmax(x, dpp_f64<0xB1>(x));
  x = fmax(x, dpp_f64<0x4E>(x));
  x = fmax(x, dpp_f64<0x141>(x));
  x = fmax(x, dpp_f64<0x140>(x));
  return fmax(fmax(lane_f64(x, 0), lane_f64(x, 16)), fmax(lane_f64(x, 32), lane_f64(x, 48)));
}
__device__ __forceinline__ int wave_min(int x) {
  x = min(x, dpp_i32<0xB1>(x));
  x = min(x, dpp_i32<0x4E>(x));
  x = min(x, dpp_i32<0x141>(x));
  x = min(x, dpp_i32<0x140>(x));
  return min(min(__builtin_amdgcn_readlane(x, 0), __builtin_amdgcn_readlane(x, 16)),
             min(__builtin_amdgcn_readlane(x, 32), __builtin_amdgcn_readlane(x, 48)));
}
// (largest value, smallest index among the lanes holding it) over the wave: the pick order
__device__ __forceinline__ void wave_pick(double& v, int& i) {
  const double vm = wave_max(v);
  const unsigned long long hit = __ballot(v == vm);
  if (__popcll(hit) == 1)  // the common case: one lane holds the largest value
    i = __builtin_amdgcn_readlane(i, __ffsll(hit) - 1);
  else
    i = wave_min(v == vm ? i : INT_MAX);
  v = vm;
}

// LDS writes of this wave visible to its own later reads, and no compiler reordering across
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The panel of block [P, P + nb): rows P..n-1 (m = n - P of them, relative row r = logical row
// P + r), thread t holding relative rows t, t + kPT, ... (Q of them). Physical row of relative
// row r: P + perm[r]. Writes ipiv (1-based) for the block's columns, *info on a zero pivot, and
// bpairs = [count, (destination, source)...], the block's row permutation.
template <int Q>
__global__ __launch_bounds__(kPT) void lu_panel_kernel(int n, int P, int nb,
                                                       double* __restrict__ A, int lda,
                                                       int* __restrict__ ipiv,
                                                       int* __restrict__ info,
                                                       int* __restrict__ bpairs,
                                                       double* __restrict__ l11out) {
  __shared__ int perm[kMaxOrder];
  __shared__ double L11[kNB][kNB + 1];   // final rows of the panel's top kNB rows
  __shared__ double X[kNB][kLeaf];       // the U rows of the current leaf's columns
  __shared__ double cand_v[2][kPW];      // per wave: the pivot candidate, double-buffered
  __shared__ int cand_i[2][kPW];         // by column parity (one barrier per column)
  __shared__ double cand_row[2][kPW][kLeaf];
  __shared__ double diag_row[2][kLeaf];  // the diagonal row before the interchange
  __shared__ int spiv[kNB];              // the block's pivots (relative logical rows)
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int m = n - P;
  const size_t base = (size_t)P * lda + P;  // A + base: the panel's top-left entry
  for (int r = t; r < m; r += kPT) perm[r] = r;
  int szero = -1;
  __syncthreads();

  // whether row t + q * kPT lies in the panel: with Q = ceil(m / kPT) every row of q < Q - 1 does
  auto exists = [&](int q) { return q < Q - 1 || t + q * kPT < m; };
  // the leaf of columns [c0, c0 + l) in logical row order, through perm
  auto load_leaf = [&](double (&dst)[Q][kLeaf], int c0, int l) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int r = t + q * kPT;
      const int pr = exists(q) ? perm[r] : 0;
#pragma unroll
      for (int c = 0; c < kLeaf; ++c)
        dst[q][c] = (exists(q) && c < l) ? A[base + (size_t)(c0 + c) * lda + pr] : 0.0;
    }
  };
  const int nleaf = (nb + kLeaf - 1) / kLeaf;
  // 1. the leaf's columns, rows in the current logical order: the first here, each later one
  //    at the end of the leaf before it, once that leaf's values are stored
  double a[Q][kLeaf];
  load_leaf(a, 0, nb < kLeaf ? nb : kLeaf);
  for (int k = 0; k < nleaf; ++k) {
    const int c0 = kLeaf * k;          // the leaf's first column and diagonal row (relative)
    const int l = nb - c0 < kLeaf ? nb - c0 : kLeaf;
    if (k > 0) {
      // 2. the U rows 0..c0-1 of the leaf's columns: wave w takes column w, lane i its row i,
      //    and the forward substitution with the unit-lower L11 runs as a wavefront over the
      //    lanes (at step s row s is final and every later row takes its term s, so each row
      //    adds its terms in index order, as the sequential solve does)
      if (wave == 0 && lane < c0) {
#pragma unroll
        for (int c = 0; c < kLeaf; ++c) X[lane][c] = a[0][c];
      }
      __syncthreads();
      if (wave < l) {
        double v = lane < c0 ? X[lane][wave] : 0.0;
        for (int s0 = 0; s0 < c0 - 1; s0 += kLeaf) {
          double lc[kLeaf];  // this lane's L11 entries for the next kLeaf steps
#pragma unroll
          for (int j = 0; j < kLeaf; ++j) lc[j] = lane < c0 ? L11[lane][s0 + j] : 0.0;
#pragma unroll
          for (int j = 0; j < kLeaf; ++j) {
            const int s = s0 + j;
            if (s < c0 - 1) {
              const double nv = __builtin_fma(-lc[j], lane_f64(v, s), v);
              v = (lane > s && lane < c0) ? nv : v;
            }
          }
        }
        if (lane < c0) X[lane][wave] = v;
      }
      __syncthreads();
      // 3. the rows below take every earlier leaf's update, column by column of L; the L
      //    entries of kCh columns for all of a thread's rows are loaded before they are used.
      //    Every load reads a row of the panel (rows past m read row m - 1) so no load needs a
      //    branch; a row past m only feeds its own padding. The U rows (r < c0, in wave 0)
      //    take the update as well and are set back from X after it, so the loop has no
      //    branch either (their lanes share a wave with rows that need it).
      constexpr int kCh = Q <= 2 ? 8 : (Q == 3 ? 4 : (Q <= 7 ? 2 : 1));
      int prq[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int r = t + q * kPT;
        prq[q] = perm[exists(q) ? r : m - 1];
        if (r < c0) {
#pragma unroll
          for (int c = 0; c < kLeaf; ++c) a[q][c] = X[r][c];
        }
      }
      auto load = [&](double (&lv)[Q][kCh], int m0) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
#pragma unroll
          for (int cc = 0; cc < kCh; ++cc) lv[q][cc] = A[base + (size_t)(m0 + cc) * lda + prq[q]];
        }
        __builtin_amdgcn_sched_barrier(0);  // the loads stay ahead of the batch applied next
      };
      auto apply = [&](const double (&lv)[Q][kCh], int m0) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
#pragma unroll
          for (int cc = 0; cc < kCh; ++cc) {
#pragma unroll
            for (int c = 0; c < kLeaf; ++c)
              a[q][c] = __builtin_fma(-lv[q][cc], X[m0 + cc][c], a[q][c]);
          }
        }
      };
      // two batches: the next one's loads are in flight while the current one is applied.
      // The last batch loads itself again instead of branching around the loads, so the
      // loads of a batch are always followed by those of the next and every wait for a
      // batch waits for that batch only
      // The loop is unrolled over the most batches a leaf can have (c0 <= kNB - kLeaf), so no
      // loop header merges the batches' outstanding loads (not for Q > 7, whose registers
      // would spill).
      constexpr int kMaxBatch = (kNB - kLeaf) / kCh;
      constexpr int kUnroll = Q <= 7 ? kMaxBatch / 2 + 1 : 1;
      double la[Q][kCh], lb[Q][kCh];
      load(la, 0);
#pragma unroll kUnroll
      for (int b = 0; b < kMaxBatch; b += 2) {  // c0 is a positive multiple of kLeaf
        const int m0 = b * kCh;                 // and kCh divides kLeaf
        const bool more = m0 + kCh < c0;
        load(lb, more ? m0 + kCh : m0);
        apply(la, m0);
        if (!more) break;
        const bool more2 = m0 + 2 * kCh < c0;
        load(la, more2 ? m0 + 2 * kCh : m0 + kCh);
        apply(lb, m0 + kCh);
        if (!more2) break;
      }
      if (t < c0) {
#pragma unroll
        for (int c = 0; c < kLeaf; ++c) a[0][c] = X[t][c];
      }
    }
    // 4. the leaf's columns factored, one workgroup barrier per column
#pragma unroll
    for (int jj = 0; jj < kLeaf; ++jj) {
      if (jj >= l) break;
      const int cj = c0 + jj;
      const int b = jj & 1;
      // the thread's candidate: its rows come in increasing order, so a strictly larger
      // magnitude replaces the pick and ties keep the first row
      double best = -1.0;
      int bi = INT_MAX;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int r = t + q * kPT;
        const double v = fabs(a[q][jj]);
        const bool take = exists(q) & ((q > 0) | (r >= cj)) & (v > best);
        best = take ? v : best;
        bi = take ? r : bi;
      }
      wave_pick(best, bi);
      // the wave's candidate row (held by thread bi % kPT as its row bi / kPT) and the
      // diagonal row (thread cj, its first row), to LDS
      if (bi != INT_MAX && t == bi % kPT) {
        const int qb = bi / kPT;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          if (q == qb) {
#pragma unroll
            for (int c = 0; c < kLeaf; ++c) cand_row[b][wave][c] = a[q][c];
          }
        }
      }
      if (t == cj) {
#pragma unroll
        for (int c = 0; c < kLeaf; ++c) diag_row[b][c] = a[0][c];
      }
      if (lane == 0) {
        cand_v[b][wave] = best;
        cand_i[b][wave] = bi;
      }
      __syncthreads();
      // the waves' candidates in lanes 0..kPW-1, picked the same way (xor 1, xor 2 and the
      // half-row mirror cover eight lanes)
      double v = -1.0;
      int p = INT_MAX;
      if (lane < kPW) {
        v = cand_v[b][lane];
        p = cand_i[b][lane];
      }
      pick(v, p, dpp_f64<0xB1>(v), dpp_i32<0xB1>(p));
      pick(v, p, dpp_f64<0x4E>(v), dpp_i32<0x4E>(p));
      pick(v, p, dpp_f64<0x141>(v), dpp_i32<0x141>(p));
      p = __builtin_amdgcn_readlane(p, 0);
      const int pt = p % kPT;  // the thread holding pivot row p
      const double* prow = cand_row[b][pt >> 6];
      double pr[kLeaf], dr[kLeaf];
#pragma unroll
      for (int c = 0; c < kLeaf; ++c) {
        pr[c] = prow[c];
        dr[c] = diag_row[b][c];
      }
      // the interchange of rows cj and p, in registers (a thread can hold both)
      if (t == cj) {
#pragma unroll
        for (int c = 0; c < kLeaf; ++c) a[0][c] = pr[c];
      }
      if (p != cj && t == pt) {
        const int qp = p / kPT;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          if (q == qp) {
#pragma unroll
            for (int c = 0; c < kLeaf; ++c) a[q][c] = dr[c];
          }
        }
      }
      if (t == 0) spiv[cj] = p;
      const double piv = pr[jj];
      if (piv != 0.0) {
        const double rp = 1.0 / piv;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          const int r = t + q * kPT;
          if (exists(q) && (q > 0 || r > cj)) {
            const double lr = a[q][jj] * rp;
            a[q][jj] = lr;
            // columns past a short last leaf (c >= l) hold zeros in every row and in the
            // pivot row, are never searched and never stored: they need no test here
#pragma unroll
            for (int c = jj + 1; c < kLeaf; ++c) a[q][c] = __builtin_fma(-lr, pr[c], a[q][c]);
          }
        }
      } else if (t == 0 && szero < 0) {
        szero = cj;  // singular: the column is left as it is, like dgetf2
      }
    }
    // the leaf's interchanges into perm at once: the swaps (c0 + j, spiv[c0 + j]) in order j
    // leave perm[x] = old perm[tau_0(tau_1(...tau_{l-1}(x)))], tau_j that swap, and move only
    // the diagonal and pivot rows. Only wave 0 takes part (spiv is written by thread 0 and
    // 2 l <= 16 threads recompose perm, which nobody reads during the column loop), so wave
    // syncs order it and one workgroup barrier publishes the result.
    if (wave == 0) {
      wave_sync();
      int px = -1, pv = 0;
      if (t < 2 * l) {
        px = t < l ? c0 + t : spiv[c0 + t - l];
        int y = px;
        for (int j = l - 1; j >= 0; --j) {
          const int cj = c0 + j, pj = spiv[cj];
          y = y == cj ? pj : (y == pj ? cj : y);
        }
        pv = perm[y];
      }
      wave_sync();
      if (px >= 0) perm[px] = pv;  // threads with the same px write the same entry
    }
    __syncthreads();
    // 5. the leaf's columns back to their physical rows; its rows are now final: their L
    //    entries (earlier leaves' from memory, this leaf's from registers) to L11. Loads and
    //    stores share one in-order counter, so the order is: the L entries, the stores, then
    //    the next leaf's loads, and the wait for the L entries waits for nothing after the
    //    stores. Neither load has a branch around it (a thread with no L entry reads row
    //    perm[c0], column 0; the next leaf's columns past the panel read its last column),
    //    so the count of loads in flight is known where the L entries are used.
    double lprev;
    int ei = -1, em = 0;
    {
      const bool mine = t < l * c0;
      const int tt = mine ? t : 0, cd = c0 > 0 ? c0 : 1;
      const int xi = tt / cd, xm = tt % cd;
      lprev = A[base + (size_t)xm * lda + perm[c0 + xi]];
      if (mine) {
        ei = xi;
        em = xm;
      }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int r = t + q * kPT;
      if (exists(q)) {
        const size_t pr = perm[r];
#pragma unroll
        for (int c = 0; c < kLeaf; ++c)
          if (c < l) A[base + (size_t)(c0 + c) * lda + pr] = a[q][c];
      }
    }
    if (t >= c0 && t < c0 + l) {
#pragma unroll
      for (int c = 0; c < kLeaf; ++c) L11[t][c0 + c] = a[0][c];
    }
    // the next leaf in logical row order, into the registers just stored (on the last leaf the
    // loads are not used). Rows past m read row m - 1 and columns past nb column nb - 1: those
    // entries are never searched, stored or used for another row
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int pr = perm[exists(q) ? t + q * kPT : m - 1];
#pragma unroll
      for (int c = 0; c < kLeaf; ++c) {
        const int col = c0 + kLeaf + c < nb ? c0 + kLeaf + c : nb - 1;
        a[q][c] = A[base + (size_t)col * lda + pr];
      }
    }
    if (ei >= 0) L11[c0 + ei][em] = lprev;
    if (t < l) ipiv[P + c0 + t] = P + spiv[c0 + t] + 1;
    __syncthreads();
  }
  if (t == 0 && szero >= 0 && *info == 0) *info = P + szero + 1;
  // the block's row permutation: only the diagonal rows and the pivot rows can have moved
  if (t < 2 * nb) {
    const int r = t < nb ? t : spiv[t - nb];
    bpairs[1 + 2 * t] = P + r;
    bpairs[2 + 2 * t] = P + perm[r];
  }
  if (t == 0) bpairs[0] = 2 * nb;
  // L11 of a full block, column-major (l11out[s * kNB + r] = L[r][s]), for the trailing solve
  if (nb == kNB) {
    for (int e = t; e < kNB * kNB; e += kPT) l11out[e] = L11[e % kNB][e / kNB];
  }
}

// The block's row pairs on every column, one wave per column, all loads of a column before its
// stores (rows that move form closed cycles). On a trailing column (j >= P + kNB) the block rows
// are then solved with L11 (A12 <- L11^-1 A12, L11 unit lower, read from l11 in column-major
// order): pair l < kNB has destination P + l, so lane l already holds that row's new value, and
// the forward substitution runs as a wavefront over the lanes (at step s, row s is final and
// every later row takes its term s, so each row adds its terms in index order, as a sequential
// solve does).
__global__ __launch_bounds__(256) void lu_swap_solve_kernel(int n, int P, double* __restrict__ A,
                                                            int lda,
                                                            const int* __restrict__ bpairs,
                                                            const double* __restrict__ l11) {
  static_assert(kNB == 64 && kMaxPairs == 2 * 64, "a lane per block row, two pairs per lane");
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= n) return;  // the whole wave
  const int cnt = bpairs[0];
  double* cp = A + (size_t)j * lda;
  const bool ok0 = lane < cnt, ok1 = lane + 64 < cnt;
  const int d0 = ok0 ? bpairs[1 + 2 * lane] : 0, s0 = ok0 ? bpairs[2 + 2 * lane] : 0;
  const int d1 = ok1 ? bpairs[1 + 2 * (lane + 64)] : 0, s1 = ok1 ? bpairs[2 + 2 * (lane + 64)] : 0;
  const double v0 = ok0 ? cp[s0] : 0.0;
  const double v1 = ok1 ? cp[s1] : 0.0;
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // every load of the column has returned
  if (j < P + kNB) {
    if (ok0) cp[d0] = v0;
    if (ok1) cp[d1] = v1;
    return;
  }
  // a trailing column (its block has kNB columns, so cnt = 2 kNB): the block rows are written
  // once, after the solve; pivot pairs into the block rows carry values lanes already hold
  if (ok1 && d1 - P >= kNB) cp[d1] = v1;
  double lc[kNB];  // L[lane][s]
#pragma unroll
  for (int s = 0; s < kNB; ++s) lc[s] = l11[s * kNB + lane];
  double v = v0;
#pragma unroll
  for (int s = 0; s < kNB - 1; ++s) {
    const double xs = lane_f64(v, s);
    if (lane > s) v = __builtin_fma(-lc[s], xs, v);
  }
  cp[P + lane] = v;
}

__global__ void lu_zero_info_kernel(int* info) { *info = 0; }

// the panel kernel for q rows per thread, q <= Q
template <int Q>
void launch_panel(int q, hipStream_t s, int n, int P, int nb, double* A, int lda, int* ipiv,
                  int* info, int* bpairs, double* l11) {
  if constexpr (Q > 1) {
    if (q < Q) {
      launch_panel<Q - 1>(q, s, n, P, nb, A, lda, ipiv, info, bpairs, l11);
      return;
    }
  }
  lu_panel_kernel<Q><<<1, kPT, 0, s>>>(n, P, nb, A, lda, ipiv, info, bpairs, l11);
}

}  // namespace

// The device's second stream for the LU (default priority, like every stream here), its own
// rocBLAS handle (host pointer mode) and two events, created on first use.
static int aux_of(Device* d, hipStream_t s, hipStream_t* B, rocblas_handle* hb) {
  (void)s;
  if (!d->aux) O2D_HIP(hipStreamCreateWithFlags(&d->aux, hipStreamNonBlocking));
  for (hipEvent_t& e : d->aux_ev)
    if (!e) O2D_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  if (!d->blas_aux) {
    rocblas_handle h;
    if (rocblas_create_handle(&h) != rocblas_status_success)
      return fail(O2D_ERR_HIP, "rocblas_create_handle failed");
    d->blas_aux = h;
    if (rocblas_set_stream(h, d->aux) != rocblas_status_success ||
        rocblas_set_pointer_mode(h, rocblas_pointer_mode_host) != rocblas_status_success)
      return fail(O2D_ERR_HIP, "rocBLAS handle set-up failed");
  }
  *B = d->aux;
  *hb = static_cast<rocblas_handle>(d->blas_aux);
  return 0;
}

// Largest order the blocked factorisation takes (a panel's rows must fit one workgroup).
int64_t lu_max_order() { return kMaxOrder; }

// In-place LU of the n x n column-major matrix A (leading dimension lda), pivots to ipiv
// (1-based), info (device int) = 0 or the first zero pivot (1-based).
int lu_factor_dev(Device* d, hipStream_t s, int n, double* A, int lda, int* ipiv, int* info) {
  if (n <= 0) return 0;
  if (n > kMaxOrder) return fail(O2D_ERR_SIZE, "blocked LU takes n <= %d", kMaxOrder);
  rocblas_handle h;
  O2D_TRY(blas_of(d, s, &h));
  if (rocblas_set_pointer_mode(h, rocblas_pointer_mode_host) != rocblas_status_success)
    return fail(O2D_ERR_HIP, "rocblas_set_pointer_mode failed");
  void* pb;  // the block's L11 (kNB x kNB doubles), then its row pairs
  O2D_TRY(slot(d, kLuPairs, sizeof(double) * kNB * kNB + sizeof(int) * (1 + 2 * kMaxPairs), s,
               &pb));
  double* l11 = static_cast<double*>(pb);
  int* bpairs = reinterpret_cast<int*>(l11 + kNB * kNB);
  lu_zero_info_kernel<<<1, 1, 0, s>>>(info);
  O2D_LAUNCH_CHECK();
  // look-ahead: a large trailing update of block k is split into the next block's columns, on
  // s, and the rest, on the second stream, so the next panel (one workgroup) runs beside it. The
  // next laswp touches every column, so s waits for the rest first. Below kLookAhead trailing
  // columns the second launch costs more than the overlap saves (at order 2049 the split
  // update took 6.80 ms against 6.58 ms unsplit; at 4097, 18.18 against 18.90).
  hipStream_t B;
  rocblas_handle hb;
  O2D_TRY(aux_of(d, s, &B, &hb));
  hipEvent_t e_next = d->aux_ev[0], e_rest = d->aux_ev[1];
  bool rest_pending = false;
  const double one = 1.0, minus_one = -1.0;
  for (int P = 0; P < n; P += kNB) {
    const int nb = std::min(kNB, n - P);
    const int q = (n - P + kPT - 1) / kPT;
    launch_panel<kMaxQ>(q, s, n, P, nb, A, lda, ipiv, info, bpairs, l11);
    O2D_LAUNCH_CHECK();
    if (rest_pending) O2D_HIP(hipStreamWaitEvent(s, e_rest, 0));
    rest_pending = false;
    lu_swap_solve_kernel<<<(unsigned)((n + 3) / 4), 256, 0, s>>>(n, P, A, lda, bpairs, l11);
    O2D_LAUNCH_CHECK();
    const int pend = P + nb;
    const int nr = n - pend;
    if (nr > 0) {  // nb = kNB here
      const int nn = nr - kNB >= kLookAhead ? kNB : nr;  // columns updated on s
      const double* L21 = A + (size_t)P * lda + pend;
      if (rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nr, nn, kNB,
                        &minus_one, L21, lda, A + (size_t)pend * lda + P, lda, &one,
                        A + (size_t)pend * lda + pend, lda) != rocblas_status_success)
        return fail(O2D_ERR_HIP, "rocblas_dgemm failed");
      if (nr > nn) {
        O2D_HIP(hipEventRecord(e_next, s));
        O2D_HIP(hipStreamWaitEvent(B, e_next, 0));
        const size_t c = (size_t)(pend + nn) * lda;
        if (rocblas_dgemm(hb, rocblas_operation_none, rocblas_operation_none, nr, nr - nn, kNB,
                          &minus_one, L21, lda, A + c + P, lda, &one, A + c + pend,
                          lda) != rocblas_status_success)
          return fail(O2D_ERR_HIP, "rocblas_dgemm failed");
        O2D_HIP(hipEventRecord(e_rest, B));
        rest_pending = true;
      }
    }
  }
  if (rest_pending) O2D_HIP(hipStreamWaitEvent(s, e_rest, 0));
  return 0;
}

}  // namespace o2d
