// lu.hip - LU factorisation with partial pivoting of the BEM matrix (LAPACK dgetrf semantics:
// P A = L U, unit lower L, 1-based row interchanges ipiv, info = first zero pivot), for the
// solve of src/BEM.h:169-234.
//
// A library dgetrf on a matrix of a few thousand rows spends its time in three kernel launches
// per column (pivot search, singularity check, scale and update), microseconds each. Here the
// factorisation is right-looking in blocks of kNB columns, each factored in leaves of kLeaf:
//   lu_leaf_kernel    one workgroup (kLeafThreads) holds the leaf's rows in registers and
//                     does the pivot search, interchange, scaling and rank-1 updates of its
//                     columns with two workgroup barriers per column; then it applies the
//                     leaf's interchanges to the block's other columns (as one permutation of at
//                     most 2 kLeaf rows), solves the block's top rows right of the leaf
//                     (U12 = L11^-1 A12), and after the block's last leaf lists the block's
//                     interchanges as (destination, source) row pairs; a pair's source is
//                     found by undoing the interchanges in reverse order, so no lane waits
//                     on another;
//   lu_update_kernel  the rank-kLeaf update of the block's columns right of the leaf;
//   lu_laswp_kernel   the block's row pairs applied to every column outside the block, all
//                     loads of a column before its stores;
//   lu_trsm_kernel    the block's top rows of the trailing columns, L11^-1 A12 (one column per
//                     lane, L11 in LDS);
//   rocBLAS dgemm     the trailing update on the fp64 matrix cores.
// Pivots are chosen as LAPACK's idamax chooses them (largest magnitude, first index on ties), so
// the factors match a library dgetrf up to rounding.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <climits>

#include "nbody.h"
#include "o2d_runtime.h"

namespace o2d {

int blas_of(Device* d, hipStream_t s, rocblas_handle* out);  // bem.hip

namespace {

constexpr int kNB = 64;          // columns per block (trailing update by dgemm)
constexpr int kLeaf = 8;         // columns per leaf (one workgroup)
#ifndef O2D_LU_THREADS
#define O2D_LU_THREADS 512
#endif
constexpr int kLeafThreads = O2D_LU_THREADS;  // one workgroup per leaf
constexpr int kWaves = kLeafThreads / 64;
constexpr int kMaxOrder = 6144;               // leaves of up to 6144 rows
constexpr int kMaxRowsPerThread = kMaxOrder / kLeafThreads;
constexpr int kMaxPairs = 2 * kNB;
constexpr int kBlock = 256;

// (value, row) of the larger magnitude; the smaller row on ties (idamax's first index)
__device__ __forceinline__ void pick(double& v, int& i, double ov, int oi) {
  if (ov > v || (ov == v && oi < i)) {
    v = ov;
    i = oi;
  }
}

// Leaf of columns [c, c + l) of block [P, pend): rows c..n-1 (m = n - c), thread t holding rows
// t, t + kLeafThreads, ... (Q of them). bpairs: after the block's last leaf, [count, (dst, src)...].
template <int Q>
__global__ __launch_bounds__(kLeafThreads) void lu_leaf_kernel(int n, int c, int l, int P,
                                                                int pend, double* __restrict__ A,
                                                                int lda, int* __restrict__ ipiv,
                                                                int* __restrict__ info,
                                                                int* __restrict__ bpairs) {
  // double-buffered by column parity, so a column needs two barriers
  __shared__ double red_v[2][kWaves];
  __shared__ int red_i[2][kWaves];
  __shared__ double prow[2][kLeaf];  // the pivot row (leaf columns)
  __shared__ double jrow[2][kLeaf];  // row j before the interchange
  __shared__ double sl[kLeaf][kLeaf];  // L11 of the leaf
  __shared__ int spiv[kLeaf];
  __shared__ int pdst[2 * kLeaf], psrc[2 * kLeaf];
  __shared__ int s_np;
  __shared__ int sbpiv[kNB];  // the block's pivot rows (0-based)
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int m = n - c;
  int szero = -1;  // first zero pivot of the leaf (thread 0)
  double a[Q][kLeaf];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int r = t + q * kLeafThreads;
#pragma unroll
    for (int k = 0; k < kLeaf; ++k)
      a[q][k] = (r < m && k < l) ? A[(size_t)(c + k) * lda + c + r] : 0.0;
  }
#pragma unroll
  for (int j = 0; j < kLeaf; ++j) {
    if (j >= l) break;
    const int b = j & 1;
    // pivot search over rows j..m-1 of column j
    double best = -1.0;
    int bi = INT_MAX;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int r = t + q * kLeafThreads;
      if (r < m && r >= j) pick(best, bi, fabs(a[q][j]), r);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) pick(best, bi, __shfl_xor(best, o), __shfl_xor(bi, o));
    if (lane == 0) {
      red_v[b][wave] = best;
      red_i[b][wave] = bi;
    }
    __syncthreads();
    int p = red_i[b][0];
    {
      double v = red_v[b][0];
#pragma unroll
      for (int w = 1; w < kWaves; ++w) pick(v, p, red_v[b][w], red_i[b][w]);
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int r = t + q * kLeafThreads;
      if (r == p) {
#pragma unroll
        for (int k = 0; k < kLeaf; ++k) prow[b][k] = a[q][k];
      }
      if (r == j) {
#pragma unroll
        for (int k = 0; k < kLeaf; ++k) jrow[b][k] = a[q][k];
      }
    }
    if (t == 0) spiv[j] = p;  // to global memory after the loop: a pending global store
                              // would hold the next barrier
    __syncthreads();
    if (p != j) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int r = t + q * kLeafThreads;
        if (r == j) {
#pragma unroll
          for (int k = 0; k < kLeaf; ++k) a[q][k] = prow[b][k];
        } else if (r == p) {
#pragma unroll
          for (int k = 0; k < kLeaf; ++k) a[q][k] = jrow[b][k];
        }
      }
    }
    const double piv = prow[b][j];
    if (piv != 0.0) {
      const double rp = 1.0 / piv;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int r = t + q * kLeafThreads;
        if (r < m && r > j) {
          const double lr = a[q][j] * rp;
          a[q][j] = lr;
#pragma unroll
          for (int k = j + 1; k < kLeaf; ++k)
            if (k < l) a[q][k] = __builtin_fma(-lr, prow[b][k], a[q][k]);
        }
      }
    } else if (t == 0 && szero < 0) {
      szero = j;  // singular: the column is left as it is, like dgetf2
    }
  }
  if (t == 0 && szero >= 0 && *info == 0) *info = c + szero + 1;
  // the factored leaf back to memory; its first l rows (L11 below the diagonal) to LDS; the
  // leaf's interchanges as one permutation (destination row <- source row, leaf-relative)
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int r = t + q * kLeafThreads;
    if (r < m) {
#pragma unroll
      for (int k = 0; k < kLeaf; ++k)
        if (k < l) A[(size_t)(c + k) * lda + c + r] = a[q][k];
    }
    if (q == 0 && t < l) {
#pragma unroll
      for (int k = 0; k < kLeaf; ++k) sl[t][k] = a[0][k];
    }
  }
  if (t < l) ipiv[c + t] = c + spiv[t] + 1;
  // The leaf's interchanges as one permutation: only rows 0..l-1 and the pivot rows move. For
  // each of those 2l candidate rows, the row whose content ends there is found by undoing the
  // interchanges in reverse order (a repeated candidate repeats its pair, which is harmless).
  if (t < 2 * l) {
    const int r = t < l ? t : spiv[t - l];
    int src = r;
    for (int j = l - 1; j >= 0; --j) {
      const int p = spiv[j];
      src = src == j ? p : (src == p ? j : src);
    }
    pdst[t] = r;
    psrc[t] = src;
  }
  if (t == 0) s_np = 2 * l;
  __syncthreads();
  // the same rows permuted in the block's other columns: every load before any store
  const int np = s_np;
  if (np > 0) {
    const int col = P + t;
    if (t < pend - P && (col < c || col >= c + l)) {
      double* cp = A + (size_t)col * lda + c;
      double v[2 * kLeaf];
#pragma unroll
      for (int i = 0; i < 2 * kLeaf; ++i)
        if (i < np) v[i] = cp[psrc[i]];
#pragma unroll
      for (int i = 0; i < 2 * kLeaf; ++i)
        if (i < np) cp[pdst[i]] = v[i];
    }
  }
  __syncthreads();  // the permuted rows are read by other threads below
  // U12 = L11^-1 A12 for the block's columns right of the leaf: one column per thread
  {
    const int col = c + l + t;
    if (col < pend) {
      double* cp = A + (size_t)col * lda + c;
      double x[kLeaf];
#pragma unroll
      for (int i = 0; i < kLeaf; ++i) {
        if (i < l) {
          double v = cp[i];
#pragma unroll
          for (int k = 0; k < i; ++k) v = __builtin_fma(-sl[i][k], x[k], v);
          x[i] = v;
          cp[i] = v;
        }
      }
    }
  }
  // after the block's last leaf: the block's interchanges (rows P..pend-1, in order) as row
  // pairs for the columns outside the block, found in the same way over the 2 nb candidates
  if (c + l == pend && pend - P < n) {
    const int nb = pend - P;
    __syncthreads();  // the leaf's own pivots are in global memory
    if (t < nb) sbpiv[t] = ipiv[P + t] - 1;
    __syncthreads();
    if (t < 2 * nb) {
      const int r = t < nb ? P + t : sbpiv[t - nb];
      int src = r;
      for (int k = nb - 1; k >= 0; --k) {
        const int p = sbpiv[k];
        src = src == P + k ? p : (src == p ? P + k : src);
      }
      bpairs[1 + 2 * t] = r;
      bpairs[2 + 2 * t] = src;
    }
    if (t == 0) bpairs[0] = 2 * nb;
  }
}

// A[r][j] -= sum_k L21[r][k] U12[k][j] for rows r >= c + l and the block's columns
// j in [c + l, pend), after the leaf [c, c + l)
__global__ __launch_bounds__(kBlock) void lu_update_kernel(int n, int c, int l,
                                                           double* __restrict__ A, int lda) {
  const int col = c + l + blockIdx.x;
  const int r = c + l + blockIdx.y * kBlock + threadIdx.x;
  if (r >= n) return;
  const double* u = A + (size_t)col * lda + c;  // U12 column (uniform)
  double v = A[(size_t)col * lda + r];
#pragma unroll
  for (int k = 0; k < kLeaf; ++k)
    if (k < l) v = __builtin_fma(-A[(size_t)(c + k) * lda + r], u[k], v);
  A[(size_t)col * lda + r] = v;
}

// the block's row pairs on every column outside [P, pend): two columns per workgroup, one
// pair per thread; a column's loads all land in LDS before any of its stores
__global__ __launch_bounds__(2 * kMaxPairs) void lu_laswp_kernel(int n, int P, int pend,
                                                                 double* __restrict__ A, int lda,
                                                                 const int* __restrict__ bpairs) {
  __shared__ double vals[2][kMaxPairs];
  const int cnt = bpairs[0];
  const int half = threadIdx.x / kMaxPairs, i = threadIdx.x % kMaxPairs;
  const int j = blockIdx.x * 2 + half;
  const int ncol = n - (pend - P);
  const bool ok = j < ncol && i < cnt;
  double* cp = A + (size_t)(j < P ? j : j + (pend - P)) * lda;
  if (ok) vals[half][i] = cp[bpairs[2 + 2 * i]];
  __syncthreads();
  if (ok) cp[bpairs[1 + 2 * i]] = vals[half][i];
}

// A12 <- L11^-1 A12 for the block rows P..P+kNB-1 and the trailing columns pend..n-1 (L11 unit
// lower, kNB x kNB): one column per lane, L11 staged in LDS
__global__ __launch_bounds__(64) void lu_trsm_kernel(int n, int P, double* __restrict__ A,
                                                     int lda) {
  __shared__ double L[kNB][kNB + 1];
  const int pend = P + kNB;
  for (int e = threadIdx.x; e < kNB * kNB; e += 64) {
    const int i = e % kNB, k = e / kNB;
    L[i][k] = A[(size_t)(P + k) * lda + P + i];
  }
  __syncthreads();
  const int col = pend + blockIdx.x * 64 + threadIdx.x;
  if (col >= n) return;
  double* cp = A + (size_t)col * lda + P;
  double x[kNB];
#pragma unroll
  for (int i = 0; i < kNB; ++i) x[i] = cp[i];
#pragma unroll
  for (int i = 1; i < kNB; ++i) {
    double v = x[i];
#pragma unroll
    for (int k = 0; k < i; ++k) v = __builtin_fma(-L[i][k], x[k], v);
    x[i] = v;
  }
#pragma unroll
  for (int i = 0; i < kNB; ++i) cp[i] = x[i];
}

__global__ void lu_zero_info_kernel(int* info) { *info = 0; }

// the leaf kernel for q rows per thread, q <= Q
template <int Q>
void launch_leaf(int q, hipStream_t s, int n, int c, int l, int P, int pend, double* A, int lda,
                 int* ipiv, int* info, int* bpairs) {
  if constexpr (Q > 1) {
    if (q < Q) {
      launch_leaf<Q - 1>(q, s, n, c, l, P, pend, A, lda, ipiv, info, bpairs);
      return;
    }
  }
  lu_leaf_kernel<Q><<<1, kLeafThreads, 0, s>>>(n, c, l, P, pend, A, lda, ipiv, info, bpairs);
}

}  // namespace

// Largest order the blocked factorisation takes (a leaf's rows must fit one workgroup).
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
  void* pb;
  O2D_TRY(slot(d, kLuPairs, sizeof(int) * (1 + 2 * kMaxPairs), s, &pb));
  int* bpairs = static_cast<int*>(pb);
  lu_zero_info_kernel<<<1, 1, 0, s>>>(info);
  O2D_LAUNCH_CHECK();
  const double one = 1.0, minus_one = -1.0;
  for (int P = 0; P < n; P += kNB) {
    const int pend = std::min(n, P + kNB);
    for (int c = P; c < pend; c += kLeaf) {
      const int l = std::min(kLeaf, pend - c);
      const int q = (n - c + kLeafThreads - 1) / kLeafThreads;
      launch_leaf<kMaxRowsPerThread>(q, s, n, c, l, P, pend, A, lda, ipiv, info, bpairs);
      O2D_LAUNCH_CHECK();
      const int rest = pend - (c + l), rows = n - (c + l);
      if (rest > 0 && rows > 0) {
        lu_update_kernel<<<dim3((unsigned)rest, (unsigned)((rows + kBlock - 1) / kBlock)), kBlock,
                           0, s>>>(n, c, l, A, lda);
        O2D_LAUNCH_CHECK();
      }
    }
    const int ncol = n - (pend - P);
    if (ncol > 0) {
      lu_laswp_kernel<<<(unsigned)((ncol + 1) / 2), 2 * kMaxPairs, 0, s>>>(n, P, pend, A, lda,
                                                                           bpairs);
      O2D_LAUNCH_CHECK();
    }
    const int nr = n - pend;
    if (nr > 0) {  // pend = P + kNB here
      lu_trsm_kernel<<<(unsigned)((nr + 63) / 64), 64, 0, s>>>(n, P, A, lda);
      O2D_LAUNCH_CHECK();
      if (rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nr, nr, kNB,
                        &minus_one, A + (size_t)P * lda + pend, lda, A + (size_t)pend * lda + P,
                        lda, &one, A + (size_t)pend * lda + pend, lda) != rocblas_status_success)
        return fail(O2D_ERR_HIP, "rocblas_dgemm failed");
    }
  }
  return 0;
}

}  // namespace o2d
