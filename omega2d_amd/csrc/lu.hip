// lu.hip - LU factorisation with partial pivoting of the BEM matrix (LAPACK dgetrf semantics:
// P A = L U, unit lower L, 1-based row interchanges ipiv, info = first zero pivot), for the
// solve of src/BEM.h:169-234.
//
// A library dgetrf on a matrix of a few thousand rows spends its time in three kernel launches
// per column (pivot search, singularity check, scale and update), microseconds each. Here the
// factorisation is right-looking in column blocks of kNB:
//   for each block of kNB columns, in leaves of kLeaf columns:
//     lu_leaf_kernel   one workgroup of 1024 threads holds the leaf's rows in registers and
//                      does the pivot search, interchange, scaling and rank-1 updates of its
//                      kLeaf columns with workgroup barriers only; it interchanges the same rows
//                      in the block's other columns and solves the block's top rows to the
//                      right of the leaf (U12 = L11^-1 A12);
//     lu_update_kernel the rank-kLeaf update of the block's remaining columns below the leaf;
//   then lu_laswp_kernel interchanges the block's rows in every column outside it, and rocBLAS
//   dtrsm and dgemm (fp64 matrix cores) update the trailing matrix.
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
constexpr int kLeafThreads = 1024;
constexpr int kMaxRowsPerThread = 6;  // leaves of up to 6144 rows
constexpr int kBlock = 256;

// (value, row) of the larger magnitude; the smaller row on ties (idamax's first index)
__device__ __forceinline__ void pick(double& v, int& i, double ov, int oi) {
  if (ov > v || (ov == v && oi < i)) {
    v = ov;
    i = oi;
  }
}

// Leaf of columns [c, c + l) of block [P, pend): rows c..n-1 (m = n - c) with thread t holding
// rows t, t + 1024, ... (Q of them).
template <int Q>
__global__ __launch_bounds__(kLeafThreads) void lu_leaf_kernel(int n, int c, int l, int P,
                                                                int pend, double* __restrict__ A,
                                                                int lda, int* __restrict__ ipiv,
                                                                int* __restrict__ info) {
  __shared__ double red_v[kLeafThreads / 64];
  __shared__ int red_i[kLeafThreads / 64];
  __shared__ double prow[kLeaf];  // the pivot row (leaf columns)
  __shared__ double jrow[kLeaf];  // row j before the interchange
  __shared__ double sl[kLeaf][kLeaf];  // L11 of the leaf
  __shared__ int s_piv;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int m = n - c;
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
      red_v[wave] = best;
      red_i[wave] = bi;
    }
    __syncthreads();
    if (t == 0) {
      double v = red_v[0];
      int i = red_i[0];
      for (int w = 1; w < kLeafThreads / 64; ++w) pick(v, i, red_v[w], red_i[w]);
      s_piv = i;
      ipiv[c + j] = c + i + 1;
    }
    __syncthreads();
    const int p = s_piv;
    // interchange rows j and p of the leaf (through LDS)
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int r = t + q * kLeafThreads;
      if (r == p) {
#pragma unroll
        for (int k = 0; k < kLeaf; ++k) prow[k] = a[q][k];
      }
      if (r == j) {
#pragma unroll
        for (int k = 0; k < kLeaf; ++k) jrow[k] = a[q][k];
      }
    }
    // and in the block's other columns, in place in global memory
    if (p != j) {
      const int col = P + t;
      if (t < pend - P && (col < c || col >= c + l)) {
        double* cp = A + (size_t)col * lda + c;
        const double x = cp[j];
        cp[j] = cp[p];
        cp[p] = x;
      }
    }
    __syncthreads();
    if (p != j) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int r = t + q * kLeafThreads;
        if (r == j) {
#pragma unroll
          for (int k = 0; k < kLeaf; ++k) a[q][k] = prow[k];
        } else if (r == p) {
#pragma unroll
          for (int k = 0; k < kLeaf; ++k) a[q][k] = jrow[k];
        }
      }
    }
    const double piv = prow[j];
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
            if (k < l) a[q][k] = __builtin_fma(-lr, prow[k], a[q][k]);
        }
      }
    } else if (t == 0 && *info == 0) {
      *info = c + j + 1;  // singular: the column is left as it is, like dgetf2
    }
    __syncthreads();  // prow and jrow are rewritten for the next column
  }
  // the factored leaf back to memory; its first l rows (L11 below the diagonal) to LDS
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
  __syncthreads();  // also orders the row interchanges above before the solve below
  // U12 = L11^-1 A12 for the block's columns right of the leaf: one column per thread
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

// A[r][j] -= sum_k L21[r][k] U12[k][j] for rows r >= c + l and the block's columns
// j in [c + l, pend), after the leaf [c, c + l)
__global__ __launch_bounds__(kBlock) void lu_update_kernel(int n, int c, int l, double* __restrict__ A,
                                                           int lda) {
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

// the block's interchanges (rows P..pend-1, in order) applied to every column outside it
__global__ __launch_bounds__(kBlock) void lu_laswp_kernel(int n, int P, int pend,
                                                          double* __restrict__ A, int lda,
                                                          const int* __restrict__ ipiv) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  const int ncol = n - (pend - P);
  if (j >= ncol) return;
  const int col = j < P ? j : j + (pend - P);
  double* cp = A + (size_t)col * lda;
  for (int k = P; k < pend; ++k) {
    const int p = ipiv[k] - 1;
    if (p != k) {
      const double x = cp[k];
      cp[k] = cp[p];
      cp[p] = x;
    }
  }
}

__global__ void lu_zero_info_kernel(int* info) { *info = 0; }

template <int Q>
void launch_leaf(hipStream_t s, int n, int c, int l, int P, int pend, double* A, int lda,
                 int* ipiv, int* info) {
  lu_leaf_kernel<Q><<<1, kLeafThreads, 0, s>>>(n, c, l, P, pend, A, lda, ipiv, info);
}

}  // namespace

// Largest order the blocked factorisation takes (a leaf's rows must fit one workgroup).
int64_t lu_max_order() { return (int64_t)kLeafThreads * kMaxRowsPerThread; }

// In-place LU of the n x n column-major matrix A (leading dimension lda), pivots to ipiv
// (1-based), info (device int) = 0 or the first zero pivot (1-based).
int lu_factor_dev(Device* d, hipStream_t s, int n, double* A, int lda, int* ipiv, int* info) {
  if (n <= 0) return 0;
  if (n > lu_max_order()) return fail(O2D_ERR_SIZE, "blocked LU takes n <= %lld", (long long)lu_max_order());
  rocblas_handle h;
  O2D_TRY(blas_of(d, s, &h));
  if (rocblas_set_pointer_mode(h, rocblas_pointer_mode_host) != rocblas_status_success)
    return fail(O2D_ERR_HIP, "rocblas_set_pointer_mode failed");
  lu_zero_info_kernel<<<1, 1, 0, s>>>(info);
  O2D_LAUNCH_CHECK();
  const double one = 1.0, minus_one = -1.0;
  for (int P = 0; P < n; P += kNB) {
    const int pend = std::min(n, P + kNB);
    for (int c = P; c < pend; c += kLeaf) {
      const int l = std::min(kLeaf, pend - c);
      const int m = n - c;
      const int q = (m + kLeafThreads - 1) / kLeafThreads;
      switch (q) {
        case 1: launch_leaf<1>(s, n, c, l, P, pend, A, lda, ipiv, info); break;
        case 2: launch_leaf<2>(s, n, c, l, P, pend, A, lda, ipiv, info); break;
        case 3: launch_leaf<3>(s, n, c, l, P, pend, A, lda, ipiv, info); break;
        case 4: launch_leaf<4>(s, n, c, l, P, pend, A, lda, ipiv, info); break;
        case 5: launch_leaf<5>(s, n, c, l, P, pend, A, lda, ipiv, info); break;
        default: launch_leaf<6>(s, n, c, l, P, pend, A, lda, ipiv, info); break;
      }
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
      lu_laswp_kernel<<<(unsigned)((ncol + kBlock - 1) / kBlock), kBlock, 0, s>>>(n, P, pend, A,
                                                                                  lda, ipiv);
      O2D_LAUNCH_CHECK();
    }
    const int nr = n - pend;
    if (nr > 0) {
      const int nb = pend - P;
      if (rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                        rocblas_diagonal_unit, nb, nr, &one, A + (size_t)P * lda + P, lda,
                        A + (size_t)pend * lda + P, lda) != rocblas_status_success)
        return fail(O2D_ERR_HIP, "rocblas_dtrsm failed");
      if (rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nr, nr, nb, &minus_one,
                        A + (size_t)P * lda + pend, lda, A + (size_t)pend * lda + P, lda, &one,
                        A + (size_t)pend * lda + pend, lda) != rocblas_status_success)
        return fail(O2D_ERR_HIP, "rocblas_dgemm failed");
    }
  }
  return 0;
}

}  // namespace o2d
