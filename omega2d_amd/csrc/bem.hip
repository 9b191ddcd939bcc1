// bem.hip - fp64 BEM influence matrix blocks, of a body acting on itself or on another body
// (src/Coefficients.h:155-573), and the LU solve of the assembled system (the reference solves
// with GMRES, src/BEM.h:169-234).
//
// Matrix layout is the reference's: column-major, column j = source unknown, row i = target
// equation; with two unknowns per panel the columns/rows interleave [vortex, source] and
// [tangential, normal]. One thread per (target panel i, source panel j): consecutive threads
// own consecutive rows of one column, so the column-major stores are coalesced. Every entry is
// the "two-way" average of the panel-j-on-midpoint-i influence and the reversed
// panel-i-on-midpoint-j influence (src/Coefficients.h:159,376-382), the diagonal is pi, and
// everything is scaled by 1/(2 pi). The augmented bottom row (panel lengths), right column
// (tangential/normal unit-rotation velocity) and corner (2 * volume) follow
// src/Coefficients.h:464-537.
#include <rocblas/rocblas.h>
#include <algorithm>
#include <rocsolver/rocsolver.h>

#include "nbody.h"
#include "o2d_math.h"
#include "o2d_runtime.h"
#include "panel_math.h"

namespace o2d {

void release_blas(Device* d);

namespace {

constexpr int kBlock = 256;
constexpr double kPi = 3.14159265358979323846;

struct alignas(16) Geom {
  double x0, y0, x1, y1, tx, ty, len, pad;
};

// Surfaces::compute_bases (src/Surfaces.h:699-724)
__global__ __launch_bounds__(kBlock) void geom_kernel(int64_t np, const double* __restrict__ nx,
                                                      const double* __restrict__ ny,
                                                      const uint32_t* __restrict__ idx,
                                                      Geom* __restrict__ g) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= np) return;
  const uint32_t a = idx[2 * p], b = idx[2 * p + 1];
  Geom q;
  q.x0 = nx[a];
  q.y0 = ny[a];
  q.x1 = nx[b];
  q.y1 = ny[b];
  const double ex = q.x1 - q.x0, ey = q.y1 - q.y0;
  const double base = sqrt(ex * ex + ey * ey);
  const double inv = 1.0 / base;
  q.tx = ex * inv;
  q.ty = ey * inv;
  q.len = base;
  q.pad = 0.0;
  g[p] = q;
}

// unit-strength vortex panel (a along the panel rotation, b across): the velocity of
// kernelu_1v_0p / kernelu_1vos_0p is (a, b) for the vortex part and (b, -a) for the source part;
// far pairs take the atanh series of panel_math.h
__device__ __forceinline__ void unit_panel(const Geom& s, double tx, double ty, double& a,
                                           double& b) {
  unit_panel_vel(s.x0, s.y0, s.x1, s.y1, tx, ty, a, b);
}

// one thread per (target panel i, source panel j); SELF: source and target are one surface,
// whose diagonal is the self term pi (src/Coefficients.h:389-401)
template <int NUNK, bool SELF>
__global__ __launch_bounds__(kBlock) void assemble_kernel(int64_t nt, const Geom* __restrict__ gs,
                                                          const Geom* __restrict__ gt,
                                                          double* __restrict__ A, int64_t lda) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;  // target panel (row)
  const int64_t j = blockIdx.y;                                   // source panel (column)
  if (i >= nt) return;
  const Geom sj = gs[j];
  const Geom ti = gt[i];
  const double xi = 0.5 * (ti.x1 + ti.x0), yi = 0.5 * (ti.y1 + ti.y0);
  const double xj = 0.5 * (sj.x0 + sj.x1), yj = 0.5 * (sj.y0 + sj.y1);
  double a1, b1, a2, b2;
  unit_panel(sj, xi, yi, a1, b1);  // panel j on the collocation point of panel i
  unit_panel(ti, xj, yj, a2, b2);  // reversed: panel i on the collocation point of panel j
  const double fact = sj.len / ti.len;
  const double fac = 1.0 / (2.0 * kPi);
  const double ntx = -ti.ty, nty = ti.tx;
  const bool diag = SELF && i == j;
  if constexpr (NUNK == 1) {
    double c = a1 * ti.tx + b1 * ti.ty;
    c -= (a2 * ti.tx + b2 * ti.ty) * fact;
    c *= 0.5;
    if (diag) c = kPi;
    A[j * lda + i] = c * fac;
  } else {
    double c11 = a1 * ti.tx + b1 * ti.ty;
    double c12 = a1 * ntx + b1 * nty;
    double c21 = b1 * ti.tx - a1 * ti.ty;
    double c22 = b1 * ntx - a1 * nty;
    c11 -= fact * (a2 * ti.tx + b2 * ti.ty);
    c12 -= fact * (a2 * ntx + b2 * nty);
    c21 -= fact * (b2 * ti.tx - a2 * ti.ty);
    c22 -= fact * (b2 * ntx - a2 * nty);
    c11 *= 0.5;
    c12 *= 0.5;
    c21 *= 0.5;
    c22 *= 0.5;
    if (diag) {
      c11 = kPi;
      c12 = 0.0;
      c21 = 0.0;
      c22 = kPi;
    }
    double* c0 = A + (2 * j) * lda + 2 * i;
    double* c1 = A + (2 * j + 1) * lda + 2 * i;
    c0[0] = c11 * fac;
    c0[1] = c12 * fac;
    c1[0] = c21 * fac;
    c1[1] = c22 * fac;
  }
}

// augmentation of a block between two different bodies (src/Coefficients.h:464-537 with
// &src != &targ): the bottom row is 0 (panel lengths count only a body's own panels), the right
// column is the source body's unit-rotation velocity on the target panels, the corner is 0
template <int NUNK>
__global__ __launch_bounds__(kBlock) void augment_pair_kernel(int64_t ns, int64_t nt,
                                                              const Geom* __restrict__ gt,
                                                              int src_aug, int targ_aug,
                                                              const double* __restrict__ uu,
                                                              const double* __restrict__ uv,
                                                              double* __restrict__ A,
                                                              int64_t lda) {
  const int64_t oldr = nt * NUNK, oldc = ns * NUNK;
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (targ_aug && k < oldc) A[k * lda + oldr] = 0.0;
  if (src_aug && k < nt) {
    const Geom q = gt[k];
    if (NUNK == 1) {
      A[oldc * lda + k] = uu[k] * q.tx + uv[k] * q.ty;
    } else {
      A[oldc * lda + 2 * k] = uu[k] * q.tx + uv[k] * q.ty;
      A[oldc * lda + 2 * k + 1] = uu[k] * (-q.ty) + uv[k] * q.tx;
    }
  }
  if (src_aug && targ_aug && k == 0) A[oldc * lda + oldr] = 0.0;
}

// bottom row (lengths), right column (unit-rotation tangential/normal velocity), corner
template <int NUNK>
__global__ __launch_bounds__(kBlock) void augment_kernel(int64_t np, const Geom* __restrict__ g,
                                                         const double* __restrict__ uu,
                                                         const double* __restrict__ uv,
                                                         double vol, double* __restrict__ A,
                                                         int64_t lda) {
  const int64_t oldn = np * NUNK;
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k < oldn) {
    double bottom;
    if (NUNK == 1) bottom = g[k].len;
    else bottom = (k % 2 == 0) ? g[k / 2].len : 0.0;
    A[k * lda + oldn] = bottom;
  }
  if (k < np) {
    const Geom q = g[k];
    if (NUNK == 1) {
      A[oldn * lda + k] = uu[k] * q.tx + uv[k] * q.ty;
    } else {
      A[oldn * lda + 2 * k] = uu[k] * q.tx + uv[k] * q.ty;
      A[oldn * lda + 2 * k + 1] = uu[k] * (-q.ty) + uv[k] * q.tx;
    }
  }
  if (k == 0) A[oldn * lda + oldn] = 2.0 * vol;
}

inline dim3 grid1(int64_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

}  // namespace

// the device's rocBLAS handle (created on first use), bound to stream s
int blas_of(Device* d, hipStream_t s, rocblas_handle* out) {
  if (!d->blas) {
    rocblas_handle h;
    if (rocblas_create_handle(&h) != rocblas_status_success)
      return fail(O2D_ERR_HIP, "rocblas_create_handle failed");
    d->blas = h;
  }
  rocblas_handle h = static_cast<rocblas_handle>(d->blas);
  if (rocblas_set_stream(h, s) != rocblas_status_success)
    return fail(O2D_ERR_HIP, "rocblas_set_stream failed");
  *out = h;
  return 0;
}

void release_blas(Device* d) {
  if (d->blas) {
    rocblas_destroy_handle(static_cast<rocblas_handle>(d->blas));
    d->blas = nullptr;
  }
  if (d->blas_aux) {
    rocblas_destroy_handle(static_cast<rocblas_handle>(d->blas_aux));
    d->blas_aux = nullptr;
  }
  for (hipEvent_t& e : d->aux_ev) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
  if (d->aux) (void)hipStreamDestroy(d->aux);
  d->aux = nullptr;
}

int bem_assemble_dev(Device* d, hipStream_t s, int64_t nnodes, const double* nx, const double* ny,
                     int64_t np, const uint32_t* idx, int nunk, int augmented,
                     const double* urot_u, const double* urot_v, double vol, double* A,
                     int64_t lda) {
  (void)nnodes;
  if (np == 0) return 0;
  void* gp;
  O2D_TRY(slot(d, kPanGeom, sizeof(Geom) * np, s, &gp));
  Geom* g = static_cast<Geom*>(gp);
  geom_kernel<<<grid1(np), kBlock, 0, s>>>(np, nx, ny, idx, g);
  O2D_LAUNCH_CHECK();
  if (np > 65535) return fail(O2D_ERR_SIZE, "at most 65535 panels per body (got %lld)", (long long)np);
  const dim3 grid((unsigned)((np + kBlock - 1) / kBlock), (unsigned)np);
  KernelTimer kt(O2D_KT_BEM_ASSEMBLE, s);
  if (nunk == 1) assemble_kernel<1, true><<<grid, kBlock, 0, s>>>(np, g, g, A, lda);
  else assemble_kernel<2, true><<<grid, kBlock, 0, s>>>(np, g, g, A, lda);
  kt.stop();
  O2D_LAUNCH_CHECK();
  if (augmented) {
    const int64_t m = np * nunk > np ? np * nunk : np;
    if (nunk == 1) augment_kernel<1><<<grid1(m), kBlock, 0, s>>>(np, g, urot_u, urot_v, vol, A, lda);
    else augment_kernel<2><<<grid1(m), kBlock, 0, s>>>(np, g, urot_u, urot_v, vol, A, lda);
    O2D_LAUNCH_CHECK();
  }
  return 0;
}

int bem_assemble_pair_dev(Device* d, hipStream_t s, const double* snx, const double* sny,
                          int64_t snp, const uint32_t* sidx, const double* tnx, const double* tny,
                          int64_t tnp, const uint32_t* tidx, int nunk, int src_aug, int targ_aug,
                          const double* urot_u, const double* urot_v, double* A, int64_t lda) {
  if (snp == 0 || tnp == 0) return 0;
  if (snp > 65535)
    return fail(O2D_ERR_SIZE, "at most 65535 source panels per block (got %lld)", (long long)snp);
  void *gsp, *gtp;
  O2D_TRY(slot(d, kPanGeom, sizeof(Geom) * snp, s, &gsp));
  O2D_TRY(slot(d, kPanGeom2, sizeof(Geom) * tnp, s, &gtp));
  Geom* gs = static_cast<Geom*>(gsp);
  Geom* gt = static_cast<Geom*>(gtp);
  geom_kernel<<<grid1(snp), kBlock, 0, s>>>(snp, snx, sny, sidx, gs);
  geom_kernel<<<grid1(tnp), kBlock, 0, s>>>(tnp, tnx, tny, tidx, gt);
  O2D_LAUNCH_CHECK();
  const dim3 grid((unsigned)((tnp + kBlock - 1) / kBlock), (unsigned)snp);
  KernelTimer kt(O2D_KT_BEM_ASSEMBLE, s);
  if (nunk == 1) assemble_kernel<1, false><<<grid, kBlock, 0, s>>>(tnp, gs, gt, A, lda);
  else assemble_kernel<2, false><<<grid, kBlock, 0, s>>>(tnp, gs, gt, A, lda);
  kt.stop();
  O2D_LAUNCH_CHECK();
  if (src_aug || targ_aug) {
    const int64_t m = std::max(snp * nunk, tnp);
    if (nunk == 1)
      augment_pair_kernel<1><<<grid1(m), kBlock, 0, s>>>(snp, tnp, gt, src_aug, targ_aug, urot_u,
                                                         urot_v, A, lda);
    else
      augment_pair_kernel<2><<<grid1(m), kBlock, 0, s>>>(snp, tnp, gt, src_aug, targ_aug, urot_u,
                                                         urot_v, A, lda);
    O2D_LAUNCH_CHECK();
  }
  return 0;
}

int bem_factor_dev(Device* d, hipStream_t s, int64_t n, const double* A, int64_t lda) {
  d->lu_ready = false;
  d->inv_ready = false;
  d->lu_solves = 0;
  void *lu, *piv, *info;
  O2D_TRY(slot(d, kLU, sizeof(double) * n * n, s, &lu));
  O2D_TRY(slot(d, kPivots, sizeof(rocblas_int) * n, s, &piv));
  O2D_TRY(slot(d, kInfo, sizeof(rocblas_int), s, &info));
  O2D_HIP(hipMemcpy2DAsync(lu, sizeof(double) * n, A, sizeof(double) * lda, sizeof(double) * n, n,
                           hipMemcpyDeviceToDevice, s));
  rocblas_handle h;
  O2D_TRY(blas_of(d, s, &h));
  if (n <= lu_max_order()) {
    O2D_TRY(lu_factor_dev(d, s, (int)n, static_cast<double*>(lu), (int)n,
                          static_cast<int*>(piv), static_cast<int*>(info)));
  } else if (rocsolver_dgetrf(h, (rocblas_int)n, (rocblas_int)n, static_cast<double*>(lu),
                              (rocblas_int)n, static_cast<rocblas_int*>(piv),
                              static_cast<rocblas_int*>(info)) != rocblas_status_success) {
    return fail(O2D_ERR_HIP, "rocsolver_dgetrf failed");
  }
  rocblas_int hinfo = 0;
  O2D_HIP(hipMemcpyAsync(&hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, s));
  O2D_HIP(hipStreamSynchronize(s));
  if (hinfo != 0) return fail(O2D_ERR_STATE, "BEM matrix is singular (getrf info=%d)", (int)hinfo);
  d->lu_n = n;
  d->lu_ready = true;
  return 0;
}

// the n x n identity, column-major (the right-hand sides whose solutions are A^-1)
__global__ __launch_bounds__(kBlock) void identity_kernel(int64_t n, double* __restrict__ a) {
  const int64_t nn = n * n;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < nn;
       e += (int64_t)gridDim.x * kBlock)
    a[e] = (e % (n + 1) == 0) ? 1.0 : 0.0;
}

// Solves A x = b with the kept factors. A matrix solved against more than once with the same
// factors is worth inverting: the second solve builds A^-1 from them, as the solution of A X = I
// (rocsolver getrs with n right-hand sides: the row interchanges, then two triangular solves
// with n columns on the fp64 matrix cores; 1.2 ms at order 2049 and 5.3 ms at 4097, where
// rocsolver getri took 4.8 and 22.7 ms), about the cost of one solve by two latency-bound
// triangular solves (1.1 and 2.8 ms). Every solve from then on is one dgemv, an HBM-bound read
// of n^2 doubles (0.09 and 0.17 ms). A body that does not move keeps its inverse for the whole
// run; a body in relative motion, refactored every step and solved once per Runge-Kutta stage,
// builds it at its second stage.
constexpr int kSolvesBeforeInverse = 1;

int bem_solve_dev(Device* d, hipStream_t s, int64_t n, const double* b, double* x) {
  if (!d->lu_ready || d->lu_n != n)
    return fail(O2D_ERR_STATE, "no LU factors of size %lld (factor first)", (long long)n);
  rocblas_handle h;
  O2D_TRY(blas_of(d, s, &h));
  if (!d->inv_ready && d->lu_solves >= kSolvesBeforeInverse) {
    void* inv;
    O2D_TRY(slot(d, kInv, sizeof(double) * n * n, s, &inv));
    const int64_t nn = n * n;
    const unsigned grid = (unsigned)std::min<int64_t>((nn + kBlock - 1) / kBlock, 65536);
    identity_kernel<<<grid, kBlock, 0, s>>>(n, static_cast<double*>(inv));
    O2D_LAUNCH_CHECK();
    if (rocsolver_dgetrs(h, rocblas_operation_none, (rocblas_int)n, (rocblas_int)n,
                         static_cast<double*>(d->buf[kLU]), (rocblas_int)n,
                         static_cast<const rocblas_int*>(d->buf[kPivots]),
                         static_cast<double*>(inv), (rocblas_int)n) != rocblas_status_success)
      return fail(O2D_ERR_HIP, "rocsolver_dgetrs failed");
    d->inv_ready = true;
  }
  ++d->lu_solves;
  if (d->inv_ready) {
    const double* src = b;
    if (x == b) {  // dgemv needs distinct input and output vectors
      void* t;
      O2D_TRY(slot(d, kSolveTmp, sizeof(double) * n, s, &t));
      O2D_HIP(hipMemcpyAsync(t, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
      src = static_cast<const double*>(t);
    }
    const double one = 1.0, zero = 0.0;
    if (rocblas_set_pointer_mode(h, rocblas_pointer_mode_host) != rocblas_status_success ||
        rocblas_dgemv(h, rocblas_operation_none, (rocblas_int)n, (rocblas_int)n, &one,
                      static_cast<const double*>(d->buf[kInv]), (rocblas_int)n, src, 1, &zero, x,
                      1) != rocblas_status_success)
      return fail(O2D_ERR_HIP, "rocblas_dgemv failed");
    return 0;
  }
  if (x != b) O2D_HIP(hipMemcpyAsync(x, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  if (rocsolver_dgetrs(h, rocblas_operation_none, (rocblas_int)n, 1, static_cast<double*>(d->buf[kLU]),
                       (rocblas_int)n, static_cast<const rocblas_int*>(d->buf[kPivots]), x,
                       (rocblas_int)n) != rocblas_status_success)
    return fail(O2D_ERR_HIP, "rocsolver_dgetrs failed");
  return 0;
}

static int check_bem(const o2d_env* env, int64_t nnodes, const void* nx, const void* ny,
                     int64_t np, const void* idx, int32_t nunk, int32_t augmented,
                     const void* uu, const void* uv, const void* A, int64_t lda) {
  if (!env) return fail(O2D_ERR_ARG, "env is NULL");
  if (nunk != 1 && nunk != 2) return fail(O2D_ERR_RESULTS, "nunk must be 1 or 2 (got %d)", nunk);
  if (np < 0 || nnodes < 0 || np > 0x3fffffffLL) return fail(O2D_ERR_SIZE, "bad panel count");
  const int64_t n = np * nunk + (augmented ? 1 : 0);
  if (lda < n) return fail(O2D_ERR_SIZE, "lda %lld < n %lld", (long long)lda, (long long)n);
  if (np > 0 && (!nx || !ny || !idx || !A)) return fail(O2D_ERR_ARG, "NULL array");
  if (np > 0 && augmented && (!uu || !uv)) return fail(O2D_ERR_ARG, "augmented needs urot_u/urot_v");
  return 0;
}

}  // namespace o2d

using namespace o2d;

extern "C" {

int o2d_hip_bem_assemble_dev_d(const o2d_env* env, int64_t nnodes, const double* nx,
                               const double* ny, int64_t np, const uint32_t* idx, int32_t nunk,
                               int32_t augmented, const double* urot_u, const double* urot_v,
                               double vol, double* A, int64_t lda) {
  clear_error();
  O2D_TRY(check_bem(env, nnodes, nx, ny, np, idx, nunk, augmented, urot_u, urot_v, A, lda));
  Device* d;
  O2D_TRY(device(env->device, &d));
  return bem_assemble_dev(d, dev_stream(env), nnodes, nx, ny, np, idx, nunk, augmented, urot_u,
                          urot_v, vol, A, lda);
}

int o2d_hip_bem_assemble_d(const o2d_env* env, int64_t nnodes, const double* nx,
                           const double* ny, int64_t np, const uint32_t* idx, int32_t nunk,
                           int32_t augmented, const double* urot_u, const double* urot_v,
                           double vol, double* A, int64_t lda) {
  clear_error();
  O2D_TRY(check_bem(env, nnodes, nx, ny, np, idx, nunk, augmented, urot_u, urot_v, A, lda));
  if (np == 0) return 0;
  for (int64_t j = 0; j < 2 * np; ++j)
    if ((int64_t)idx[j] >= nnodes) return fail(O2D_ERR_SIZE, "panel node index out of range");
  Device* d;
  O2D_TRY(device(env->device, &d));
  hipStream_t s = stream_of(env, d);
  const int64_t n = np * nunk + (augmented ? 1 : 0);
  double *dnx, *dny, *duu = nullptr, *duv = nullptr;
  uint32_t* didx;
  O2D_TRY(upload<double>(d, kPanX, nx, nnodes, s, &dnx));
  O2D_TRY(upload<double>(d, kPanY, ny, nnodes, s, &dny));
  O2D_TRY(upload<uint32_t>(d, kPanIdx, idx, 2 * np, s, &didx));
  if (augmented) {
    O2D_TRY(upload<double>(d, kPanVs, urot_u, np, s, &duu));
    O2D_TRY(upload<double>(d, kPanSs, urot_v, np, s, &duv));
  }
  void* dA;
  O2D_TRY(slot(d, kMatA, sizeof(double) * lda * n, s, &dA));
  O2D_TRY(bem_assemble_dev(d, s, nnodes, dnx, dny, np, didx, nunk, augmented, duu, duv, vol,
                           static_cast<double*>(dA), lda));
  O2D_HIP(hipMemcpy2DAsync(A, sizeof(double) * lda, dA, sizeof(double) * lda, sizeof(double) * n, n,
                           hipMemcpyDeviceToHost, s));
  O2D_HIP(hipStreamSynchronize(s));
  return 0;
}

static int check_pair(const o2d_env* env, int64_t snn, const void* snx, const void* sny,
                      int64_t snp, const void* sidx, int64_t tnn, const void* tnx,
                      const void* tny, int64_t tnp, const void* tidx, int32_t nunk,
                      int32_t src_aug, int32_t targ_aug, const void* uu, const void* uv,
                      const void* A, int64_t lda) {
  if (!env) return fail(O2D_ERR_ARG, "env is NULL");
  if (nunk != 1 && nunk != 2) return fail(O2D_ERR_RESULTS, "nunk must be 1 or 2 (got %d)", nunk);
  if (snp < 0 || tnp < 0 || snn < 0 || tnn < 0 || snp > 0x3fffffffLL || tnp > 0x3fffffffLL)
    return fail(O2D_ERR_SIZE, "bad panel count");
  if (lda < tnp * nunk + (targ_aug ? 1 : 0)) return fail(O2D_ERR_SIZE, "lda too small");
  if ((snp > 0 && (!snx || !sny || !sidx)) || (tnp > 0 && (!tnx || !tny || !tidx)) || !A)
    return fail(O2D_ERR_ARG, "NULL array");
  if (src_aug && tnp > 0 && (!uu || !uv)) return fail(O2D_ERR_ARG, "src_aug needs urot_u/urot_v");
  return 0;
}

int o2d_hip_bem_assemble_pair_dev_d(const o2d_env* env, int64_t snn, const double* snx,
                                    const double* sny, int64_t snp, const uint32_t* sidx,
                                    int64_t tnn, const double* tnx, const double* tny,
                                    int64_t tnp, const uint32_t* tidx, int32_t nunk,
                                    int32_t src_aug, int32_t targ_aug, const double* urot_u,
                                    const double* urot_v, double* A, int64_t lda) {
  clear_error();
  O2D_TRY(check_pair(env, snn, snx, sny, snp, sidx, tnn, tnx, tny, tnp, tidx, nunk, src_aug,
                     targ_aug, urot_u, urot_v, A, lda));
  Device* d;
  O2D_TRY(device(env->device, &d));
  return bem_assemble_pair_dev(d, dev_stream(env), snx, sny, snp, sidx, tnx, tny, tnp, tidx,
                               nunk, src_aug, targ_aug, urot_u, urot_v, A, lda);
}

int o2d_hip_bem_assemble_pair_d(const o2d_env* env, int64_t snn, const double* snx,
                                const double* sny, int64_t snp, const uint32_t* sidx,
                                int64_t tnn, const double* tnx, const double* tny, int64_t tnp,
                                const uint32_t* tidx, int32_t nunk, int32_t src_aug,
                                int32_t targ_aug, const double* urot_u, const double* urot_v,
                                double* A, int64_t lda) {
  clear_error();
  O2D_TRY(check_pair(env, snn, snx, sny, snp, sidx, tnn, tnx, tny, tnp, tidx, nunk, src_aug,
                     targ_aug, urot_u, urot_v, A, lda));
  for (int64_t j = 0; j < 2 * snp; ++j)
    if ((int64_t)sidx[j] >= snn) return fail(O2D_ERR_SIZE, "source node index out of range");
  for (int64_t j = 0; j < 2 * tnp; ++j)
    if ((int64_t)tidx[j] >= tnn) return fail(O2D_ERR_SIZE, "target node index out of range");
  if (snp == 0 || tnp == 0) return 0;
  Device* d;
  O2D_TRY(device(env->device, &d));
  hipStream_t s = stream_of(env, d);
  const int64_t nr = tnp * nunk + (targ_aug ? 1 : 0);
  const int64_t nc = snp * nunk + (src_aug ? 1 : 0);
  double *dsx, *dsy, *dtx, *dty, *duu = nullptr, *duv = nullptr;
  uint32_t *dsi, *dti;
  O2D_TRY(upload<double>(d, kPanX, snx, snn, s, &dsx));
  O2D_TRY(upload<double>(d, kPanY, sny, snn, s, &dsy));
  O2D_TRY(upload<uint32_t>(d, kPanIdx, sidx, 2 * snp, s, &dsi));
  O2D_TRY(upload<double>(d, kPan2X, tnx, tnn, s, &dtx));
  O2D_TRY(upload<double>(d, kPan2Y, tny, tnn, s, &dty));
  O2D_TRY(upload<uint32_t>(d, kPan2Idx, tidx, 2 * tnp, s, &dti));
  if (src_aug) {
    O2D_TRY(upload<double>(d, kPanVs, urot_u, tnp, s, &duu));
    O2D_TRY(upload<double>(d, kPanSs, urot_v, tnp, s, &duv));
  }
  void* dA;
  O2D_TRY(slot(d, kMatA, sizeof(double) * lda * nc, s, &dA));
  O2D_TRY(bem_assemble_pair_dev(d, s, dsx, dsy, snp, dsi, dtx, dty, tnp, dti, nunk, src_aug,
                                targ_aug, duu, duv, static_cast<double*>(dA), lda));
  O2D_HIP(hipMemcpy2DAsync(A, sizeof(double) * lda, dA, sizeof(double) * lda,
                           sizeof(double) * nr, nc, hipMemcpyDeviceToHost, s));
  O2D_HIP(hipStreamSynchronize(s));
  return 0;
}

int o2d_hip_bem_factor_dev_d(const o2d_env* env, int64_t n, const double* A, int64_t lda) {
  clear_error();
  if (!env) return fail(O2D_ERR_ARG, "env is NULL");
  if (n <= 0 || n > 0x7fffffffLL || lda < n) return fail(O2D_ERR_SIZE, "bad n/lda");
  if (!A) return fail(O2D_ERR_ARG, "NULL matrix");
  Device* d;
  O2D_TRY(device(env->device, &d));
  return bem_factor_dev(d, dev_stream(env), n, A, lda);
}

int o2d_hip_bem_factor_d(const o2d_env* env, int64_t n, const double* A, int64_t lda) {
  clear_error();
  if (!env) return fail(O2D_ERR_ARG, "env is NULL");
  if (n <= 0 || n > 0x7fffffffLL || lda < n) return fail(O2D_ERR_SIZE, "bad n/lda");
  if (!A) return fail(O2D_ERR_ARG, "NULL matrix");
  Device* d;
  O2D_TRY(device(env->device, &d));
  hipStream_t s = stream_of(env, d);
  void* dA;
  O2D_TRY(slot(d, kMatA, sizeof(double) * n * n, s, &dA));
  O2D_HIP(hipMemcpy2DAsync(dA, sizeof(double) * n, A, sizeof(double) * lda, sizeof(double) * n, n,
                           hipMemcpyHostToDevice, s));
  return bem_factor_dev(d, s, n, static_cast<double*>(dA), n);
}

int o2d_hip_bem_factors_d(const o2d_env* env, int64_t n, double* lu, int32_t* ipiv) {
  clear_error();
  if (!env) return fail(O2D_ERR_ARG, "env is NULL");
  if (!lu || !ipiv) return fail(O2D_ERR_ARG, "NULL output");
  Device* d;
  O2D_TRY(device(env->device, &d));
  if (!d->lu_ready || d->lu_n != n)
    return fail(O2D_ERR_STATE, "no LU factors of size %lld (factor first)", (long long)n);
  hipStream_t s = stream_of(env, d);
  O2D_HIP(hipMemcpyAsync(lu, d->buf[kLU], sizeof(double) * n * n, hipMemcpyDeviceToHost, s));
  O2D_HIP(hipMemcpyAsync(ipiv, d->buf[kPivots], sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
  O2D_HIP(hipStreamSynchronize(s));
  return 0;
}

int o2d_hip_bem_solve_dev_d(const o2d_env* env, int64_t n, const double* b, double* x) {
  clear_error();
  if (!env) return fail(O2D_ERR_ARG, "env is NULL");
  if (!b || !x) return fail(O2D_ERR_ARG, "NULL vector");
  Device* d;
  O2D_TRY(device(env->device, &d));
  return bem_solve_dev(d, dev_stream(env), n, b, x);
}

int o2d_hip_bem_solve_d(const o2d_env* env, int64_t n, const double* b, double* x) {
  clear_error();
  if (!env) return fail(O2D_ERR_ARG, "env is NULL");
  if (!b || !x) return fail(O2D_ERR_ARG, "NULL vector");
  if (n <= 0) return fail(O2D_ERR_SIZE, "bad n");
  Device* d;
  O2D_TRY(device(env->device, &d));
  hipStream_t s = stream_of(env, d);
  double* db;
  O2D_TRY(upload<double>(d, kRhs, b, n, s, &db));
  O2D_TRY(bem_solve_dev(d, s, n, db, db));
  O2D_TRY(download<double>(x, db, n, s));
  O2D_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // extern "C"
