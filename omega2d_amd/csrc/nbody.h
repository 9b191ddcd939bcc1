// nbody.h - internal interfaces between the kernel modules and the C ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "o2d_runtime.h"

namespace o2d {

// particles on points, device pointers, enqueued on s (tw is the vorticity or shear output);
// temporaries come from the device's buffer pool, or from *ws when it is given
int nbody_dev(Device* d, hipStream_t s, int core, int prec, int res, int64_t ns, const double* sx,
              const double* sy, const double* ss, const double* sr, int64_t nt, const double* tx,
              const double* ty, const double* tr, double* tu, double* tv, double* tw,
              Workspace* ws = nullptr);

// argument checks of the particles-on-points entry points (counts, NULL arrays, outputs the
// result type needs)
int check_pp(const o2d_env* env, int64_t ns, const void* sx, const void* sy, const void* ss,
             const void* sr, int64_t nt, const void* tx, const void* ty, const void* tu,
             const void* tv, const void* tw, const void* te);

// 32-bit Morton keys and stable order (bbox: derive the box on device from x, y)
int morton_order(Device* d, hipStream_t s, int64_t n, const double* x, const double* y,
                 bool bbox, double x0, double y0, double inv, uint32_t* keys_out,
                 uint32_t* order_out, Workspace* ws = nullptr);

int panels_on_points_dev(Device* d, hipStream_t s, int64_t nnodes, const double* nx,
                         const double* ny, int64_t np, const uint32_t* idx, const double* vs,
                         const double* ss, int64_t nt, const double* tx, const double* ty,
                         double* tu, double* tv);

int points_on_panels_dev(Device* d, hipStream_t s, int64_t ns, const double* sx, const double* sy,
                         const double* ss, int64_t nnodes, const double* nx, const double* ny,
                         int64_t np, const uint32_t* idx, double* pu, double* pv);

int bem_assemble_dev(Device* d, hipStream_t s, int64_t nnodes, const double* nx, const double* ny,
                     int64_t np, const uint32_t* idx, int nunk, int augmented,
                     const double* urot_u, const double* urot_v, double vol, double* A,
                     int64_t lda);

int bem_factor_dev(Device* d, hipStream_t s, int64_t n, const double* A, int64_t lda);

// blocked LU with partial pivoting (lu.hip), n <= lu_max_order(); ipiv 1-based, info device int
int64_t lu_max_order();
int lu_factor_dev(Device* d, hipStream_t s, int n, double* A, int lda, int* ipiv, int* info);
int bem_solve_dev(Device* d, hipStream_t s, int64_t n, const double* b, double* x);

int rk_move_dev(Device* d, hipStream_t s, int64_t n, double dt, int nstage, const double* w,
                const double* const* u, const double* const* v, double* x, double* y);

int adjoint_dev(Device* d, hipStream_t s, int core, int64_t ns, const double* sx,
                const double* sy, const double* ss, const double* sr, int64_t nt,
                const double* tx, const double* ty, const double* gu, const double* gv,
                double* gx, double* gy, double* gs, double* gr);

// Gaussian vorticity of particles at targets (fp64, +=), device pointers
int vorticity_dev(Device* d, hipStream_t s, int64_t ns, const double* sx, const double* sy,
                  const double* ss, const double* sr, int64_t nt, const double* tx,
                  const double* ty, double* tw);

// particle/panel proximity (mode 0 nearest distances, 1 reflect, 2 clear), device pointers
int near_dev(Device* d, hipStream_t s, int mode, int method, double cutoff_mult, double ips,
             int64_t nnodes, const double* nx, const double* ny, int64_t np,
             const uint32_t* idx, const double* pnx, const double* pny, int64_t nt, double* tx,
             double* ty, double* ts, const double* tr, double* dist, double* value,
             int64_t* counts);

// host <-> device staging for the synchronous entry points
template <class T>
int upload(Device* d, Slot sl, const T* h, int64_t n, hipStream_t s, T** out) {
  void* p;
  O2D_TRY(slot(d, sl, sizeof(T) * (n > 0 ? n : 1), s, &p));
  if (n > 0) O2D_HIP(hipMemcpyAsync(p, h, sizeof(T) * n, hipMemcpyHostToDevice, s));
  *out = static_cast<T*>(p);
  return 0;
}

template <class T>
int download(T* h, const T* dptr, int64_t n, hipStream_t s) {
  if (n > 0) O2D_HIP(hipMemcpyAsync(h, dptr, sizeof(T) * n, hipMemcpyDeviceToHost, s));
  return 0;
}

// float host array -> double device array (through a float staging slot)
int upload_f2d(Device* d, Slot stage, Slot dst, const float* h, int64_t n, hipStream_t s,
               double** out);
// double device array -> float host array (through a float staging slot)
int download_d2f(Device* d, Slot stage, float* h, const double* dptr, int64_t n, hipStream_t s);

}  // namespace o2d
