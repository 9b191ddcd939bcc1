/*
 * o2d_oracle.h - CPU restatement of the Omega2D direct Biot-Savart path, in fp64.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, called from tests/, from
 * __graft_entry__.smoke() and from bench.py's cpu_baseline leg. The product never links it.
 *
 * Every function follows a reference routine (file:line in o2d_oracle.c) and is pinned by
 * the checks listed there; see tests/test_oracle.py.
 */
#ifndef O2D_ORACLE_H
#define O2D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* core and result enums share their values with include/omega2d_hip.h */
int oc_points_on_points(int core, int restype, int64_t ns, const double* sx, const double* sy,
                        const double* ss, const double* sr, int64_t nt, const double* tx,
                        const double* ty, const double* tr, double* tu, double* tv, double* tw,
                        double* te);

int oc_panels_on_points(int64_t nnodes, const double* nx, const double* ny, int64_t np,
                        const uint32_t* idx, const double* vs, const double* ss, int64_t nt,
                        const double* tx, const double* ty, double* tu, double* tv);

int oc_points_on_panels(int64_t ns, const double* sx, const double* sy, const double* ss,
                        int64_t nnodes, const double* nx, const double* ny, int64_t np,
                        const uint32_t* idx, double* pu, double* pv);

int oc_panel_geometry(int64_t nnodes, const double* nx, const double* ny, int64_t np,
                      const uint32_t* idx, double* tang_x, double* tang_y, double* norm_x,
                      double* norm_y, double* area);

int oc_body_volume(int64_t nnodes, const double* nx, const double* ny, int64_t np,
                   const uint32_t* idx, double* vol, double* cx, double* cy);

int oc_unit_rot_strengths(int64_t nnodes, const double* nx, const double* ny, int64_t np,
                          const uint32_t* idx, double cx, double cy, double* vs, double* ss);

int oc_bem_assemble(int64_t nnodes, const double* nx, const double* ny, int64_t np,
                    const uint32_t* idx, int nunk, int augmented, const double* urot_u,
                    const double* urot_v, double vol, double* A, int64_t lda);

/* the block of a source surface acting on a different target surface (no self term) */
int oc_bem_assemble_pair(int64_t snn, const double* snx, const double* sny, int64_t snp,
                         const uint32_t* sidx, int64_t tnn, const double* tnx, const double* tny,
                         int64_t tnp, const uint32_t* tidx, int nunk, int src_aug, int targ_aug,
                         const double* urot_u, const double* urot_v, double* A, int64_t lda);

int oc_vels_to_rhs(int64_t np, int nunk, const double* pu, const double* pv,
                   const double* tang_x, const double* tang_y, const double* norm_bc,
                   double* rhs);

int oc_finalize(int64_t n, double* u, double* v, double fsx, double fsy);

int oc_rk_combine(int64_t n, int nstage, const double* w, const double* const* us,
                  const double* const* vs, double dt, double* x, double* y);

int oc_points_on_points_adjoint(int core, int64_t ns, const double* sx, const double* sy,
                                const double* ss, const double* sr, int64_t nt, const double* tx,
                                const double* ty, const double* gu, const double* gv,
                                double* gx, double* gy, double* gs, double* gr);

int oc_morton_keys(int64_t n, const double* x, const double* y, double x0, double y0,
                   double inv_extent, uint32_t* keys);

/* src/Reflect.h: nearest distance from each target to a set of panels, reflection of targets
 * that lie behind the nearest panel, and the near-wall clearing pass (method 1 pushes targets
 * out to the cutoff layer, method 0 trims strength with the cut tables). pnx/pny are the
 * panels' unit normals (Surfaces::get_norm). */
int oc_nearest_distances(int64_t nnodes, const double* nx, const double* ny, int64_t np,
                         const uint32_t* idx, int64_t nt, const double* tx, const double* ty,
                         double* dist);

int oc_reflect(int64_t nnodes, const double* nx, const double* ny, int64_t np,
               const uint32_t* idx, const double* pnx, const double* pny, int64_t nt, double* tx,
               double* ty, int64_t* num_reflected);

int oc_cut_table(double dx, int cap, double* pos, double* smult, double* dshift);

int oc_clear_inner(int method, int64_t nnodes, const double* nx, const double* ny, int64_t np,
                   const uint32_t* idx, const double* pnx, const double* pny, double cutoff_mult,
                   double ips, int64_t nt, double* tx, double* ty, double* ts, const double* tr,
                   int max_passes, double* circ_removed, int64_t* num_cropped,
                   int64_t* unsettled);

/* src/InfluenceVort.h:47-126: unscaled Gaussian vorticity of particles at targets (+=) */
int oc_points_vorticity(int64_t ns, const double* sx, const double* sy, const double* ss,
                        const double* sr, int64_t nt, const double* tx, const double* ty,
                        double* tw);

#ifdef __cplusplus
}
#endif

#endif
