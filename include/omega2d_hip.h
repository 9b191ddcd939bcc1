/*
 * omega2d_hip.h - C ABI of the MI355X (gfx950) direct Biot-Savart library.
 *
 * Plain pointers and sizes only. Every entry point returns 0 on success or a negative
 * O2D_ERR_* code; o2d_hip_last_error_string() describes the last failure on the calling thread.
 *
 * Conventions kept from the reference (src/Influence.h, src/ElementBase.h:200-234):
 *  - velocity sums are UNSCALED (no 1/(2 pi), no freestream) and ACCUMULATED (+=) into
 *    caller-zeroed outputs; the caller finalizes with u = fs + u/(2 pi);
 *  - sources are const; the self pair (d = 0) contributes exactly 0;
 *  - one caller thread per device (the reference is single-caller).
 *
 * Two families of entry points:
 *  - o2d_hip_*_d / _f: host arrays, synchronous (results are in the host arrays on return).
 *    _d takes double storage, _f float storage (the reference application's STORE=float);
 *    both compute in the precision env->precision selects. Device mirrors are kept by the
 *    library per device and reused across calls.
 *  - o2d_hip_*_dev_d: device pointers, enqueued on env->stream (asynchronous). The library's
 *    scratch buffers (sort keys, partial sums, staged panels, LU factors) are per device and
 *    shared by every call on that device, so device-pointer calls on different streams of one
 *    device must not overlap in time: order them (one stream, or events between streams).
 */
#ifndef OMEGA2D_HIP_H
#define OMEGA2D_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define O2D_HIP_ABI_VERSION 1

/* status codes */
#define O2D_OK 0
#define O2D_ERR_ARG (-1)       /* a required pointer is NULL */
#define O2D_ERR_SIZE (-2)      /* negative/too large count, or a panel index out of range */
#define O2D_ERR_CORE (-3)      /* unknown core function */
#define O2D_ERR_RESULTS (-4)   /* unsupported result type or option */
#define O2D_ERR_PRECISION (-5) /* unknown precision form */
#define O2D_ERR_HIP (-6)       /* HIP runtime or library failure */
#define O2D_ERR_STATE (-7)     /* e.g. solve before factor */

/* regularised velocity cores, in the order of src/CoreFunc.h:35-39 */
#define O2D_CORE_RM 1       /* Rosenhead-Moore            src/CoreFunc.h:42-82  */
#define O2D_CORE_GAUSSIAN 2 /* exponential (Gaussian)     src/CoreFunc.h:85-181 */
#define O2D_CORE_WL 3       /* Winckelmans-Leonard        src/CoreFunc.h:184-230 */
#define O2D_CORE_V2 4       /* Vatistas n=2 (ref default) src/CoreFunc.h:233-275 */
#define O2D_CORE_V3 5       /* Vatistas n=3               src/CoreFunc.h:278-322 */

/* result sets: the values of results_t, src/ResultsType.h:26-32 */
#define O2D_VELONLY 1
#define O2D_VELANDVORT 4
#define O2D_VELANDSHEAR 5

/* precision forms */
#define O2D_PLAIN_FP64 1 /* fp64 pair arithmetic and accumulation */
#define O2D_SPLIT_FP32 2 /* fp32 pair arithmetic on coordinates rebased to each source tile's
                            fp64 centroid, fp64 accumulation per target; all result types */

/* execution environment: the gpu_hip counterpart of ExecEnv (src/ExecEnv.h:45-108) plus the
 * run-time choices the reference makes at compile time */
typedef struct o2d_env {
  int32_t core;      /* O2D_CORE_* */
  int32_t precision; /* O2D_PLAIN_FP64 | O2D_SPLIT_FP32 */
  int32_t results;   /* O2D_VELONLY | O2D_VELANDVORT | O2D_VELANDSHEAR */
  int32_t device;    /* HIP device ordinal */
  void* stream;      /* hipStream_t. Device-pointer entry points enqueue on exactly this stream
                        (NULL = the device's null stream); the synchronous host-array entry
                        points use it when non-NULL, else a stream of the library's own */
} o2d_env;

/* ---- runtime -------------------------------------------------------------------------- */

/* Select the device, create the library stream and buffer pool for it. Optional: every entry
 * point initialises its device on first use. */
int o2d_hip_init(int32_t device);
/* Release every device's buffers, LU factors and streams. */
int o2d_hip_release(void);
const char* o2d_hip_last_error_string(void);
int32_t o2d_hip_abi_version(void);
/* Algorithmic flops per pair interaction in the reference's own convention
 * (src/Kernels.h:81,95,312,362, src/CoreFunc.h); panel = 0 for point sources, 1 for vortex
 * panels (kernelu_1v_0p), 2 for vortex+source panels (kernelu_1vs_0p). */
int32_t o2d_hip_flops_per_pair(int32_t core, int32_t point_targets, int32_t panel);

/* ---- kernel timing ------------------------------------------------------------------------- */

/* Kernel classes timed by o2d_hip_timing_*: the main kernel of each entry point. */
#define O2D_KT_NBODY_SPLIT 1   /* split fp32 particles-on-points kernel, particle targets */
#define O2D_KT_NBODY_PLAIN 2   /* plain fp64 particles-on-points kernel, particle targets */
#define O2D_KT_PANELS 3        /* panels on points */
#define O2D_KT_PANEL_RHS 4     /* points on panels */
#define O2D_KT_BEM_ASSEMBLE 5  /* influence-matrix assembly */
#define O2D_KT_ADJOINT 6       /* adjoint kernel */
#define O2D_KT_SPLIT_PREP 7    /* Morton order + rebasing of the split form */
#define O2D_KT_NBODY_SPLIT_PT 8 /* split fp32 kernel, point targets */
#define O2D_KT_NBODY_PLAIN_PT 9 /* plain fp64 kernel, point targets */
#define O2D_KT_NEAR 10          /* particle/panel proximity scan (distances, reflect, clear) */
#define O2D_KT_VORT 11          /* Gaussian vorticity of particles */
#define O2D_KT_TREE_BUILD 12    /* treecode: Morton orders, leaves, moments (all build kernels) */
#define O2D_KT_TREE_EVAL 13     /* treecode: tree walk and evaluation */
#define O2D_KT_COUNT 14

/* enable = 1: from now on record a HIP event pair on the launch stream around every timed
 * kernel launch (results are unaffected); enable = 0: stop recording. */
int o2d_hip_timing_enable(int32_t enable);
/* Total milliseconds and number of launches recorded for a kernel class since the last read
 * (waits for the recorded events, then clears the class). */
int o2d_hip_timing_read(int32_t kernel_class, double* total_ms, int64_t* launches);

/* ---- particles on points: points_affect_points (src/Influence.h:56-416) ------------------ */

/* Sources: ns particles (sx, sy, strength sstr, core radius srad).
 * Targets: nt points (tx, ty); trad = NULL means point targets (inert: kernelu_0v_0p),
 * otherwise particle targets with core radii (kernelu_0v_0b).
 * tu, tv accumulate the unscaled velocity; tw (vorticity, results = VELANDVORT) and te (shear
 * rate, results = VELANDSHEAR) accumulate likewise and may be NULL otherwise.
 * Replaces the internal CPU loops and the external_vel_solver_f_/_tr_f_/_d_ C ABI
 * (src/Influence.h:31-38, 74-98) for env->precision's form. */
int o2d_hip_points_on_points_d(const o2d_env* env, int64_t ns, const double* sx, const double* sy,
                               const double* sstr, const double* srad, int64_t nt,
                               const double* tx, const double* ty, const double* trad,
                               double* tu, double* tv, double* tw, double* te);
int o2d_hip_points_on_points_f(const o2d_env* env, int64_t ns, const float* sx, const float* sy,
                               const float* sstr, const float* srad, int64_t nt, const float* tx,
                               const float* ty, const float* trad, float* tu, float* tv,
                               float* tw, float* te);
int o2d_hip_points_on_points_dev_d(const o2d_env* env, int64_t ns, const double* sx,
                                   const double* sy, const double* sstr, const double* srad,
                                   int64_t nt, const double* tx, const double* ty,
                                   const double* trad, double* tu, double* tv, double* tw,
                                   double* te);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the same sum as
 * o2d_hip_points_on_points_dev_d (same arguments, both precision forms, every core and result
 * type) with its temporaries in a fixed 64 MiB array per device instead of the library's
 * buffer pool; it calls no hipMalloc, hipFree or allocating library routine. At most 262,144
 * sources and 262,144 targets (O2D_ERR_SIZE beyond). Calls that overlap on one device race on
 * that array. */
int o2d_hip_points_on_points_noalloc_dev_d(const o2d_env* env, int64_t ns, const double* sx,
                                           const double* sy, const double* sstr,
                                           const double* srad, int64_t nt, const double* tx,
                                           const double* ty, const double* trad, double* tu,
                                           double* tv, double* tw, double* te);

/* ---- panels on points: panels_affect_points (src/Influence.h:422-614) -------------------- */

/* Panels: nnodes node coordinates (nx, ny), np panels whose endpoints are nodes
 * idx[2j], idx[2j+1]; vortex sheet strength vstr[j], source strength sstr[j] per unit length
 * (sstr = NULL: vortex-only panels, kernelu_1v_0p; else kernelu_1vs_0p). Unscaled, accumulated.
 * Velocity only (env->results must be O2D_VELONLY); computed in fp64 whatever env->precision. */
int o2d_hip_panels_on_points_d(const o2d_env* env, int64_t nnodes, const double* nx,
                               const double* ny, int64_t np, const uint32_t* idx,
                               const double* vstr, const double* sstr, int64_t nt,
                               const double* tx, const double* ty, double* tu, double* tv);
int o2d_hip_panels_on_points_dev_d(const o2d_env* env, int64_t nnodes, const double* nx,
                                   const double* ny, int64_t np, const uint32_t* idx,
                                   const double* vstr, const double* sstr, int64_t nt,
                                   const double* tx, const double* ty, double* tu, double* tv);

/* ---- particles on panels: points_affect_panels (src/Influence.h:633-771), the BEM RHS ----- */

/* Panel-averaged velocity of singular point vortices (radii are ignored, as in the reference):
 * pu[i] -= (1/L_i) sum_j kernelu_1v_0p(panel i, sstr[j], particle j). Unscaled, accumulated. */
int o2d_hip_points_on_panels_d(const o2d_env* env, int64_t ns, const double* sx,
                               const double* sy, const double* sstr, int64_t nnodes,
                               const double* nx, const double* ny, int64_t np,
                               const uint32_t* idx, double* pu, double* pv);
int o2d_hip_points_on_panels_dev_d(const o2d_env* env, int64_t ns, const double* sx,
                                   const double* sy, const double* sstr, int64_t nnodes,
                                   const double* nx, const double* ny, int64_t np,
                                   const uint32_t* idx, double* pu, double* pv);

/* ---- BEM: panels_on_panels_coeff (src/Coefficients.h:155-573) + solve (src/BEM.h:116-234) */

/* Self-influence block of one body, fp64, written column-major into A (n x n, leading
 * dimension lda >= n) with n = np*nunk + (augmented ? 1 : 0). nunk = 1: vortex unknowns only;
 * nunk = 2: interleaved [vortex, source] unknowns per panel. When augmented, urot_u/urot_v are
 * the finalized panel-centre velocities induced by the body's unit-rotation strengths
 * (src/BEMHelper.h:203-210) and vol the body's signed area (src/Surfaces.h:623-677). */
int o2d_hip_bem_assemble_d(const o2d_env* env, int64_t nnodes, const double* nx,
                           const double* ny, int64_t np, const uint32_t* idx, int32_t nunk,
                           int32_t augmented, const double* urot_u, const double* urot_v,
                           double vol, double* A, int64_t lda);
int o2d_hip_bem_assemble_dev_d(const o2d_env* env, int64_t nnodes, const double* nx,
                               const double* ny, int64_t np, const uint32_t* idx, int32_t nunk,
                               int32_t augmented, const double* urot_u, const double* urot_v,
                               double vol, double* A, int64_t lda);
/* The block of a source body's panels acting on a DIFFERENT target body's panels
 * (panels_on_panels_coeff with &src != &targ, src/Coefficients.h:155-573): rows are the target's
 * equations (tnp * nunk, plus 1 if targ_aug), columns the source's unknowns (snp * nunk, plus 1 if
 * src_aug), column-major with leading dimension lda, scaled by 1/(2 pi), two-way averaged and
 * with no self term. targ_aug adds a bottom row of zeros (the lengths row counts only a body's own
 * panels); src_aug adds the right column from urot_u/urot_v, the velocity of the source body
 * rotating at unit rate on the target panels (computed by the caller as solve_bem does,
 * src/BEMHelper.h:202-210), and a zero corner. Both bodies use the same nunk. */
int o2d_hip_bem_assemble_pair_d(const o2d_env* env, int64_t snn, const double* snx,
                                const double* sny, int64_t snp, const uint32_t* sidx,
                                int64_t tnn, const double* tnx, const double* tny, int64_t tnp,
                                const uint32_t* tidx, int32_t nunk, int32_t src_aug,
                                int32_t targ_aug, const double* urot_u, const double* urot_v,
                                double* A, int64_t lda);
int o2d_hip_bem_assemble_pair_dev_d(const o2d_env* env, int64_t snn, const double* snx,
                                    const double* sny, int64_t snp, const uint32_t* sidx,
                                    int64_t tnn, const double* tnx, const double* tny,
                                    int64_t tnp, const uint32_t* tidx, int32_t nunk,
                                    int32_t src_aug, int32_t targ_aug, const double* urot_u,
                                    const double* urot_v, double* A, int64_t lda);

/* LU-factor the n x n column-major fp64 matrix on the device and keep the factors for
 * o2d_hip_bem_solve_d (the reference re-initialises its solver only when A changes,
 * src/BEM.h:182-195). A is a host array (_d) or device array (_dev_d, not modified). */
int o2d_hip_bem_factor_d(const o2d_env* env, int64_t n, const double* A, int64_t lda);
int o2d_hip_bem_factor_dev_d(const o2d_env* env, int64_t n, const double* A, int64_t lda);
/* The stored factors to host memory in LAPACK getrf's layout: lu (n x n, column-major) holds
 * L (unit diagonal, not stored) and U, ipiv the 1-based row interchanges. No counterpart in the
 * reference, whose GMRES (src/BEM.h:182-202) keeps no factors: it lets a caller check the
 * factors and pivots against a library getrf. */
int o2d_hip_bem_factors_d(const o2d_env* env, int64_t n, double* lu, int32_t* ipiv);
/* x = A^-1 b with the stored factors (x may alias b). */
int o2d_hip_bem_solve_d(const o2d_env* env, int64_t n, const double* b, double* x);
int o2d_hip_bem_solve_dev_d(const o2d_env* env, int64_t n, const double* b, double* x);

/* ---- RK position update: ElementBase::move (src/ElementBase.h:276-359) -------------------- */

/* x += dt * sum_s w[s] * u_s (nstage = 1, 2 or 3), fp64. u and v are arrays of nstage
 * pointers to finalized stage velocities. Host arrays (_d) or device arrays (_dev_d). */
int o2d_hip_rk_move_d(const o2d_env* env, int64_t n, double dt, int32_t nstage, const double* w,
                      const double* const* u, const double* const* v, double* x, double* y);
int o2d_hip_rk_move_dev_d(const o2d_env* env, int64_t n, double dt, int32_t nstage,
                          const double* w, const double* const* u, const double* const* v,
                          double* x, double* y);

/* ---- adjoint (no reference counterpart) -------------------------------------------------- */

/* Vector-Jacobian product of the unscaled point-target velocity sum
 *   tu_k = -sum_j sstr_j K(d_kj^2; srad_j) dy_kj,  tv_k = sum_j sstr_j K dx_kj
 * given gu = dJ/dtu, gv = dJ/dtv at the nt point targets: accumulates dJ/dsx, dJ/dsy,
 * dJ/dsstr, dJ/dsrad into gx, gy, gs, gr. The interaction matrix is recomputed, never stored.
 * fp64 (env->precision must be O2D_PLAIN_FP64). */
int o2d_hip_points_on_points_adjoint_d(const o2d_env* env, int64_t ns, const double* sx,
                                       const double* sy, const double* sstr, const double* srad,
                                       int64_t nt, const double* tx, const double* ty,
                                       const double* gu, const double* gv, double* gx,
                                       double* gy, double* gs, double* gr);
int o2d_hip_points_on_points_adjoint_dev_d(const o2d_env* env, int64_t ns, const double* sx,
                                           const double* sy, const double* sstr,
                                           const double* srad, int64_t nt, const double* tx,
                                           const double* ty, const double* gu, const double* gv,
                                           double* gx, double* gy, double* gs, double* gr);

/* ---- treecode summation: summation_t barneshut (src/ExecEnv.h:26-31) ---------------------- */

/* Parameters of the O(N log N) particles-on-points sum. A cell of sources (radius r about its
 * centre c) acts on a target z through its complex moments when r <= theta |z - c| and the
 * core equals 1/d^2 for every pair to within core_tol (exactly, for the Gaussian, beyond its
 * compact cutoff); otherwise it is opened, and leaves of 64 sources are summed directly with
 * the exact core in fp64. The truncation error of a used cell is about (theta)^order of its
 * contribution. order: 1..32; theta in (0, 1); core_tol >= 0 (ignored by the Gaussian). */
typedef struct o2d_tree_opts {
  double theta;
  int32_t order;
  int32_t pad;
  double core_tol;
} o2d_tree_opts;

/* Same operator, arguments and accumulate semantics as o2d_hip_points_on_points_d (velocity
 * only: env->results must be O2D_VELONLY; env->precision is not used, the sum is fp64). Host
 * arrays (_d) or device arrays (_dev_d). */
int o2d_hip_points_on_points_tree_d(const o2d_env* env, const o2d_tree_opts* opt, int64_t ns,
                                    const double* sx, const double* sy, const double* sstr,
                                    const double* srad, int64_t nt, const double* tx,
                                    const double* ty, const double* trad, double* tu,
                                    double* tv);
int o2d_hip_points_on_points_tree_dev_d(const o2d_env* env, const o2d_tree_opts* opt,
                                        int64_t ns, const double* sx, const double* sy,
                                        const double* sstr, const double* srad, int64_t nt,
                                        const double* tx, const double* ty, const double* trad,
                                        double* tu, double* tv);

/* ---- Gaussian vorticity: points_affect_points_vorticity (src/InfluenceVort.h:47-126) ------ */

/* tw_i += sum_j 2 sstr_j exp(-d_ij^2 / srad_j^2) / srad_j^2 at the nt targets (unscaled, as the
 * reference's x86 loop; finalize_vels multiplies by 1/(2 pi)). Computed in fp64 whatever
 * env->precision says; env->core is not used (the reference always uses the Gaussian here).
 * Host arrays (_d) or device arrays (_dev_d). */
int o2d_hip_points_vorticity_d(const o2d_env* env, int64_t ns, const double* sx,
                               const double* sy, const double* sstr, const double* srad,
                               int64_t nt, const double* tx, const double* ty, double* tw);
int o2d_hip_points_vorticity_dev_d(const o2d_env* env, int64_t ns, const double* sx,
                                   const double* sy, const double* sstr, const double* srad,
                                   int64_t nt, const double* tx, const double* ty, double* tw);

/* ---- particle/panel proximity: src/Reflect.h ---------------------------------------------- */
/*
 * Panels are straight segments (node arrays nx, ny; idx = 2 node indices per panel, as in
 * the velocity entry points); pnx, pny are the panels' unit normals pointing into the fluid
 * as Surfaces::compute_bases stores them (src/Surfaces.h:682-725). Each target's nearest
 * element (a panel interior or a node) is found by a scan over panels in index order with the
 * reference's tie rule (a distance below nextafter(min, 0) replaces the set of nearest
 * elements, one below nextafter(min, +max) joins it), in IEEE double without fused
 * multiply-adds. A node's normal is the sum of the normals of the panels that share it.
 *
 * The host entry points check node indices and zero-length panels before launching (the
 * reference asserts on a zero-length panel, src/Reflect.h:131). The device entry points
 * count such panels into counts[2] instead (an out-of-range index is clamped, never read).
 */

/* Distance from each target to its nearest panel or node (get_nearest_distances,
 * src/Reflect.h:792-826); with no panels every distance is sqrt(DBL_MAX). */
int o2d_hip_nearest_distances_d(const o2d_env* env, int64_t nnodes, const double* nx,
                                const double* ny, int64_t np, const uint32_t* idx, int64_t nt,
                                const double* tx, const double* ty, double* dist);
int o2d_hip_nearest_distances_dev_d(const o2d_env* env, int64_t nnodes, const double* nx,
                                    const double* ny, int64_t np, const uint32_t* idx,
                                    int64_t nt, const double* tx, const double* ty,
                                    double* dist);

/* Move each target that lies behind its nearest element to its mirror image in front of it
 * (reflect_panp2, src/Reflect.h:167-311); tx, ty are updated in place. *num_reflected (host
 * entry, may be NULL) receives the count; the device entry writes counts[0] = reflected,
 * counts[1] = 0, counts[2] = bad panels (counts may be NULL). */
int o2d_hip_reflect_d(const o2d_env* env, int64_t nnodes, const double* nx, const double* ny,
                      int64_t np, const uint32_t* idx, const double* pnx, const double* pny,
                      int64_t nt, double* tx, double* ty, int64_t* num_reflected);
int o2d_hip_reflect_dev_d(const o2d_env* env, int64_t nnodes, const double* nx,
                          const double* ny, int64_t np, const uint32_t* idx, const double* pnx,
                          const double* pny, int64_t nt, double* tx, double* ty,
                          int64_t* counts);

/* Near-wall clearing pass (clear_inner_panp2, src/Reflect.h:487-727), called after every
 * position update with method 1 and cutoff_mult = 0.5/sqrt(2 pi) (src/Convection.h:301).
 * With dotp the height of a target above the layer cutoff_mult * ips over its nearest element:
 *   method 1: a target with dotp < -0.001 ips moves out along the mean normal by -dotp;
 *   method 0: a target with dotp < its radius (ips for field points, trad == NULL) loses
 *             strength and moves out by the cut-table entry at dotp/radius (particles only).
 * A target that moved is tested again, as the reference's outer loop does, up to 64 passes.
 * tx, ty (and tstr for method 0) are updated in place. Host entry: *circ_removed and
 * *num_cropped (each may be NULL) receive the circulation removed and the number of moves;
 * returns O2D_ERR_STATE, after writing every result, if a target is still moving after 64
 * passes (the reference would loop for ever). Device entry: *circ_removed (device, may be
 * NULL); counts (device int64[3], may be NULL) = moves, unsettled targets, bad panels. */
int o2d_hip_clear_inner_d(const o2d_env* env, int32_t method, double cutoff_mult, double ips,
                          int64_t nnodes, const double* nx, const double* ny, int64_t np,
                          const uint32_t* idx, const double* pnx, const double* pny, int64_t nt,
                          double* tx, double* ty, double* tstr, const double* trad,
                          double* circ_removed, int64_t* num_cropped);
int o2d_hip_clear_inner_dev_d(const o2d_env* env, int32_t method, double cutoff_mult,
                              double ips, int64_t nnodes, const double* nx, const double* ny,
                              int64_t np, const uint32_t* idx, const double* pnx,
                              const double* pny, int64_t nt, double* tx, double* ty,
                              double* tstr, const double* trad, double* circ_removed,
                              int64_t* counts);

/* ---- index work of the split-precision form (exposed for testing) ------------------------ */

/* 32-bit Morton keys of (x, y) on the square box [x0, x0 + 1/inv_extent)^2 (16 bits per axis,
 * x in the even bits) and the stable key order of 0..n-1. Device arrays. */
int o2d_hip_morton_order_dev_d(const o2d_env* env, int64_t n, const double* x, const double* y,
                               double x0, double y0, double inv_extent, uint32_t* keys,
                               uint32_t* order);

/* ---- the reference's external velocity solver hook ------------------------------------- */

/* external_vel_solver_f_ / _tr_f_ / _d_: the exact signatures the reference declares for an
 * external particle velocity solver (src/Influence.h:31-38) and calls from points_affect_points
 * when built with EXTERNAL_VEL_SOLVE and run with a non-internal ExecEnv
 * (src/Influence.h:74-98): counts by pointer, float or double arrays, point targets (_f_, _d_)
 * or particle targets with radii tr (_tr_f_). Linking the reference against this library
 * routes that call to o2d_hip_points_on_points_f/_d. Core, precision and device come from
 * O2D_EXT_CORE (rm|gauss|wl|v2|v3, default v2), O2D_EXT_PRECISION (split|fp64; default split
 * for the float entries, fp64 for the double one) and O2D_EXT_DEVICE (default 0), read on the
 * first call. tu, tv accumulate the unscaled velocity. Returns the flop count
 * nt (2 + flops per pair * ns) of the reference's convention; negative on an error, whose
 * reason is printed to stderr. */
float external_vel_solver_f_(int* ns, const float* sx, const float* sy, const float* ss,
                             const float* sr, int* nt, const float* tx, const float* ty, float* tu,
                             float* tv);
float external_vel_solver_tr_f_(int* ns, const float* sx, const float* sy, const float* ss,
                                const float* sr, int* nt, const float* tx, const float* ty,
                                const float* tr, float* tu, float* tv);
float external_vel_solver_d_(int* ns, const double* sx, const double* sy, const double* ss,
                             const double* sr, int* nt, const double* tx, const double* ty,
                             double* tu, double* tv);

#ifdef __cplusplus
}
#endif

#endif /* OMEGA2D_HIP_H */
