/*
 * o2d_oracle.c - CPU restatement, in fp64, of Omega2D's direct Biot-Savart path.
 *
 * TEST INFRASTRUCTURE ONLY. Called from tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg, always as the checker or the timed CPU baseline, never as the product.
 * Built by oracle/Makefile with -ffp-contract=off (plain IEEE double, no fused operations).
 *
 * Each routine names the reference lines it restates, and what pins it (tests/test_oracle.py):
 *   oc_points_on_points    src/Influence.h:56-416, src/Kernels.h:81-231, src/CoreFunc.h:42-322
 *                          pinned: compiled-reference golden vectors (V2 core, vel/vort/shear,
 *                          0p and 0b); closed-form single-vortex profiles for RM, Gaussian
 *                          (all three branches), WL, V2, V3; translation/rotation equivariance;
 *                          linearity in strength; long-double brute force.
 *   oc_panels_on_points    src/Influence.h:422-614, src/Kernels.h:241-405
 *                          pinned: golden vectors; Gauss-Legendre quadrature of point vortices
 *                          and sources along the panel; a uniform vortex circle (0 inside,
 *                          Gamma/(2 pi r) outside).
 *   oc_points_on_panels    src/Influence.h:633-771      pinned: golden vectors; reciprocity
 *                          with oc_panels_on_points.
 *   oc_panel_geometry      src/Surfaces.h:682-725      pinned: golden bases and lengths.
 *   oc_body_volume         src/Surfaces.h:623-677      pinned: golden volume; polygon area.
 *   oc_unit_rot_strengths  src/Surfaces.h:570-620      pinned: golden augmented column.
 *   oc_bem_assemble        src/Coefficients.h:155-573  pinned: golden matrices (vortex-only
 *                          and vortex+source, augmented and not); diagonal 1/2; bottom row =
 *                          panel lengths.
 *   oc_bem_assemble_pair   src/Coefficients.h:155-573 (two different surfaces)  pinned:
 *                          golden blocks of one body on another (vortex-only and
 *                          vortex+source, with and without augmentation of either).
 *   oc_vels_to_rhs         src/RHS.h:54-115            pinned: golden RHS vectors.
 *   oc_finalize            src/ElementBase.h:215-234   pinned: compiled reference.
 *   oc_rk_combine          src/ElementBase.h:276-359   pinned: compiled reference (2 and 3
 *                          stages); co-rotating vortex pair convergence order.
 *   oc_points_on_points_adjoint  (no reference counterpart)  pinned: fp64 central finite
 *                          differences of oc_points_on_points; linearity (strength gradient
 *                          equals the transposed sum).
 *   oc_morton_keys         (no reference counterpart; index work for the split-precision
 *                          form)  pinned: brute-force bit interleave in Python.
 *   oc_nearest_distances   src/Reflect.h:792-826       pinned: golden vectors; brute force
 *                          over segments in Python; distance to a circle's polygon.
 *   oc_reflect             src/Reflect.h:167-311       pinned: golden vectors; mirror image
 *                          across a flat wall; reflected points lie outside a convex body.
 *   oc_cut_table           src/Reflect.h:333-394       pinned: golden table; the strength
 *                          fraction is the 2D Gaussian's cumulative weight (-> 1 at +3 radii).
 *   oc_points_vorticity    src/InfluenceVort.h:47-126  pinned: golden vectors; closed-form
 *                          single blob; the vorticity integrates to the circulation.
 *   oc_clear_inner         src/Reflect.h:487-727       pinned: golden vectors (method 1 push
 *                          and method 0 trim); pushed points sit at the cutoff layer of a flat
 *                          wall; the removed circulation equals the strength lost.
 *
 * Deliberate difference from the reference, kept identical in the HIP library: at a coincident
 * pair (d^2 == 0) the Gaussian core's gradient factor is taken as 0 (the reference evaluates
 * inf*0 there and returns NaN vorticity; any finite value gives the same result because the
 * factor multiplies dx and dy, which are 0).
 */
#include "o2d_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define O2D_PI 3.14159265358979323846

enum { CORE_RM = 1, CORE_GAUSS = 2, CORE_WL = 3, CORE_V2 = 4, CORE_V3 = 5 };

/* The Gaussian core's two cutoffs on reld2 = d^2/sigma^2. The build without Vc compiles the
 * generic exp_cond, which compares against the float literals 16.0f and 0.001f
 * (src/CoreFunc.h:119-131); at S=double they are promoted, and (double)0.001f is
 * 0.001000000047497451305389404296875, not 0.001. Between the two the core jumps by about
 * 5e-4 relative, so the literal matters. */
#define GAUSS_FAR_CUT ((double)16.0f)
#define GAUSS_NEAR_CUT ((double)0.001f)
enum { RES_VEL = 1, RES_VORT = 4, RES_SHEAR = 5 };

/* velocity factor K(d^2; sr, tr) (src/CoreFunc.h), tr2 = tr*tr (0 for a point target) */
static double core_k(int core, double d2, double sr, double tr) {
  switch (core) {
    case CORE_RM: return 1.0 / (d2 + sr * sr + tr * tr);
    case CORE_GAUSS: {
      const double ood2 = 1.0 / d2;
      const double corefac = 1.0 / (sr * sr + tr * tr);
      const double reld2 = corefac / ood2;
      if (reld2 > GAUSS_FAR_CUT) return ood2;
      if (reld2 < GAUSS_NEAR_CUT) return corefac;
      return ood2 * (1.0 - exp(-reld2));
    }
    case CORE_WL: {
      const double s2 = sr * sr + tr * tr;
      const double dd = d2 + s2;
      return (dd + s2) / (dd * dd);
    }
    case CORE_V2: {
      const double s2 = sr * sr, t2 = tr * tr;
      return 1.0 / sqrt(d2 * d2 + s2 * s2 + t2 * t2);
    }
    case CORE_V3: {
      const double s2 = sr * sr, t2 = tr * tr;
      return 1.0 / cbrt(d2 * d2 * d2 + s2 * s2 * s2 + t2 * t2 * t2);
    }
  }
  return NAN;
}

/* K and its gradient factor bbb = 2 dK/d(d^2), in the forms of the reference's gradient
 * overloads (src/CoreFunc.h:64-81, 156-180, 208-229, 256-274, 301-321) */
static void core_kg(int core, double d2, double sr, double tr, double* k, double* bbb) {
  switch (core) {
    case CORE_RM: {
      const double r = 1.0 / (d2 + sr * sr + tr * tr);
      *k = r;
      *bbb = -2.0 * r * r;
      return;
    }
    case CORE_GAUSS: {
      const double ood2 = 1.0 / d2;
      const double corefac = 1.0 / (sr * sr + tr * tr);
      const double reld2 = corefac / ood2;
      double r;
      if (reld2 > GAUSS_FAR_CUT) r = ood2;
      else if (reld2 < GAUSS_NEAR_CUT) r = corefac;
      else r = ood2 * (1.0 - exp(-reld2));
      *k = r;
      *bbb = (d2 == 0.0) ? 0.0 : (-2.0 * r * ood2 + 2.0 * corefac * ood2 * exp(-reld2));
      return;
    }
    case CORE_WL: {
      const double s2 = sr * sr + tr * tr;
      const double dd = d2 + s2;
      const double ood2s = 1.0 / (dd * dd);
      const double r = (dd + s2) * ood2s;
      *k = r;
      *bbb = (2.0 - 4.0 * r * dd) * ood2s;
      return;
    }
    case CORE_V2: {
      const double s2 = sr * sr, t2 = tr * tr;
      const double r = 1.0 / sqrt(d2 * d2 + s2 * s2 + t2 * t2);
      *k = r;
      *bbb = -2.0 * r * r * r * d2;
      return;
    }
    case CORE_V3: {
      const double s2 = sr * sr, t2 = tr * tr;
      const double ds6 = d2 * d2 * d2 + s2 * s2 * s2 + t2 * t2 * t2;
      const double r = 1.0 / cbrt(ds6);
      *k = r;
      *bbb = -2.0 * d2 * d2 * r / ds6;
      return;
    }
  }
  *k = NAN;
  *bbb = NAN;
}

/* points_affect_points<double,double> (src/Influence.h:196-249, 346-403) */
int oc_points_on_points(int core, int restype, int64_t ns, const double* sx, const double* sy,
                        const double* ss, const double* sr, int64_t nt, const double* tx,
                        const double* ty, const double* tr, double* tu, double* tv, double* tw,
                        double* te) {
  if (core < CORE_RM || core > CORE_V3) return -3;
  if (restype != RES_VEL && restype != RES_VORT && restype != RES_SHEAR) return -4;
  if (ns < 0 || nt < 0) return -2;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nt; ++i) {
    const double x = tx[i], y = ty[i], r = tr ? tr[i] : 0.0;
    double au = 0.0, av = 0.0, aw = 0.0, ae = 0.0;
    for (int64_t j = 0; j < ns; ++j) {
      const double dx = x - sx[j];
      const double dy = y - sy[j];
      const double d2 = dx * dx + dy * dy;
      if (restype == RES_VEL) {
        const double k = ss[j] * core_k(core, d2, sr[j], r);
        au -= k * dy;
        av += k * dx;
      } else {
        double k, bbb;
        core_kg(core, d2, sr[j], r, &k, &bbb);
        k *= ss[j];
        bbb *= ss[j];
        au -= k * dy;
        av += k * dx;
        if (restype == RES_VORT) {
          const double dudy = -bbb * dy * dy - k;
          const double dvdx = bbb * dx * dx + k;
          aw += dvdx - dudy;
        } else {
          const double dudx = -bbb * dx * dy;
          const double asym = bbb * (dx * dx - dy * dy);
          ae += sqrt(dudx * dudx + 0.25 * asym * asym);
        }
      }
    }
    tu[i] += au;
    tv[i] += av;
    if (restype == RES_VORT) tw[i] += aw;
    if (restype == RES_SHEAR) te[i] += ae;
  }
  return 0;
}

/* get_ustar (src/Kernels.h:262-268): angle subtended, wrapped to [-pi, pi] */
static double ustar_of(double dx0, double dy0, double dx1, double dy1) {
  double us = atan2(dx1, dy1) - atan2(dx0, dy0);
  if (us < -O2D_PI) us += 2.0 * O2D_PI;
  if (us > O2D_PI) us -= 2.0 * O2D_PI;
  return us;
}

/* kernelu_1v_0p (src/Kernels.h:314-354) */
static void panel_1v(double sx0, double sy0, double sx1, double sy1, double str, double tx,
                     double ty, double* u, double* v) {
  const double dx0 = tx - sx0, dy0 = ty - sy0, dx1 = tx - sx1, dy1 = ty - sy1;
  const double px = sx1 - sx0, py = sy1 - sy0;
  const double rij2 = dx0 * dx0 + dy0 * dy0;
  const double rij12 = dx1 * dx1 + dy1 * dy1;
  const double panl = px * px + py * py;
  const double us = ustar_of(dx0, dy0, dx1, dy1);
  const double vs = 0.5 * log(rij2 / rij12);
  const double velx = us * px - vs * py;
  const double vely = us * py + vs * px;
  const double mult = str * (1.0 / sqrt(panl));
  *u = mult * velx;
  *v = mult * vely;
}

/* kernelu_1vos_0p (src/Kernels.h:415-457); kernelu_1vs_0p is its sum */
static void panel_1vos(double sx0, double sy0, double sx1, double sy1, double vstr, double sstr,
                       double tx, double ty, double* vu, double* vv, double* su, double* sv) {
  const double dx0 = tx - sx0, dy0 = ty - sy0, dx1 = tx - sx1, dy1 = ty - sy1;
  double px = sx1 - sx0, py = sy1 - sy0;
  const double rij2 = dx0 * dx0 + dy0 * dy0;
  const double rij12 = dx1 * dx1 + dy1 * dy1;
  const double panl = px * px + py * py;
  const double mult = 1.0 / sqrt(panl);
  px *= mult;
  py *= mult;
  const double us = ustar_of(dx0, dy0, dx1, dy1);
  const double vs = 0.5 * log(rij2 / rij12);
  *vu = vstr * (us * px - vs * py);
  *vv = vstr * (us * py + vs * px);
  *su = sstr * (us * py + vs * px);
  *sv = sstr * (vs * py - us * px);
}

/* panels_affect_points<double,double> (src/Influence.h:548-602) */
int oc_panels_on_points(int64_t nnodes, const double* nx, const double* ny, int64_t np,
                        const uint32_t* idx, const double* vs, const double* ss, int64_t nt,
                        const double* tx, const double* ty, double* tu, double* tv) {
  for (int64_t j = 0; j < 2 * np; ++j)
    if ((int64_t)idx[j] >= nnodes) return -2;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nt; ++i) {
    double au = 0.0, av = 0.0;
    for (int64_t j = 0; j < np; ++j) {
      const uint32_t a = idx[2 * j], b = idx[2 * j + 1];
      double ru, rv;
      if (ss) {
        double vu, vv, su, sv;
        panel_1vos(nx[a], ny[a], nx[b], ny[b], vs[j], ss[j], tx[i], ty[i], &vu, &vv, &su, &sv);
        ru = vu + su;
        rv = vv + sv;
      } else {
        panel_1v(nx[a], ny[a], nx[b], ny[b], vs[j], tx[i], ty[i], &ru, &rv);
      }
      au += ru;
      av += rv;
    }
    tu[i] += au;
    tv[i] += av;
  }
  return 0;
}

/* points_affect_panels<double,double> (src/Influence.h:719-763): panel-averaged velocity of
 * singular point vortices, by running the panel kernel backwards */
int oc_points_on_panels(int64_t ns, const double* sx, const double* sy, const double* ss,
                        int64_t nnodes, const double* nx, const double* ny, int64_t np,
                        const uint32_t* idx, double* pu, double* pv) {
  for (int64_t j = 0; j < 2 * np; ++j)
    if ((int64_t)idx[j] >= nnodes) return -2;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < np; ++i) {
    const uint32_t a = idx[2 * i], b = idx[2 * i + 1];
    const double ex = nx[b] - nx[a], ey = ny[b] - ny[a];
    const double plen = 1.0 / sqrt(ex * ex + ey * ey);
    double au = 0.0, av = 0.0;
    for (int64_t j = 0; j < ns; ++j) {
      double ru, rv;
      panel_1v(nx[a], ny[a], nx[b], ny[b], ss[j], sx[j], sy[j], &ru, &rv);
      au += ru;
      av += rv;
    }
    pu[i] -= plen * au;
    pv[i] -= plen * av;
  }
  return 0;
}

/* Surfaces::compute_bases (src/Surfaces.h:682-725) */
int oc_panel_geometry(int64_t nnodes, const double* nx, const double* ny, int64_t np,
                      const uint32_t* idx, double* tang_x, double* tang_y, double* norm_x,
                      double* norm_y, double* area) {
  for (int64_t i = 0; i < np; ++i) {
    const uint32_t a = idx[2 * i], b = idx[2 * i + 1];
    if ((int64_t)a >= nnodes || (int64_t)b >= nnodes) return -2;
    double x1 = nx[b] - nx[a], y1 = ny[b] - ny[a];
    const double base = sqrt(x1 * x1 + y1 * y1);
    x1 *= (1.0 / base);
    y1 *= (1.0 / base);
    area[i] = base;
    tang_x[i] = x1;
    tang_y[i] = y1;
    if (norm_x) norm_x[i] = -y1;
    if (norm_y) norm_y[i] = x1;
  }
  return 0;
}

/* Surfaces::set_geom_center (src/Surfaces.h:623-677): signed area and centroid of the closed
 * panel loop from the fan of triangles (origin, node a, node b) via Heron's formula */
int oc_body_volume(int64_t nnodes, const double* nx, const double* ny, int64_t np,
                   const uint32_t* idx, double* vol, double* cx, double* cy) {
  double asum = 0.0, xsum = 0.0, ysum = 0.0;
  for (int64_t i = 0; i < np; ++i) {
    const uint32_t j = idx[2 * i], jp1 = idx[2 * i + 1];
    if ((int64_t)j >= nnodes || (int64_t)jp1 >= nnodes) return -2;
    const double xc = (0.0 + nx[j] + nx[jp1]) / 3.0;
    const double yc = (0.0 + ny[j] + ny[jp1]) / 3.0;
    const double panelx = nx[jp1] - nx[j];
    const double panely = ny[jp1] - ny[j];
    const double a = sqrt(pow(nx[j], 2) + pow(ny[j], 2));
    const double b = sqrt(pow(panelx, 2) + pow(panely, 2));
    const double c = sqrt(pow(nx[jp1], 2) + pow(ny[jp1], 2));
    const double hs = 0.5 * (a + b + c);
    const double asqrd = hs * (hs - a) * (hs - b) * (hs - c);
    double thisarea = 0.0;
    if (asqrd > 0.0) thisarea = sqrt(asqrd);
    if (ny[j] * panelx - nx[j] * panely < 0.0) thisarea = -thisarea;
    asum += thisarea;
    xsum += xc * thisarea;
    ysum += yc * thisarea;
  }
  *vol = asum;
  *cx = xsum / asum;
  *cy = ysum / asum;
  return 0;
}

/* Surfaces::add_rot_strengths_base(1.0) on zeroed strengths (src/Surfaces.h:570-620) */
int oc_unit_rot_strengths(int64_t nnodes, const double* nx, const double* ny, int64_t np,
                          const uint32_t* idx, double cx, double cy, double* vs, double* ss) {
  for (int64_t i = 0; i < np; ++i) {
    const uint32_t j = idx[2 * i], jp1 = idx[2 * i + 1];
    if ((int64_t)j >= nnodes || (int64_t)jp1 >= nnodes) return -2;
    const double dx = 0.5 * (nx[j] + nx[jp1]) - cx;
    const double dy = 0.5 * (ny[j] + ny[jp1]) - cy;
    const double ui = -1.0 * dy;
    const double vi = 1.0 * dx;
    double panelx = nx[jp1] - nx[j];
    double panely = ny[jp1] - ny[j];
    const double len = sqrt(panelx * panelx + panely * panely);
    const double oopanell = 1.0 / len;
    panelx *= oopanell;
    panely *= oopanell;
    vs[i] = -1.0 * (ui * panelx + vi * panely);
    ss[i] = -1.0 * (ui * panely - vi * panelx);
  }
  return 0;
}

/* panels_on_panels_coeff<double> for a body acting on itself (src/Coefficients.h:155-573),
 * written straight into the column-major (augmented) matrix A with leading dimension lda.
 * urot_u/urot_v are the finalized unit-rotation panel velocities (src/BEMHelper.h:203-210). */
int oc_bem_assemble(int64_t nnodes, const double* nx, const double* ny, int64_t np,
                    const uint32_t* idx, int nunk, int augmented, const double* urot_u,
                    const double* urot_v, double vol, double* A, int64_t lda) {
  if (nunk != 1 && nunk != 2) return -4;
  const int64_t oldn = np * nunk;
  const int64_t n = oldn + (augmented ? 1 : 0);
  if (lda < n) return -2;
  double* tt0 = (double*)malloc(sizeof(double) * np);
  double* tt1 = (double*)malloc(sizeof(double) * np);
  double* tn0 = (double*)malloc(sizeof(double) * np);
  double* tn1 = (double*)malloc(sizeof(double) * np);
  double* ar = (double*)malloc(sizeof(double) * np);
  int rc = oc_panel_geometry(nnodes, nx, ny, np, idx, tt0, tt1, tn0, tn1, ar);
  if (rc) goto done;
  const double fac = 1.0 / (2.0 * O2D_PI);
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < np; ++j) {
    const uint32_t sa_ = idx[2 * j], sb_ = idx[2 * j + 1];
    const double sx0 = nx[sa_], sy0 = ny[sa_], sx1 = nx[sb_], sy1 = ny[sb_];
    for (int64_t i = 0; i < np; ++i) {
      const uint32_t ta_ = idx[2 * i], tb_ = idx[2 * i + 1];
      const double tx0 = nx[ta_], ty0 = ny[ta_], tx1 = nx[tb_], ty1 = ny[tb_];
      const double xi = 0.5 * (tx1 + tx0);
      const double yi = 0.5 * (ty1 + ty0);
      if (nunk == 2) {
        double vu, vv, su, sv;
        panel_1vos(sx0, sy0, sx1, sy1, 1.0, 1.0, xi, yi, &vu, &vv, &su, &sv);
        double c11 = vu * tt0[i] + vv * tt1[i];
        double c12 = vu * tn0[i] + vv * tn1[i];
        double c21 = su * tt0[i] + sv * tt1[i];
        double c22 = su * tn0[i] + sv * tn1[i];
        panel_1vos(tx0, ty0, tx1, ty1, 1.0, 1.0, 0.5 * (sx0 + sx1), 0.5 * (sy0 + sy1), &vu, &vv,
                   &su, &sv);
        const double fact = ar[j] / ar[i];
        c11 -= fact * (vu * tt0[i] + vv * tt1[i]);
        c12 -= fact * (vu * tn0[i] + vv * tn1[i]);
        c21 -= fact * (su * tt0[i] + sv * tt1[i]);
        c22 -= fact * (su * tn0[i] + sv * tn1[i]);
        c11 *= 0.5;
        c12 *= 0.5;
        c21 *= 0.5;
        c22 *= 0.5;
        if (i == j) {
          c11 = O2D_PI;
          c12 = 0.0;
          c21 = 0.0;
          c22 = O2D_PI;
        }
        A[(2 * j) * lda + 2 * i] = c11 * fac;
        A[(2 * j) * lda + 2 * i + 1] = c12 * fac;
        A[(2 * j + 1) * lda + 2 * i] = c21 * fac;
        A[(2 * j + 1) * lda + 2 * i + 1] = c22 * fac;
      } else {
        double vu, vv;
        panel_1v(sx0, sy0, sx1, sy1, 1.0, xi, yi, &vu, &vv);
        double c = vu * tt0[i] + vv * tt1[i];
        panel_1v(tx0, ty0, tx1, ty1, 1.0, 0.5 * (sx0 + sx1), 0.5 * (sy0 + sy1), &vu, &vv);
        c -= (vu * tt0[i] + vv * tt1[i]) * (ar[j] / ar[i]);
        c *= 0.5;
        if (i == j) c = O2D_PI;
        A[j * lda + i] = c * fac;
      }
    }
  }
  if (augmented) {
    for (int64_t j = 0; j < oldn; ++j) {
      double bottom;
      if (nunk == 1) bottom = ar[j];
      else bottom = (j % 2 == 0) ? ar[j / 2] : 0.0;
      A[j * lda + oldn] = bottom;
    }
    for (int64_t i = 0; i < np; ++i) {
      if (nunk == 1) {
        A[oldn * lda + i] = urot_u[i] * tt0[i] + urot_v[i] * tt1[i];
      } else {
        A[oldn * lda + 2 * i] = urot_u[i] * tt0[i] + urot_v[i] * tt1[i];
        A[oldn * lda + 2 * i + 1] = urot_u[i] * tn0[i] + urot_v[i] * tn1[i];
      }
    }
    A[oldn * lda + oldn] = 2.0 * vol;
  }
done:
  free(tt0);
  free(tt1);
  free(tn0);
  free(tn1);
  free(ar);
  return rc;
}

/* vels_to_rhs_panels (src/RHS.h:54-115) */
int oc_vels_to_rhs(int64_t np, int nunk, const double* pu, const double* pv,
                   const double* tang_x, const double* tang_y, const double* norm_bc,
                   double* rhs) {
  if (nunk == 1) {
    for (int64_t i = 0; i < np; ++i) rhs[i] = -(pu[i] * tang_x[i] + pv[i] * tang_y[i]);
  } else if (nunk == 2) {
    for (int64_t i = 0; i < np; ++i) {
      const double nxv = -tang_y[i], nyv = tang_x[i];
      rhs[2 * i] = -(pu[i] * tang_x[i] + pv[i] * tang_y[i]);
      rhs[2 * i + 1] = -(pu[i] * nxv + pv[i] * nyv);
      rhs[2 * i + 1] += norm_bc ? norm_bc[i] : 0.0;
    }
  } else {
    return -4;
  }
  return 0;
}

/* ElementBase::finalize_vels (src/ElementBase.h:215-221) */
int oc_finalize(int64_t n, double* u, double* v, double fsx, double fsy) {
  const double factor = 0.5 / O2D_PI;
  for (int64_t i = 0; i < n; ++i) u[i] = fsx + u[i] * factor;
  for (int64_t i = 0; i < n; ++i) v[i] = fsy + v[i] * factor;
  return 0;
}

/* ElementBase::move with 1, 2 or 3 weighted stage velocities (src/ElementBase.h:276-359) */
int oc_rk_combine(int64_t n, int nstage, const double* w, const double* const* us,
                  const double* const* vs, double dt, double* x, double* y) {
  if (nstage < 1 || nstage > 3) return -4;
  for (int64_t i = 0; i < n; ++i) {
    if (nstage == 1) {
      x[i] += dt * w[0] * us[0][i];
      y[i] += dt * w[0] * vs[0][i];
    } else {
      double u = w[0] * us[0][i] + w[1] * us[1][i];
      double v = w[0] * vs[0][i] + w[1] * vs[1][i];
      if (nstage == 3) {
        u += w[2] * us[2][i];
        v += w[2] * vs[2][i];
      }
      x[i] += dt * u;
      y[i] += dt * v;
    }
  }
  return 0;
}

/* derivative of K with respect to the source core radius, and the gradient factor consistent
 * with the branches of K actually taken (for the exact adjoint of the forward sum) */
static void core_adj(int core, double d2, double sr, double* k, double* bbb, double* dks) {
  switch (core) {
    case CORE_RM: {
      const double r = 1.0 / (d2 + sr * sr);
      *k = r;
      *bbb = -2.0 * r * r;
      *dks = -2.0 * sr * r * r;
      return;
    }
    case CORE_GAUSS: {
      const double ood2 = 1.0 / d2;
      const double corefac = 1.0 / (sr * sr);
      const double reld2 = corefac / ood2;
      if (reld2 > GAUSS_FAR_CUT) {
        *k = ood2;
        *bbb = -2.0 * ood2 * ood2;
        *dks = 0.0;
      } else if (reld2 < GAUSS_NEAR_CUT) {
        *k = corefac;
        *bbb = 0.0;
        *dks = -2.0 * corefac / sr;
      } else {
        const double e = exp(-reld2);
        const double r = ood2 * (1.0 - e);
        *k = r;
        *bbb = -2.0 * r * ood2 + 2.0 * corefac * ood2 * e;
        *dks = -2.0 * e * corefac / sr;
      }
      return;
    }
    case CORE_WL: {
      const double s2 = sr * sr;
      const double dd = d2 + s2;
      const double r = (dd + s2) / (dd * dd);
      *k = r;
      *bbb = -2.0 * (dd + 2.0 * s2) / (dd * dd * dd);
      *dks = -4.0 * s2 * sr / (dd * dd * dd);
      return;
    }
    case CORE_V2: {
      const double s2 = sr * sr;
      const double r = 1.0 / sqrt(d2 * d2 + s2 * s2);
      *k = r;
      *bbb = -2.0 * r * r * r * d2;
      *dks = -2.0 * s2 * sr * r * r * r;
      return;
    }
    case CORE_V3: {
      const double s2 = sr * sr;
      const double ds6 = d2 * d2 * d2 + s2 * s2 * s2;
      const double r = 1.0 / cbrt(ds6);
      *k = r;
      *bbb = -2.0 * d2 * d2 * r / ds6;
      *dks = -2.0 * s2 * s2 * sr * r / ds6;
      return;
    }
  }
  *k = *bbb = *dks = NAN;
}

/* vector-Jacobian product of the unscaled point-target velocity sum
 *   tu_k = -sum_j s_j K(d_kj^2; r_j) dy_kj,  tv_k = sum_j s_j K dx_kj,  dx_kj = tx_k - sx_j
 * with respect to every source's x, y, strength and core radius, given gu = dJ/dtu and
 * gv = dJ/dtv at the point targets. Outputs are accumulated (+=). */
int oc_points_on_points_adjoint(int core, int64_t ns, const double* sx, const double* sy,
                                const double* ss, const double* sr, int64_t nt, const double* tx,
                                const double* ty, const double* gu, const double* gv,
                                double* gx, double* gy, double* gs, double* gr) {
  if (core < CORE_RM || core > CORE_V3) return -3;
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < ns; ++j) {
    double ax = 0.0, ay = 0.0, as = 0.0, arr = 0.0;
    const double s = ss[j];
    for (int64_t k = 0; k < nt; ++k) {
      const double dx = tx[k] - sx[j];
      const double dy = ty[k] - sy[j];
      const double d2 = dx * dx + dy * dy;
      double kk, bbb, dks;
      core_adj(core, d2, sr[j], &kk, &bbb, &dks);
      const double w = gv[k] * dx - gu[k] * dy;
      as += kk * w;
      arr += s * dks * w;
      ax += s * (gu[k] * bbb * dx * dy - gv[k] * (kk + bbb * dx * dx));
      ay += s * (gu[k] * (kk + bbb * dy * dy) - gv[k] * bbb * dx * dy);
    }
    gx[j] += ax;
    gy[j] += ay;
    gs[j] += as;
    gr[j] += arr;
  }
  return 0;
}

static uint32_t spread16(uint32_t v) {
  v &= 0xFFFFu;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

/* 32-bit Morton key: 16 bits per axis, x in the even bits. Quantisation
 *   q = clamp(floor(((x - x0) * inv_extent) * 65536), 0, 65535)
 * uses two IEEE double operations and no fused multiply-add, so it is reproducible. */
int oc_morton_keys(int64_t n, const double* x, const double* y, double x0, double y0,
                   double inv_extent, uint32_t* keys) {
  for (int64_t i = 0; i < n; ++i) {
    double qx = floor(((x[i] - x0) * inv_extent) * 65536.0);
    double qy = floor(((y[i] - y0) * inv_extent) * 65536.0);
    qx = qx < 0.0 ? 0.0 : (qx > 65535.0 ? 65535.0 : qx);
    qy = qy < 0.0 ? 0.0 : (qy > 65535.0 ? 65535.0 : qy);
    keys[i] = spread16((uint32_t)qx) | (spread16((uint32_t)qy) << 1);
  }
  return 0;
}

/* ------------------------------------------------------------------------------------------
 * Particle-panel proximity (src/Reflect.h). Panels are straight two-node segments; a target's
 * nearest element is a panel interior or a node, found by a scan over panels in index order
 * with the reference's one-ulp tie window.
 * ------------------------------------------------------------------------------------------ */

typedef struct {
  int is_node;      /* 0: panel interior, 1: node */
  int64_t jidx;     /* panel index, or node index for a node hit */
  double distsq, cpx, cpy;
} oc_hit;

/* panel_point_distance (src/Reflect.h:110-165). The reference's std::pow(v, 2) is v*v. */
static oc_hit panel_point(double sx0, double sy0, double sx1, double sy1, double tx, double ty) {
  oc_hit h;
  const double bx = sx1 - sx0, by = sy1 - sy0;
  const double blensq = bx * bx + by * by;
  const double ax = tx - sx0, ay = ty - sy0;
  const double t = (ax * bx + ay * by) / blensq;
  if (t > 0.0 && t < 1.0) {
    const double c = ay * bx - ax * by;
    h.is_node = 0;
    h.jidx = -1;
    h.distsq = (c * c) / blensq;
    h.cpx = (1.0 - t) * sx0 + t * sx1;
    h.cpy = (1.0 - t) * sy0 + t * sy1;
    return h;
  }
  h.is_node = 1;
  const double rij2 = ax * ax + ay * ay;
  const double ex = tx - sx1, ey = ty - sy1;
  const double rij12 = ex * ex + ey * ey;
  if (rij2 < rij12) {
    h.jidx = 0;
    h.distsq = rij2;
    h.cpx = sx0;
    h.cpy = sy0;
  } else {
    h.jidx = 1;
    h.distsq = rij12;
    h.cpx = sx1;
    h.cpy = sy1;
  }
  return h;
}

/* node normals: the sum of the normals of the panels that use the node, added in panel order,
 * end 0 before end 1 (src/Reflect.h:186-197, 510-521) */
static void node_normals(int64_t nnodes, int64_t np, const uint32_t* idx, const double* pnx,
                         const double* pny, double* nnx, double* nny) {
  for (int64_t k = 0; k < nnodes; ++k) nnx[k] = nny[k] = 0.0;
  for (int64_t j = 0; j < np; ++j) {
    const uint32_t a = idx[2 * j], b = idx[2 * j + 1];
    nnx[a] += pnx[j];
    nny[a] += pny[j];
    nnx[b] += pnx[j];
    nny[b] += pny[j];
  }
}

/* The scan over panels shared by reflect_panp2 and clear_inner_panp2 (src/Reflect.h:207-240,
 * 553-584): a result strictly below nextafter(min, 0) replaces the hit list; one below
 * nextafter(min, +max) joins it. Returns the number of hits; the first hit goes to *first and
 * the sums of the hits' normals and contact points (in hit order) to sums[0..3]. */
static int64_t scan_hits(double mindist, int64_t np, const uint32_t* idx, const double* nx,
                         const double* ny, const double* pnx, const double* pny,
                         const double* nnx, const double* nny, double tx, double ty,
                         oc_hit* first, double* sums) {
  int64_t nh = 0;
  for (int64_t j = 0; j < np; ++j) {
    const uint32_t a = idx[2 * j], b = idx[2 * j + 1];
    oc_hit h = panel_point(nx[a], ny[a], nx[b], ny[b], tx, ty);
    int take = 0;
    if (h.distsq < nextafter(mindist, 0.0)) {
      mindist = h.distsq;
      nh = 0;
      sums[0] = sums[1] = sums[2] = sums[3] = 0.0;
      take = 1;
    } else if (h.distsq < nextafter(mindist, DBL_MAX)) {
      take = 1;
    }
    if (!take) continue;
    h.jidx = h.is_node ? (int64_t)idx[2 * j + h.jidx] : j;
    if (nh == 0) *first = h;
    sums[0] += h.is_node ? nnx[h.jidx] : pnx[h.jidx];
    sums[1] += h.is_node ? nny[h.jidx] : pny[h.jidx];
    sums[2] += h.cpx;
    sums[3] += h.cpy;
    ++nh;
  }
  return nh;
}

int oc_nearest_distances(int64_t nnodes, const double* nx, const double* ny, int64_t np,
                         const uint32_t* idx, int64_t nt, const double* tx, const double* ty,
                         double* dist) {
  for (int64_t j = 0; j < np; ++j)
    if ((int64_t)idx[2 * j] >= nnodes || (int64_t)idx[2 * j + 1] >= nnodes) return -2;
  for (int64_t i = 0; i < nt; ++i) {
    double mindist = DBL_MAX;
    for (int64_t j = 0; j < np; ++j) {
      const uint32_t a = idx[2 * j], b = idx[2 * j + 1];
      const oc_hit h = panel_point(nx[a], ny[a], nx[b], ny[b], tx[i], ty[i]);
      if (h.distsq < nextafter(mindist, 0.0)) mindist = h.distsq;
    }
    dist[i] = sqrt(mindist);
  }
  return 0;
}

int oc_reflect(int64_t nnodes, const double* nx, const double* ny, int64_t np,
               const uint32_t* idx, const double* pnx, const double* pny, int64_t nt, double* tx,
               double* ty, int64_t* num_reflected) {
  for (int64_t j = 0; j < np; ++j)
    if ((int64_t)idx[2 * j] >= nnodes || (int64_t)idx[2 * j + 1] >= nnodes) return -2;
  double* nnx = (double*)calloc((size_t)(nnodes > 0 ? nnodes : 1), sizeof(double));
  double* nny = (double*)calloc((size_t)(nnodes > 0 ? nnodes : 1), sizeof(double));
  if (!nnx || !nny) {
    free(nnx);
    free(nny);
    return -1;
  }
  node_normals(nnodes, np, idx, pnx, pny, nnx, nny);
  int64_t nref = 0;
  for (int64_t i = 0; i < nt; ++i) {
    oc_hit h0;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    const int64_t nh = scan_hits(DBL_MAX, np, idx, nx, ny, pnx, pny, nnx, nny, tx[i], ty[i],
                                 &h0, s);
    if (nh == 0) continue; /* the reference asserts here */
    double normx = s[0], normy = s[1], cpx = s[2], cpy = s[3];
    if (normx * normx + normy * normy == 0.0) {
      normx += h0.is_node ? nnx[h0.jidx] : pnx[h0.jidx];
      normy += h0.is_node ? nny[h0.jidx] : pny[h0.jidx];
      cpx = h0.cpx;
      cpy = h0.cpy;
    }
    const double normilen = 1.0 / sqrt(normx * normx + normy * normy);
    normx *= normilen;
    normy *= normilen;
    cpx /= (double)nh;
    cpy /= (double)nh;
    const double dotp = normx * (tx[i] - cpx) + normy * (ty[i] - cpy);
    if (dotp < 0.0) {
      const double dist = sqrt(h0.distsq);
      tx[i] = cpx + dist * normx;
      ty[i] = cpy + dist * normy;
      ++nref;
    }
  }
  free(nnx);
  free(nny);
  if (num_reflected) *num_reflected = nref;
  return 0;
}

/* init_cut_tables (src/Reflect.h:333-394): (position, kept strength fraction, shift) rows of the
 * cumulative 2D Gaussian, positions in units of the particle radius. Returns the row count. */
int oc_cut_table(double dxin, int cap, double* pos, double* smult, double* dshift) {
  const double max_rad = 3.0;
  const int nx = (int)(0.5 + max_rad / dxin);
  const double dx = max_rad / (double)(float)nx;
  const int rows = 2 * nx + 3;
  if (rows > cap) return -rows;
  int r = 0;
  pos[r] = (double)(-nx - 0.5) * dx;
  smult[r] = 0.0;
  dshift[r] = 0.0;
  ++r;
  double twgt = 0.0, tmom = 0.0;
  for (int i = -nx; i < nx + 1; ++i) {
    const double xi = (double)i * dx;
    const double distxs = xi * xi;
    double rwgt = 0.0;
    for (int j = -nx; j < nx + 1; ++j) {
      const double yj = (double)j * dx;
      const double distxys = distxs + yj * yj;
      rwgt += exp(-distxys);
    }
    rwgt *= dx * dx;
    rwgt *= 1.0 / O2D_PI;
    twgt += rwgt;
    tmom += ((double)i * dx) * rwgt;
    pos[r] = (double)(i + 0.5) * dx;
    smult[r] = twgt;
    dshift[r] = -tmom / twgt;
    ++r;
  }
  pos[r] = (double)(nx + 1.5) * dx;
  smult[r] = 1.0;
  dshift[r] = 0.0;
  ++r;
  return r;
}

/* get_cut_entry (src/Reflect.h:400-432) */
static void cut_entry(int rows, const double* pos, const double* sm, const double* ds, double p,
                      double* smult, double* dshift) {
  *smult = 1.0;
  *dshift = 0.0;
  if (p < pos[0]) {
    *smult = 0.0;
    *dshift = 1.0 - p;
  } else if (p > pos[rows - 1]) {
    *smult = 1.0;
    *dshift = 0.0;
  } else {
    for (int i = 1; i < rows; ++i) {
      if (p < pos[i]) {
        const double frac = (p - pos[i - 1]) / (pos[i] - pos[i - 1]);
        *smult = frac * sm[i] + (1.0 - frac) * sm[i - 1];
        *dshift = frac * ds[i] + (1.0 - frac) * ds[i - 1];
        break;
      }
    }
  }
}

/* clear_inner_panp2 (src/Reflect.h:487-727). Each target is re-tested until it passes, as the
 * reference's outer while loop does; a target still moving after max_passes tests is counted
 * in *unsettled and left where it is. The removed circulation is summed per target in pass
 * order, then over targets in index order. */
int oc_clear_inner(int method, int64_t nnodes, const double* nx, const double* ny, int64_t np,
                   const uint32_t* idx, const double* pnx, const double* pny, double cutoff_mult,
                   double ips, int64_t nt, double* tx, double* ty, double* ts, const double* tr,
                   int max_passes, double* circ_removed, int64_t* num_cropped,
                   int64_t* unsettled) {
  if (method != 0 && method != 1) return -1;
  if (method == 0 && tr && !ts) return -1;
  for (int64_t j = 0; j < np; ++j)
    if ((int64_t)idx[2 * j] >= nnodes || (int64_t)idx[2 * j + 1] >= nnodes) return -2;
  double cpos[128], csm[128], cds[128];
  const int rows = oc_cut_table(0.1, 128, cpos, csm, cds);
  if (rows <= 0) return -1;
  double* nnx = (double*)calloc((size_t)(nnodes > 0 ? nnodes : 1), sizeof(double));
  double* nny = (double*)calloc((size_t)(nnodes > 0 ? nnodes : 1), sizeof(double));
  if (!nnx || !nny) {
    free(nnx);
    free(nny);
    return -1;
  }
  node_normals(nnodes, np, idx, pnx, pny, nnx, nny);
  const double start = 0.1 * DBL_MAX;
  double circ = 0.0;
  int64_t ncrop = 0, nuns = 0;
  for (int64_t i = 0; i < nt; ++i) {
    double mine = 0.0;
    int settled = 0;
    for (int pass = 0; pass < max_passes && !settled; ++pass) {
      oc_hit h0;
      double s[4] = {0.0, 0.0, 0.0, 0.0};
      const int64_t nh = scan_hits(start, np, idx, nx, ny, pnx, pny, nnx, nny, tx[i], ty[i],
                                   &h0, s);
      if (nh == 0) continue; /* the reference skips the target and tests it again */
      double normx = s[0], normy = s[1], cpx = s[2], cpy = s[3];
      const double normilen = 1.0 / sqrt(normx * normx + normy * normy);
      normx *= normilen;
      normy *= normilen;
      cpx /= (double)nh;
      cpy /= (double)nh;
      const double dotp =
          normx * (tx[i] - cpx) + normy * (ty[i] - cpy) - cutoff_mult * ips;
      if (method == 0) {
        /* field points (no radii) take the spacing as radius and are never modified, so
         * one below the layer stays unsettled (the reference loops on it for ever) */
        const double rad = tr ? tr[i] : ips;
        if (dotp < rad) {
          if (tr) {
            double sm, dsh;
            cut_entry(rows, cpos, csm, cds, dotp / rad, &sm, &dsh);
            mine += ts[i] * (1.0 - sm);
            ts[i] *= sm;
            tx[i] += dsh * rad * normx;
            ty[i] += dsh * rad * normy;
          }
          ++ncrop;
        } else {
          settled = 1;
        }
      } else {
        if (dotp < -0.001 * ips) {
          tx[i] -= dotp * normx;
          ty[i] -= dotp * normy;
          ++ncrop;
        } else {
          settled = 1;
        }
      }
    }
    if (!settled) ++nuns;
    circ += mine;
  }
  free(nnx);
  free(nny);
  if (circ_removed) *circ_removed = circ;
  if (num_cropped) *num_cropped = ncrop;
  if (unsettled) *unsettled = nuns;
  return 0;
}

/* points_affect_points_vorticity (src/InfluenceVort.h:47-126, the x86 loop :103-117): the
 * unscaled Gaussian vorticity sum tw_i += sum_j 2 s_j exp(-d^2/r_j^2) / r_j^2 at every target;
 * finalize_vels multiplies it by 1/(2 pi) (src/ElementBase.h:221-226). */
int oc_points_vorticity(int64_t ns, const double* sx, const double* sy, const double* ss,
                        const double* sr, int64_t nt, const double* tx, const double* ty,
                        double* tw) {
  for (int64_t i = 0; i < nt; ++i) {
    double acc = 0.0;
    for (int64_t j = 0; j < ns; ++j) {
      const double dx = tx[i] - sx[j], dy = ty[i] - sy[j];
      const double r2 = sr[j] * sr[j];
      const double distsq = (dx * dx + dy * dy) / r2;
      acc += 2.0 * ss[j] * exp(-distsq) / r2;
    }
    tw[i] += acc;
  }
  return 0;
}

/* panels_on_panels_coeff for two different surfaces (src/Coefficients.h:155-573 with
 * &src != &targ): the block of target equations (rows) by source unknowns (columns), two-way
 * averaged, scaled by 1/(2 pi), with no self term. targ_aug adds the bottom row, which is 0
 * here because the lengths row counts only a body's own panels (:474-491); src_aug adds the
 * right column, the source body's unit-rotation velocity urot_u/urot_v on the target panels
 * projected on their tangents (and normals) (:509-531), and the corner, 0 here (:533-537). */
int oc_bem_assemble_pair(int64_t snn, const double* snx, const double* sny, int64_t snp,
                         const uint32_t* sidx, int64_t tnn, const double* tnx, const double* tny,
                         int64_t tnp, const uint32_t* tidx, int nunk, int src_aug, int targ_aug,
                         const double* urot_u, const double* urot_v, double* A, int64_t lda) {
  if (nunk != 1 && nunk != 2) return -4;
  const int64_t oldr = tnp * nunk, oldc = snp * nunk;
  const int64_t nr = oldr + (targ_aug ? 1 : 0);
  if (lda < nr) return -2;
  double* tt0 = (double*)malloc(sizeof(double) * (tnp > 0 ? tnp : 1));
  double* tt1 = (double*)malloc(sizeof(double) * (tnp > 0 ? tnp : 1));
  double* tn0 = (double*)malloc(sizeof(double) * (tnp > 0 ? tnp : 1));
  double* tn1 = (double*)malloc(sizeof(double) * (tnp > 0 ? tnp : 1));
  double* ta = (double*)malloc(sizeof(double) * (tnp > 0 ? tnp : 1));
  double* st0 = (double*)malloc(sizeof(double) * (snp > 0 ? snp : 1));
  double* st1 = (double*)malloc(sizeof(double) * (snp > 0 ? snp : 1));
  double* sa = (double*)malloc(sizeof(double) * (snp > 0 ? snp : 1));
  int rc = oc_panel_geometry(tnn, tnx, tny, tnp, tidx, tt0, tt1, tn0, tn1, ta);
  if (!rc) rc = oc_panel_geometry(snn, snx, sny, snp, sidx, st0, st1, NULL, NULL, sa);
  if (rc) goto done;
  const double fac = 1.0 / (2.0 * O2D_PI);
  for (int64_t j = 0; j < snp; ++j) {
    const uint32_t sa_ = sidx[2 * j], sb_ = sidx[2 * j + 1];
    const double sx0 = snx[sa_], sy0 = sny[sa_], sx1 = snx[sb_], sy1 = sny[sb_];
    for (int64_t i = 0; i < tnp; ++i) {
      const uint32_t ta_ = tidx[2 * i], tb_ = tidx[2 * i + 1];
      const double tx0 = tnx[ta_], ty0 = tny[ta_], tx1 = tnx[tb_], ty1 = tny[tb_];
      const double xi = 0.5 * (tx1 + tx0), yi = 0.5 * (ty1 + ty0);
      const double fact = sa[j] / ta[i];
      if (nunk == 2) {
        double vu, vv, su, sv;
        panel_1vos(sx0, sy0, sx1, sy1, 1.0, 1.0, xi, yi, &vu, &vv, &su, &sv);
        double c11 = vu * tt0[i] + vv * tt1[i];
        double c12 = vu * tn0[i] + vv * tn1[i];
        double c21 = su * tt0[i] + sv * tt1[i];
        double c22 = su * tn0[i] + sv * tn1[i];
        panel_1vos(tx0, ty0, tx1, ty1, 1.0, 1.0, 0.5 * (sx0 + sx1), 0.5 * (sy0 + sy1), &vu, &vv,
                   &su, &sv);
        c11 = 0.5 * (c11 - fact * (vu * tt0[i] + vv * tt1[i]));
        c12 = 0.5 * (c12 - fact * (vu * tn0[i] + vv * tn1[i]));
        c21 = 0.5 * (c21 - fact * (su * tt0[i] + sv * tt1[i]));
        c22 = 0.5 * (c22 - fact * (su * tn0[i] + sv * tn1[i]));
        A[(2 * j) * lda + 2 * i] = c11 * fac;
        A[(2 * j) * lda + 2 * i + 1] = c12 * fac;
        A[(2 * j + 1) * lda + 2 * i] = c21 * fac;
        A[(2 * j + 1) * lda + 2 * i + 1] = c22 * fac;
      } else {
        double vu, vv;
        panel_1v(sx0, sy0, sx1, sy1, 1.0, xi, yi, &vu, &vv);
        double c = vu * tt0[i] + vv * tt1[i];
        panel_1v(tx0, ty0, tx1, ty1, 1.0, 0.5 * (sx0 + sx1), 0.5 * (sy0 + sy1), &vu, &vv);
        c = 0.5 * (c - (vu * tt0[i] + vv * tt1[i]) * fact);
        A[j * lda + i] = c * fac;
      }
    }
  }
  if (targ_aug)
    for (int64_t j = 0; j < oldc; ++j) A[j * lda + oldr] = 0.0;
  if (src_aug) {
    for (int64_t i = 0; i < tnp; ++i) {
      if (nunk == 1) {
        A[oldc * lda + i] = urot_u[i] * tt0[i] + urot_v[i] * tt1[i];
      } else {
        A[oldc * lda + 2 * i] = urot_u[i] * tt0[i] + urot_v[i] * tt1[i];
        A[oldc * lda + 2 * i + 1] = urot_u[i] * tn0[i] + urot_v[i] * tn1[i];
      }
    }
    if (targ_aug) A[oldc * lda + oldr] = 0.0;
  }
done:
  free(tt0);
  free(tt1);
  free(tn0);
  free(tn1);
  free(ta);
  free(st0);
  free(st1);
  free(sa);
  return rc;
}
