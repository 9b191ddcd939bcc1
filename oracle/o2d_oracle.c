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
 *   oc_vels_to_rhs         src/RHS.h:54-115            pinned: golden RHS vectors.
 *   oc_finalize            src/ElementBase.h:215-234   pinned: compiled reference.
 *   oc_rk_combine          src/ElementBase.h:276-359   pinned: compiled reference (2 and 3
 *                          stages); co-rotating vortex pair convergence order.
 *   oc_points_on_points_adjoint  (no reference counterpart)  pinned: fp64 central finite
 *                          differences of oc_points_on_points; linearity (strength gradient
 *                          equals the transposed sum).
 *   oc_morton_keys         (no reference counterpart; index work for the split-precision
 *                          form)  pinned: brute-force bit interleave in Python.
 *
 * Deliberate difference from the reference, kept identical in the HIP library: at a coincident
 * pair (d^2 == 0) the Gaussian core's gradient factor is taken as 0 (the reference evaluates
 * inf*0 there and returns NaN vorticity; any finite value gives the same result because the
 * factor multiplies dx and dy, which are 0).
 */
#include "o2d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define O2D_PI 3.14159265358979323846

enum { CORE_RM = 1, CORE_GAUSS = 2, CORE_WL = 3, CORE_V2 = 4, CORE_V3 = 5 };
enum { RES_VEL = 1, RES_VORT = 4, RES_SHEAR = 5 };

/* velocity factor K(d^2; sr, tr) (src/CoreFunc.h), tr2 = tr*tr (0 for a point target) */
static double core_k(int core, double d2, double sr, double tr) {
  switch (core) {
    case CORE_RM: return 1.0 / (d2 + sr * sr + tr * tr);
    case CORE_GAUSS: {
      const double ood2 = 1.0 / d2;
      const double corefac = 1.0 / (sr * sr + tr * tr);
      const double reld2 = corefac / ood2;
      if (reld2 > 16.0) return ood2;
      if (reld2 < 0.001) return corefac;
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
      if (reld2 > 16.0) r = ood2;
      else if (reld2 < 0.001) r = corefac;
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
      if (reld2 > 16.0) {
        *k = ood2;
        *bbb = -2.0 * ood2 * ood2;
        *dks = 0.0;
      } else if (reld2 < 0.001) {
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
