// panel_math.h - velocity of a straight constant-strength panel at a point, in fp64, shared by
// the panel kernels (panels on points, points on panels) and the BEM matrix assembly.
//
// The reference's kernels (src/Kernels.h:314-457) build the velocity of a unit vortex panel
// from u* = atan2(dx1,dy1) - atan2(dx0,dy0) wrapped to [-pi, pi] (src/Kernels.h:262-268) and
// v* = log(r0/r1), rotated into the panel frame: a = u* tx - v* ty, b = u* ty + v* tx with
// (tx, ty) the unit tangent. A vortex sheet of strength vs and a source sheet of strength ss
// then give (u, v) = (vs a + ss b, vs b - ss a).
//
// In complex form L(z) = log((z - z0)/(z - z1)) = v* + i u* and b + i a = conj(e) L(z), e the
// unit tangent. About the midpoint m, with h = (z1 - z0)/2 and q = h/(z - m),
// L = 2 atanh(q) = 2 (q + q^3/3 + q^5/5 + ...). A point at least 8|h| from the midpoint
// (|q| <= 1/8) takes that series, summed until a term falls below 1e-17 of the first (at most
// eleven terms; a distant point needs two or three). Nearer points take the closed form with
// one atan2 (of the cross and dot products of the endpoint vectors, which is the wrapped
// difference of the two angles) and one log.
#pragma once

#include <hip/hip_runtime.h>

#include "o2d_math.h"

namespace o2d {

__device__ __forceinline__ void unit_panel_vel(double x0, double y0, double x1, double y1,
                                               double tx, double ty, double& a, double& b) {
  const double px = x1 - x0, py = y1 - y0;
  const double hx = 0.5 * px, hy = 0.5 * py;
  const double zx = tx - (x0 + hx), zy = ty - (y0 + hy);
  const double z2 = __builtin_fma(zx, zx, zy * zy);
  const double h2 = __builtin_fma(hx, hx, hy * hy);
  const double mult = rsq64(4.0 * h2);
  const double tpx = px * mult, tpy = py * mult;
  if (z2 >= 64.0 * h2) {
    const double iz = rcp64(z2);
    const double qx = __builtin_fma(hx, zx, hy * zy) * iz;  // q = h conj(z - m) / |z - m|^2
    const double qy = __builtin_fma(hy, zx, -hx * zy) * iz;
    const double q2x = __builtin_fma(qx, qx, -qy * qy), q2y = 2.0 * qx * qy;
    const double a2 = __builtin_fma(qx, qx, qy * qy);
    double sx = 1.0, sy = 0.0, wx = 1.0, wy = 0.0, mag = 1.0;
    for (int j = 1; j < 12; ++j) {  // sum_j q^(2j) / (2j + 1)
      const double nx = __builtin_fma(wx, q2x, -wy * q2y);
      wy = __builtin_fma(wx, q2y, wy * q2x);
      wx = nx;
      mag *= a2;
      const double c = 1.0 / (double)(2 * j + 1);
      sx = __builtin_fma(c, wx, sx);
      sy = __builtin_fma(c, wy, sy);
      if (mag * c < 1e-17) break;
    }
    const double lx = 2.0 * __builtin_fma(qx, sx, -qy * sy);  // L = 2 q S
    const double ly = 2.0 * __builtin_fma(qx, sy, qy * sx);
    b = __builtin_fma(tpx, lx, tpy * ly);  // conj(e) L = b + i a
    a = __builtin_fma(tpx, ly, -tpy * lx);
    return;
  }
  const double dx0 = tx - x0, dy0 = ty - y0;
  const double dx1 = tx - x1, dy1 = ty - y1;
  const double rij2 = __builtin_fma(dx0, dx0, dy0 * dy0);
  const double rij12 = __builtin_fma(dx1, dx1, dy1 * dy1);
  const double us = atan2(__builtin_fma(py, dx0, -px * dy0), __builtin_fma(dy1, dy0, dx1 * dx0));
  const double vstar = 0.5 * log(rij2 * rcp64(rij12));
  a = us * tpx - vstar * tpy;
  b = us * tpy + vstar * tpx;
}

// A panel with its per-panel quantities worked out once, for kernels that apply one panel to
// many points: endpoints, midpoint m, half-vector h, 64|h|^2 (the series threshold) and a
// complex factor W folded into the result (see panel_log).
struct alignas(16) PanelPre {
  double x0, y0, x1, y1;
  double mx, my, hx, hy;
  double h64, wr, wi, pad;
};

// W = (vs + i ss) conj(e), e the unit tangent: then v + i u = W L(z) is the velocity of the
// panel with vortex strength vs and source strength ss (with vs = 1, ss = 0, W = conj(e))
__device__ __forceinline__ PanelPre panel_pre(double x0, double y0, double x1, double y1,
                                              double vs, double ss) {
  PanelPre p;
  p.x0 = x0;
  p.y0 = y0;
  p.x1 = x1;
  p.y1 = y1;
  const double px = x1 - x0, py = y1 - y0;
  p.hx = 0.5 * px;
  p.hy = 0.5 * py;
  p.mx = x0 + p.hx;
  p.my = y0 + p.hy;
  const double h2 = __builtin_fma(p.hx, p.hx, p.hy * p.hy);
  p.h64 = 64.0 * h2;
  const double mult = rsq64(4.0 * h2);
  const double tpx = px * mult, tpy = py * mult;
  p.wr = __builtin_fma(vs, tpx, ss * tpy);
  p.wi = __builtin_fma(ss, tpx, -vs * tpy);
  p.pad = 0.0;
  return p;
}

// L(z) = log((z - z0)/(z - z1)) = v* + i u* at the point (tx, ty): the atanh series about the
// midpoint beyond 8 half-lengths, else the closed form (the same branches as unit_panel_vel)
__device__ __forceinline__ void panel_log(const PanelPre& p, double tx, double ty, double& lr,
                                          double& li) {
  const double zx = tx - p.mx, zy = ty - p.my;
  const double z2 = __builtin_fma(zx, zx, zy * zy);
  if (z2 >= p.h64) {
    const double iz = rcp64(z2);
    const double qx = __builtin_fma(p.hx, zx, p.hy * zy) * iz;
    const double qy = __builtin_fma(p.hy, zx, -p.hx * zy) * iz;
    const double q2x = __builtin_fma(qx, qx, -qy * qy), q2y = 2.0 * qx * qy;
    const double a2 = __builtin_fma(qx, qx, qy * qy);
    double sx = 1.0, sy = 0.0, wx = 1.0, wy = 0.0, mag = 1.0;
    for (int j = 1; j < 12; ++j) {
      const double nx = __builtin_fma(wx, q2x, -wy * q2y);
      wy = __builtin_fma(wx, q2y, wy * q2x);
      wx = nx;
      mag *= a2;
      const double c = 1.0 / (double)(2 * j + 1);
      sx = __builtin_fma(c, wx, sx);
      sy = __builtin_fma(c, wy, sy);
      if (mag * c < 1e-17) break;
    }
    lr = 2.0 * __builtin_fma(qx, sx, -qy * sy);
    li = 2.0 * __builtin_fma(qx, sy, qy * sx);
    return;
  }
  const double px = 2.0 * p.hx, py = 2.0 * p.hy;
  const double dx0 = tx - p.x0, dy0 = ty - p.y0;
  const double dx1 = tx - p.x1, dy1 = ty - p.y1;
  const double rij2 = __builtin_fma(dx0, dx0, dy0 * dy0);
  const double rij12 = __builtin_fma(dx1, dx1, dy1 * dy1);
  li = atan2(__builtin_fma(py, dx0, -px * dy0), __builtin_fma(dy1, dy0, dx1 * dx0));
  lr = 0.5 * log(rij2 * rcp64(rij12));
}

}  // namespace o2d
