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

// velocity at (tx, ty) of the panel (x0,y0)-(x1,y1) with vortex strength vs and source strength
// ss per unit length (src/Kernels.h:364-405); vortex-only panels pass ss = 0
__device__ __forceinline__ void panel_vel(double x0, double y0, double x1, double y1, double vs,
                                          double ss, double tx, double ty, double& u, double& v) {
  double a, b;
  unit_panel_vel(x0, y0, x1, y1, tx, ty, a, b);
  u = __builtin_fma(vs, a, ss * b);
  v = __builtin_fma(vs, b, -ss * a);
}

}  // namespace o2d
