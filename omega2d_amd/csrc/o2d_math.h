// o2d_math.h - device-side regularised core functions (src/CoreFunc.h) in fp64 and fp32.
//
// fp64 forms follow the reference's arithmetic closely enough for 1e-12 relative agreement:
// reciprocals and reciprocal square roots are the hardware seeds refined by one third-order step
// (about 1 ulp, as two Newton steps), exp/cbrt come from the device math library. fp32 forms are the split-precision
// pair math: one hardware reciprocal (or rsq) per pair, the Gaussian's exp only inside its
// core (d^2 <= 16 sigma^2), evaluated as an expm1-style series where 1 - e^-x would cancel.
#pragma once

#include <hip/hip_runtime.h>

#include "omega2d_hip.h"

namespace o2d {

// 1/x from the hardware seed r (relative error below 5e-8 on gfx950) and one third-order
// step: with e = 1 - x r, 1/x = r / (1 - e) = r (1 + e + e^2) + O(e^3). The e^3 term is below
// 1.3e-22, so the result is within rounding of two Newton steps (measured on 2^20 log-uniform
// x in [e^-20, e^20]: largest relative error 1.1e-16 for both) with three fused ops, not four.
// rcp64(0) is NaN, not inf (the seed is inf and e = fma(-0, inf, 1) is NaN): callers either
// never pass 0 or discard the result, as the Gaussian's coincident pair does.
__device__ __forceinline__ double rcp64(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  const double e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(__builtin_fma(e, e, e), r, r);
}

// 1/sqrt(x) from the hardware seed y (relative error below 6e-8) and one third-order step:
// with d = 1 - x y^2, 1/sqrt(x) = y (1 - d)^(-1/2) = y (1 + d/2 + 3 d^2/8) + O(d^3)
// (largest relative error 1.4e-16 on the same sample, as for two Newton steps).
__device__ __forceinline__ double rsq64(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double d = __builtin_fma(-x * y, y, 1.0);
  return __builtin_fma(y * d, __builtin_fma(d, 0.375, 0.5), y);
}

// ---- radius term ---------------------------------------------------------------------------
// Every core depends on the two core radii only through w = f(sr) + f(tr): f(r) = r^2 for RM,
// Gaussian and WL, r^4 for V2, r^6 for V3 (src/CoreFunc.h:48,56,136-147,190-200,239-249,283-294).
// Sources carry f(sr), targets f(tr) (0 for point targets). Products are rounded one by one
// (no fused multiply-add), as the reference writes them, so that the fp64 Gaussian's cutoff
// predicate below sees the same w.
template <int CORE, class T>
__host__ __device__ __forceinline__ T rterm(T r) {
#pragma clang fp contract(off)
  const T r2 = r * r;
  if constexpr (CORE == O2D_CORE_V2) return r2 * r2;
  else if constexpr (CORE == O2D_CORE_V3) return r2 * r2 * r2;
  else return r2;
}

// the sum of two radius terms, rounded once (never fused with the products that formed them)
__host__ __device__ __forceinline__ double wsum(double a, double b) {
#pragma clang fp contract(off)
  return a + b;
}

// ---- the Gaussian's compact cutoffs --------------------------------------------------------
// The reference without Vc compiles the generic exp_cond (src/CoreFunc.h:117-131): with
// reld2 = (1/w) / (1/d^2) it returns 1/d^2 for reld2 > 16.0f, 1/w for reld2 < 0.001f and
// (1 - e^-reld2)/d^2 between. The literals are floats; at S = double the second is promoted to
// 0.001000000047497451..., not 0.001, and the core jumps by about 5e-4 relative across it.
constexpr double kGaussFar = 16.0;
constexpr double kGaussNear = (double)0.001f;
// The kernels test x = d^2 * rcp64(w), within a few ulps of reld2. Where x lies within this
// relative margin of a cutoff, gauss_branch_ref evaluates the reference's own expression.
constexpr double kCutMargin = 0x1p-40;

// The branch as the reference computes it: IEEE-rounded divisions (1/w, 1/d^2 and their
// quotient) and d^2 = dx*dx + dy*dy with each product rounded. 1: beyond the far cutoff,
// -1: below the near one, 0: between. Out of line: it runs only for pairs at a cutoff.
__device__ __noinline__ int gauss_branch_ref(double dx, double dy, double w) {
#pragma clang fp contract(off)
  const double d2 = dx * dx + dy * dy;
  const double corefac = 1.0 / w;
  const double ood2 = 1.0 / d2;
  const double reld2 = corefac / ood2;
  return reld2 > kGaussFar ? 1 : (reld2 < kGaussNear ? -1 : 0);
}

// the branch from x = d^2 / w (a few ulps from reld2), exact near the cutoffs
__device__ __forceinline__ int gauss_branch(double x, double dx, double dy, double w) {
  if (x < kGaussNear * (1.0 - kCutMargin)) return -1;
  if (x > kGaussNear * (1.0 + kCutMargin) && x < kGaussFar * (1.0 - kCutMargin)) return 0;
  if (x > kGaussFar * (1.0 + kCutMargin)) return 1;
  return gauss_branch_ref(dx, dy, w);
}

// ---- fp64 velocity factor K(d2; w) ---------------------------------------------------------

// Gaussian inside (or at) its compact cutoff (src/CoreFunc.h:117-131), out of line like
// gauss_core32
__device__ __noinline__ double gauss_core64(double dx, double dy, double d2, double w) {
  const double cf = rcp64(w);
  const double x = d2 * cf;
  const int br = gauss_branch(x, dx, dy, w);
  if (br < 0) return cf;
  const double ood2 = rcp64(d2);
  return br > 0 ? ood2 : (1.0 - exp(-x)) * ood2;
}

// dx, dy: the pair's offsets; d2 their squared length (possibly fused); w the radius term
template <int CORE>
__device__ __forceinline__ double core_k64(double dx, double dy, double d2, double w) {
  if constexpr (CORE == O2D_CORE_RM) {
    return rcp64(d2 + w);
  } else if constexpr (CORE == O2D_CORE_GAUSSIAN) {
    if (d2 > kGaussFar * (1.0 + kCutMargin) * w) return rcp64(d2);
    return gauss_core64(dx, dy, d2, w);
  } else if constexpr (CORE == O2D_CORE_WL) {
    const double dd = d2 + w;
    const double r = rcp64(dd);
    return (dd + w) * r * r;
  } else if constexpr (CORE == O2D_CORE_V2) {
    return rsq64(__builtin_fma(d2, d2, w));
  } else {
    return rcbrt(__builtin_fma(d2 * d2, d2, w));
  }
}

// K and bbb = 2 dK/d(d^2) in the reference's gradient forms (src/CoreFunc.h:64-81, 156-180,
// 208-229, 256-274, 301-321). Gaussian at d2 == 0: bbb = 0 (the reference yields NaN there).
template <int CORE>
__device__ __forceinline__ void core_kg64(double dx, double dy, double d2, double w, double& k,
                                          double& bbb) {
  if constexpr (CORE == O2D_CORE_RM) {
    const double r = rcp64(d2 + w);
    k = r;
    bbb = -2.0 * r * r;
  } else if constexpr (CORE == O2D_CORE_GAUSSIAN) {
    const double cf = rcp64(w);
    const double x = d2 * cf;
    const double ood2 = rcp64(d2);
    const double e = exp(-x);
    const int br = gauss_branch(x, dx, dy, w);
    const double r = br > 0 ? ood2 : (br < 0 ? cf : ood2 * (1.0 - e));
    k = r;
    bbb = (d2 == 0.0) ? 0.0 : (-2.0 * r * ood2 + 2.0 * cf * ood2 * e);
  } else if constexpr (CORE == O2D_CORE_WL) {
    const double dd = d2 + w;
    const double ood2s = rcp64(dd * dd);
    const double r = (dd + w) * ood2s;
    k = r;
    bbb = (2.0 - 4.0 * r * dd) * ood2s;
  } else if constexpr (CORE == O2D_CORE_V2) {
    const double r = rsq64(__builtin_fma(d2, d2, w));
    k = r;
    bbb = -2.0 * r * r * r * d2;
  } else {
    const double ds6 = __builtin_fma(d2 * d2, d2, w);
    const double r = rcbrt(ds6);
    k = r;
    bbb = -2.0 * d2 * d2 * r * rcp64(ds6);
  }
}

struct KAdj {
  double k, bbb, dks;
};

// Gaussian inside (or at) its compact cutoff, for the adjoint; out of line so that the pair
// loop keeps only the far-field registers (beyond the cutoff K = 1/d^2 and dK/dsigma = 0)
__device__ __noinline__ KAdj gauss_adj_core64(double dx, double dy, double d2, double sr) {
  const double s2 = rterm<O2D_CORE_GAUSSIAN>(sr);
  const double cf = rcp64(s2);
  const double x = d2 * cf;
  const int br = gauss_branch(x, dx, dy, s2);
  if (br < 0) return KAdj{cf, 0.0, -2.0 * cf * rcp64(sr)};
  const double ood2 = rcp64(d2);
  if (br > 0) return KAdj{ood2, -2.0 * ood2 * ood2, 0.0};
  const double e = exp(-x);
  const double r = ood2 * (1.0 - e);
  return KAdj{r, -2.0 * r * ood2 + 2.0 * cf * ood2 * e, -2.0 * e * cf * rcp64(sr)};
}

// K, bbb consistent with the branch of K actually taken, and dK/dsigma_source, for the adjoint
// of point-target sums (t2 = 0).
template <int CORE>
__device__ __forceinline__ void core_adj64(double dx, double dy, double d2, double sr, double& k,
                                           double& bbb, double& dks) {
  const double s2 = sr * sr;
  if constexpr (CORE == O2D_CORE_RM) {
    const double r = rcp64(d2 + s2);
    k = r;
    bbb = -2.0 * r * r;
    dks = -2.0 * sr * r * r;
  } else if constexpr (CORE == O2D_CORE_GAUSSIAN) {
    if (d2 > kGaussFar * (1.0 + kCutMargin) * s2) {
      const double ood2 = rcp64(d2);
      k = ood2;
      bbb = -2.0 * ood2 * ood2;
      dks = 0.0;
    } else {
      const KAdj a = gauss_adj_core64(dx, dy, d2, sr);
      k = a.k;
      bbb = a.bbb;
      dks = a.dks;
    }
  } else if constexpr (CORE == O2D_CORE_WL) {
    const double dd = d2 + s2;
    const double r1 = rcp64(dd);
    const double r3 = r1 * r1 * r1;
    k = (dd + s2) * r1 * r1;
    bbb = -2.0 * (dd + 2.0 * s2) * r3;
    dks = -4.0 * s2 * sr * r3;
  } else if constexpr (CORE == O2D_CORE_V2) {
    const double r = rsq64(__builtin_fma(d2, d2, s2 * s2));
    const double r3 = r * r * r;
    k = r;
    bbb = -2.0 * r3 * d2;
    dks = -2.0 * s2 * sr * r3;
  } else {
    const double ds6 = d2 * d2 * d2 + s2 * s2 * s2;
    const double r = rcbrt(ds6);
    const double q = r * rcp64(ds6);
    k = r;
    bbb = -2.0 * d2 * d2 * q;
    dks = -2.0 * s2 * s2 * sr * q;
  }
}

// ---- fp32 split-precision velocity factor -------------------------------------------------

// 1 - e^-x for 0.001 <= x <= 16 without cancellation: Taylor series below 0.25 (truncation
// < 2e-9 relative there), otherwise 1 - exp2(-x log2 e).
__device__ __forceinline__ float one_minus_expm(float x) {
  if (x < 0.25f) {
    float p = __builtin_fmaf(x, -1.0f / 720.0f, 1.0f / 120.0f);
    p = __builtin_fmaf(x, p, -1.0f / 24.0f);
    p = __builtin_fmaf(x, p, 1.0f / 6.0f);
    p = __builtin_fmaf(x, p, -0.5f);
    p = __builtin_fmaf(x, p, 1.0f);
    return x * p;
  }
  return 1.0f - __builtin_amdgcn_exp2f(-x * 1.4426950408889634f);
}

// Inside the Gaussian core (d2 <= 16 w, rare for all but the nearest pairs); kept out of line
// so the unrolled far-field loop does not carry its registers.
__device__ __forceinline__ float gauss_core32(float d2, float w, float rd2) {
  const float cf = __builtin_amdgcn_rcpf(w);
  const float x = d2 * cf;
  return (x < 0.001f) ? cf : one_minus_expm(x) * rd2;
}

// w = radius term of the source plus that of the target (see rterm)
template <int CORE>
__device__ __forceinline__ float core_k32(float dx, float dy, float w) {
  if constexpr (CORE == O2D_CORE_RM) {
    return __builtin_amdgcn_rcpf(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, w)));
  } else if constexpr (CORE == O2D_CORE_GAUSSIAN) {
    const float d2 = __builtin_fmaf(dx, dx, dy * dy);
    float k = __builtin_amdgcn_rcpf(d2);
    if (d2 <= 16.0f * w) k = gauss_core32(d2, w, k);
    return k;
  } else if constexpr (CORE == O2D_CORE_WL) {
    const float dd = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, w));
    const float r = __builtin_amdgcn_rcpf(dd);
    return (dd + w) * r * r;
  } else if constexpr (CORE == O2D_CORE_V2) {
    const float d2 = __builtin_fmaf(dx, dx, dy * dy);
    return __builtin_amdgcn_rsqf(__builtin_fmaf(d2, d2, w));
  } else {
    const float d2 = __builtin_fmaf(dx, dx, dy * dy);
    const float v = __builtin_fmaf(d2 * d2, d2, w);
    return __builtin_amdgcn_exp2f(-0.33333333333f * __builtin_amdgcn_logf(v));
  }
}

// ---- fp32 split-precision vorticity and shear factors --------------------------------------
//
// With bbb = 2 dK/d(d^2), the reference's per-pair vorticity (bbb dx^2 + K) - (-bbb dy^2 - K)
// (src/Kernels.h:139-179) is bbb d^2 + 2K, and its per-pair shear
// sqrt((bbb dx dy)^2 + (bbb (dx^2 - dy^2))^2 / 4) (src/Kernels.h:185-231) is |bbb| d^2 / 2, since
// dx^2 dy^2 + (dx^2 - dy^2)^2 / 4 = d^4 / 4. Each core's pair of factors is written here in a
// form free of cancellation:
//   RM  K = 1/(d^2+w):           vort 2 w K^2,          shear d^2 K^2
//   WL  K = (dd+w)/dd^2, dd=d^2+w: vort 4 w^2 / dd^3,   shear (d^2 + 3w) d^2 / dd^3
//   V2  K = (d^4+w)^-1/2:        vort 2 w K^3,          shear d^4 K^3
//   V3  K = (d^6+w)^-1/3:        vort 2 w K^4,          shear d^6 K^4
//   Gaussian, x = d^2/w:         vort 2 e^-x / w,       shear |(1+x) e^-x - 1| / (x w)
//     (for x < 0.001 the reference's K is the constant 1/w, which makes its shear (1 - e^-x)/w)
// Beyond the Gaussian's compact cutoff (x > 16) K = 1/d^2, the shear factor is K and the
// vorticity factor 2 e^-x / w is below 2.3e-7 of the in-core value 2/w; the split form drops it
// (the reference's own fp32 pair arithmetic rounds the same term at that level).
// Returns K; aux receives the vorticity (RES == O2D_VELANDVORT) or shear factor, to be scaled by
// the source strength (shear by its magnitude).

// (1+x) e^-x - 1 = -x^2 (1/2 - x/3 + x^2/8 - x^3/30 + ...) for small x, as |.| / x
__device__ __forceinline__ float gauss_shear_series(float x) {
  // coefficients (m - 1) / m! of x^(m-2), m = 2..12, alternating; truncation below 1e-8
  // relative for x < 1
  float p = __builtin_fmaf(x, 11.0f / 479001600.0f, -10.0f / 39916800.0f);
  p = __builtin_fmaf(x, p, 9.0f / 3628800.0f);
  p = __builtin_fmaf(x, p, -8.0f / 362880.0f);
  p = __builtin_fmaf(x, p, 7.0f / 40320.0f);
  p = __builtin_fmaf(x, p, -6.0f / 5040.0f);
  p = __builtin_fmaf(x, p, 5.0f / 720.0f);
  p = __builtin_fmaf(x, p, -4.0f / 120.0f);
  p = __builtin_fmaf(x, p, 3.0f / 24.0f);
  p = __builtin_fmaf(x, p, -2.0f / 6.0f);
  p = __builtin_fmaf(x, p, 0.5f);
  return x * p;
}

template <int CORE, int RES>
__device__ __forceinline__ float core_kw32(float dx, float dy, float w, float& aux) {
  const float d2 = __builtin_fmaf(dx, dx, dy * dy);
  if constexpr (CORE == O2D_CORE_RM) {
    const float k = __builtin_amdgcn_rcpf(d2 + w);
    aux = (RES == O2D_VELANDVORT ? 2.0f * w : d2) * k * k;
    return k;
  } else if constexpr (CORE == O2D_CORE_GAUSSIAN) {
    const float rd2 = __builtin_amdgcn_rcpf(d2);
    if (d2 > 16.0f * w) {
      aux = RES == O2D_VELANDVORT ? 0.0f : rd2;
      return rd2;
    }
    const float cf = __builtin_amdgcn_rcpf(w);
    const float x = d2 * cf;
    const float e = __builtin_amdgcn_exp2f(-x * 1.4426950408889634f);
    float k;
    if (x < 0.001f) {
      k = cf;
      // 1 - e^-x = x (1 - x/2) to fp32 precision here
      aux = RES == O2D_VELANDVORT ? 2.0f * cf * e : cf * x * __builtin_fmaf(-0.5f, x, 1.0f);
    } else {
      k = one_minus_expm(x) * rd2;
      if constexpr (RES == O2D_VELANDVORT)
        aux = 2.0f * cf * e;
      else
        aux = cf * (x < 1.0f ? gauss_shear_series(x)
                             : __builtin_fabsf(__builtin_fmaf(1.0f + x, e, -1.0f)) *
                                   __builtin_amdgcn_rcpf(x));
    }
    return k;
  } else if constexpr (CORE == O2D_CORE_WL) {
    const float dd = d2 + w;
    const float r = __builtin_amdgcn_rcpf(dd);
    const float r3 = r * r * r;
    aux = (RES == O2D_VELANDVORT ? 4.0f * w * w : __builtin_fmaf(3.0f, w, d2) * d2) * r3;
    return (dd + w) * r * r;
  } else if constexpr (CORE == O2D_CORE_V2) {
    const float k = __builtin_amdgcn_rsqf(__builtin_fmaf(d2, d2, w));
    aux = (RES == O2D_VELANDVORT ? 2.0f * w : d2 * d2) * k * k * k;
    return k;
  } else {
    // the exp2/log2 estimate of v^-1/3 carries the log's error times |log2 v| / 3; one Newton
    // step y (4 - v y^3) / 3 brings it to fp32 rounding before the fourth power
    const float d6 = d2 * d2 * d2;
    const float v = d6 + w;
    const float k0 = __builtin_amdgcn_exp2f(-0.33333333333f * __builtin_amdgcn_logf(v));
    const float k = k0 * __builtin_fmaf(-v, k0 * k0 * k0, 4.0f) * 0.33333334f;
    const float k2 = k * k;
    aux = (RES == O2D_VELANDVORT ? 2.0f * w : d6) * k2 * k2;
    return k;
  }
}

}  // namespace o2d
