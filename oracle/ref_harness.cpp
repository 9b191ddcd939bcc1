// ref_harness.cpp - a C-ABI shim over the Omega2D reference's own templates.
//
// TEST INFRASTRUCTURE ONLY. Compiled by oracle/Makefile against the reference headers where
// they lie (src/Influence.h, src/Coefficients.h, src/RHS.h, src/Reflect.h, src/Points.h,
// src/Surfaces.h, src/InfluenceVort.h),
// never copied. The shared library lands in oracle/_ref/ and is used only to write the golden
// fixtures under tests/golden/, to validate the C restatement in oracle/o2d_oracle.c, and as
// the "reference" CPU baseline timed beside the GPU numbers. Nothing in the product links it.
//
// The reference core function is a compile-time choice (src/CoreFunc.h:35-39, Vatistas n=2 is
// hard-wired). oracle/Makefile builds this shim once with the headers as they are (V2) and once
// per other core with only that define switched; ref_core_id() reports which core a build
// carries, using the numbering of the oracle and the HIP library (1 RM, 2 Gaussian, 3 WL, 4 V2,
// 5 V3).

#include "Collection.h"
#include "Influence.h"
#include "Coefficients.h"
#include "RHS.h"
#include "Reflect.h"
#include "InfluenceVort.h"

#include <unistd.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

namespace {

// Surfaces with the protected unknown-count switch exposed (src/Surfaces.h:69,260,1411)
template <class S>
struct HarnessSurfaces : public Surfaces<S> {
  using Surfaces<S>::Surfaces;
  void set_vortex_only() { this->source_str_is_unknown = false; }
};

template <class S>
Points<S> make_points(int64_t n, const S* x, const S* y, const S* str, const S* rad, bool inert_pts) {
  std::vector<S> xs(2 * n);
  for (int64_t i = 0; i < n; ++i) { xs[2 * i] = x[i]; xs[2 * i + 1] = y[i]; }
  std::vector<S> vals;
  if (!inert_pts) vals.assign(str, str + n);
  ElementPacket<S> pk(xs, std::vector<Int>(), vals, (size_t)n, 0);
  Points<S> p(pk, inert_pts ? inert : active, lagrangian, nullptr, 0.01f);
  if (!inert_pts) {
    Vector<S>& r = p.get_rad();
    for (int64_t i = 0; i < n; ++i) r[i] = rad[i];
  }
  return p;
}

// Panels built from float node coordinates (the reference Surfaces ctor takes ElementPacket<float>)
template <class S>
HarnessSurfaces<S> make_surfaces(int64_t nnodes, const float* x, const float* y, int64_t np,
                                 const uint32_t* idx, elem_t e, move_t m,
                                 std::shared_ptr<Body> body) {
  std::vector<float> xs(2 * nnodes);
  for (int64_t i = 0; i < nnodes; ++i) { xs[2 * i] = x[i]; xs[2 * i + 1] = y[i]; }
  std::vector<Int> id(idx, idx + 2 * np);
  std::vector<float> vals(np, 0.0f);
  ElementPacket<float> pk(xs, id, vals, (size_t)np, 1);
  return HarnessSurfaces<S>(pk, e, m, body);
}

template <class S>
void copy_out(const Vector<S>& v, S* out) { std::memcpy(out, v.data(), v.size() * sizeof(S)); }

template <class S, class A>
int pts_on_pts(int64_t ns, const S* sx, const S* sy, const S* ss, const S* sr,
               int64_t nt, const S* tx, const S* ty, const S* tr, int restype,
               S* tu, S* tv, S* tw, S* te) {
  Points<S> src = make_points<S>(ns, sx, sy, ss, sr, false);
  const bool targ_inert = (tr == nullptr);
  std::vector<S> dummy(nt, S(0));
  Points<S> targ = make_points<S>(nt, tx, ty, targ_inert ? nullptr : dummy.data(), tr, targ_inert);
  targ.zero_vels();
  if (restype == velandvort) (void)targ.get_vort();
  if (restype == velandshear) (void)targ.get_shear();
  points_affect_points<S, A>(src, targ, ResultsType((results_t)restype), ExecEnv(true, direct, cpu_x86));
  copy_out<S>(targ.get_vel()[0], tu);
  copy_out<S>(targ.get_vel()[1], tv);
  if (restype == velandvort && tw) copy_out<S>(targ.get_vort(), tw);
  if (restype == velandshear && te) copy_out<S>(targ.get_shear(), te);
  return 0;
}

template <class S, class A>
int pan_on_pts(int64_t nnodes, const float* nx, const float* ny, int64_t np, const uint32_t* idx,
               const S* vs, const S* ss, int64_t nt, const S* tx, const S* ty, S* tu, S* tv) {
  HarnessSurfaces<S> src = make_surfaces<S>(nnodes, nx, ny, np, idx, active, fixed, nullptr);
  for (int64_t j = 0; j < np; ++j) {
    src.get_str()[j] = vs[j];
    src.get_src_str()[j] = ss ? ss[j] : S(0);
  }
  Points<S> targ = make_points<S>(nt, tx, ty, nullptr, nullptr, true);
  targ.zero_vels();
  panels_affect_points<S, A>(src, targ, ResultsType(velonly), ExecEnv(true, direct, cpu_x86));
  copy_out<S>(targ.get_vel()[0], tu);
  copy_out<S>(targ.get_vel()[1], tv);
  return 0;
}

template <class S, class A>
int pts_on_pan(int64_t ns, const S* sx, const S* sy, const S* ss, const S* sr,
               int64_t nnodes, const float* nx, const float* ny, int64_t np, const uint32_t* idx,
               S* pu, S* pv) {
  Points<S> src = make_points<S>(ns, sx, sy, ss, sr, false);
  HarnessSurfaces<S> targ = make_surfaces<S>(nnodes, nx, ny, np, idx, reactive, fixed, nullptr);
  targ.zero_vels();
  points_affect_panels<S, A>(src, targ, ResultsType(velonly), ExecEnv(true, direct, cpu_x86));
  copy_out<S>(targ.get_vel()[0], pu);
  copy_out<S>(targ.get_vel()[1], pv);
  return 0;
}

// Full BEM block for one reactive body acting on itself, exactly as solve_bem builds it
// (src/BEMHelper.h:202-216): unit-rotation velocities for the augmented column first, then
// panels_on_panels_coeff. body_kind: 0 = no Body (never augmented), 1 = "ground" body,
// 2 = a named moving body. Also returns the panel bases/areas, the body volume and the
// unit-rotation panel velocities so the restatement can be fed the same inputs.
template <class S, class A>
int bem_coeffs(int64_t nnodes, const float* nx, const float* ny, int64_t np, const uint32_t* idx,
               int nunk, int body_kind, S* coeffs, int64_t* nrows_out, int64_t* ncols_out,
               S* tang_x, S* tang_y, S* area, S* vol_out, S* urot_u, S* urot_v) {
  std::shared_ptr<Body> body;
  if (body_kind == 1) { body = std::make_shared<Body>(); body->set_name("ground"); }
  if (body_kind == 2) { body = std::make_shared<Body>(); body->set_name("rotor"); }
  HarnessSurfaces<S> surf = make_surfaces<S>(nnodes, nx, ny, np, idx, reactive,
                                             body ? bodybound : fixed, body);
  if (nunk == 1) surf.set_vortex_only();
  surf.transform(0.0);
  if (surf.is_augmented()) {
    surf.zero_vels();
    surf.zero_strengths();
    surf.add_unit_rot_strengths();
    // panels_affect_panels (src/Influence.h:774-795) only compiles at S=float, because it builds
    // Points<S> from an ElementPacket<float>; the same steps are spelled out here with the
    // packet converted to S, calling the reference's own represent_as_particles and
    // panels_affect_points.
    ElementPacket<float> fp = surf.represent_as_particles(0.0001, -1.0);
    ElementPacket<S> sp(std::vector<S>(fp.x.begin(), fp.x.end()), fp.idx,
                        std::vector<S>(fp.val.begin(), fp.val.end()), fp.nelem, fp.ndim);
    Points<S> temppts(sp, active, lagrangian, nullptr, 0.0001f);
    temppts.zero_vels();
    const Surfaces<S>& csrc = surf;
    panels_affect_points<S, A>(csrc, temppts, ResultsType(velonly), ExecEnv(true, direct, cpu_x86));
    for (size_t d = 0; d < Dimensions; ++d)
      for (int64_t i = 0; i < np; ++i) surf.get_vel()[d][i] += temppts.get_vel()[d][i];
    surf.zero_strengths();
    surf.finalize_vels(std::array<double, Dimensions>({0.0, 0.0}));
  }
  for (int64_t i = 0; i < np; ++i) {
    urot_u[i] = surf.get_vel()[0][i];
    urot_v[i] = surf.get_vel()[1][i];
  }
  const Surfaces<S>& csurf = surf;
  Vector<S> c = panels_on_panels_coeff<S>(csurf, static_cast<Surfaces<S>&>(surf));
  const int64_t nr = surf.get_num_rows();
  *nrows_out = nr;
  *ncols_out = nr;
  if (coeffs) std::memcpy(coeffs, c.data(), c.size() * sizeof(S));
  copy_out<S>(surf.get_tang()[0], tang_x);
  copy_out<S>(surf.get_tang()[1], tang_y);
  copy_out<S>(surf.get_area(), area);
  *vol_out = surf.get_vol();
  return (int64_t)c.size() == nr * nr ? 0 : -1;
}

// RHS from panel-centre velocities (src/RHS.h:54-115) for a reactive surface with given normal BCs
// One off-diagonal block, source surface on a different target surface, exactly as solve_bem
// builds it (src/BEMHelper.h:184-216): when the source body is augmented, the velocity of its
// unit-rotation strengths at the target's collocation points, then panels_on_panels_coeff.
// kind: 0 = no Body, 2 = a named rotating body (augmented). Returns the target velocities in
// urot_u/urot_v (zeros when the source is not augmented).
template <class S, class A>
int bem_block(int64_t snn, const float* snx, const float* sny, int64_t snp, const uint32_t* sidx,
              int skind, int64_t tnn, const float* tnx, const float* tny, int64_t tnp,
              const uint32_t* tidx, int tkind, int nunk, S* coeffs, int64_t* nrows_out,
              int64_t* ncols_out, S* urot_u, S* urot_v) {
  std::shared_ptr<Body> sb, tb;
  if (skind == 2) { sb = std::make_shared<Body>(); sb->set_name("rotor_a"); }
  if (tkind == 2) { tb = std::make_shared<Body>(); tb->set_name("rotor_b"); }
  HarnessSurfaces<S> src = make_surfaces<S>(snn, snx, sny, snp, sidx, reactive,
                                            sb ? bodybound : fixed, sb);
  HarnessSurfaces<S> targ = make_surfaces<S>(tnn, tnx, tny, tnp, tidx, reactive,
                                             tb ? bodybound : fixed, tb);
  if (nunk == 1) {
    src.set_vortex_only();
    targ.set_vortex_only();
  }
  src.transform(0.0);
  targ.transform(0.0);
  targ.zero_vels();
  if (src.is_augmented()) {
    src.zero_strengths();
    src.add_unit_rot_strengths();
    ElementPacket<float> fp = targ.represent_as_particles(0.0001, -1.0);
    ElementPacket<S> sp(std::vector<S>(fp.x.begin(), fp.x.end()), fp.idx,
                        std::vector<S>(fp.val.begin(), fp.val.end()), fp.nelem, fp.ndim);
    Points<S> temppts(sp, active, lagrangian, nullptr, 0.0001f);
    temppts.zero_vels();
    const Surfaces<S>& csrc = src;
    panels_affect_points<S, A>(csrc, temppts, ResultsType(velonly), ExecEnv(true, direct, cpu_x86));
    for (size_t d = 0; d < Dimensions; ++d)
      for (int64_t i = 0; i < tnp; ++i) targ.get_vel()[d][i] += temppts.get_vel()[d][i];
    src.zero_strengths();
    targ.finalize_vels(std::array<double, Dimensions>({0.0, 0.0}));
  }
  for (int64_t i = 0; i < tnp; ++i) {
    urot_u[i] = targ.get_vel()[0][i];
    urot_v[i] = targ.get_vel()[1][i];
  }
  const Surfaces<S>& csrc = src;
  Vector<S> c = panels_on_panels_coeff<S>(csrc, static_cast<Surfaces<S>&>(targ));
  *nrows_out = targ.get_num_rows();
  *ncols_out = src.get_num_rows();
  if ((int64_t)c.size() != *nrows_out * *ncols_out) return -1;
  std::memcpy(coeffs, c.data(), c.size() * sizeof(S));
  return 0;
}

template <class S>
int rhs_panels(int64_t nnodes, const float* nx, const float* ny, int64_t np, const uint32_t* idx,
               int nunk, const S* pu, const S* pv, const float* norm_bc, S* rhs) {
  std::vector<float> xs(2 * nnodes);
  for (int64_t i = 0; i < nnodes; ++i) { xs[2 * i] = nx[i]; xs[2 * i + 1] = ny[i]; }
  std::vector<Int> id(idx, idx + 2 * np);
  std::vector<float> vals(2 * np);
  for (int64_t i = 0; i < np; ++i) { vals[2 * i] = 0.0f; vals[2 * i + 1] = norm_bc[i]; }
  ElementPacket<float> pk(xs, id, vals, (size_t)np, 1);
  HarnessSurfaces<S> surf(pk, reactive, fixed, nullptr);
  if (nunk == 1) surf.set_vortex_only();
  for (int64_t i = 0; i < np; ++i) { surf.get_vel()[0][i] = pu[i]; surf.get_vel()[1][i] = pv[i]; }
  std::vector<S> r = vels_to_rhs_panels<S>(surf);
  std::memcpy(rhs, r.data(), r.size() * sizeof(S));
  return 0;
}

}  // namespace

extern "C" {

int ref_core_id() {
  int id = 0, n = 0;
#ifdef USE_RM_KERNEL
  id = 1; ++n;
#endif
#ifdef USE_EXPONENTIAL_KERNEL
  id = 2; ++n;
#endif
#ifdef USE_WL_KERNEL
  id = 3; ++n;
#endif
#ifdef USE_V2_KERNEL
  id = 4; ++n;
#endif
#ifdef USE_V3_KERNEL
  id = 5; ++n;
#endif
  return n == 1 ? id : -n;
}

int ref_points_on_points_d(int64_t ns, const double* sx, const double* sy, const double* ss,
                           const double* sr, int64_t nt, const double* tx, const double* ty,
                           const double* tr, int restype, double* tu, double* tv, double* tw,
                           double* te) {
  return pts_on_pts<double, double>(ns, sx, sy, ss, sr, nt, tx, ty, tr, restype, tu, tv, tw, te);
}

int ref_points_on_points_f(int64_t ns, const float* sx, const float* sy, const float* ss,
                           const float* sr, int64_t nt, const float* tx, const float* ty,
                           const float* tr, int restype, float* tu, float* tv, float* tw,
                           float* te) {
  return pts_on_pts<float, double>(ns, sx, sy, ss, sr, nt, tx, ty, tr, restype, tu, tv, tw, te);
}

int ref_panels_on_points_d(int64_t nnodes, const float* nx, const float* ny, int64_t np,
                           const uint32_t* idx, const double* vs, const double* ss, int64_t nt,
                           const double* tx, const double* ty, double* tu, double* tv) {
  return pan_on_pts<double, double>(nnodes, nx, ny, np, idx, vs, ss, nt, tx, ty, tu, tv);
}

int ref_panels_on_points_f(int64_t nnodes, const float* nx, const float* ny, int64_t np,
                           const uint32_t* idx, const float* vs, const float* ss, int64_t nt,
                           const float* tx, const float* ty, float* tu, float* tv) {
  return pan_on_pts<float, double>(nnodes, nx, ny, np, idx, vs, ss, nt, tx, ty, tu, tv);
}

int ref_points_on_panels_d(int64_t ns, const double* sx, const double* sy, const double* ss,
                           const double* sr, int64_t nnodes, const float* nx, const float* ny,
                           int64_t np, const uint32_t* idx, double* pu, double* pv) {
  return pts_on_pan<double, double>(ns, sx, sy, ss, sr, nnodes, nx, ny, np, idx, pu, pv);
}

int ref_points_on_panels_f(int64_t ns, const float* sx, const float* sy, const float* ss,
                           const float* sr, int64_t nnodes, const float* nx, const float* ny,
                           int64_t np, const uint32_t* idx, float* pu, float* pv) {
  return pts_on_pan<float, double>(ns, sx, sy, ss, sr, nnodes, nx, ny, np, idx, pu, pv);
}

int ref_bem_coeffs_d(int64_t nnodes, const float* nx, const float* ny, int64_t np,
                     const uint32_t* idx, int nunk, int body_kind, double* coeffs,
                     int64_t* nrows, int64_t* ncols, double* tang_x, double* tang_y,
                     double* area, double* vol, double* urot_u, double* urot_v) {
  return bem_coeffs<double, double>(nnodes, nx, ny, np, idx, nunk, body_kind, coeffs, nrows,
                                    ncols, tang_x, tang_y, area, vol, urot_u, urot_v);
}

int ref_bem_coeffs_f(int64_t nnodes, const float* nx, const float* ny, int64_t np,
                     const uint32_t* idx, int nunk, int body_kind, float* coeffs,
                     int64_t* nrows, int64_t* ncols, float* tang_x, float* tang_y,
                     float* area, float* vol, float* urot_u, float* urot_v) {
  return bem_coeffs<float, double>(nnodes, nx, ny, np, idx, nunk, body_kind, coeffs, nrows,
                                   ncols, tang_x, tang_y, area, vol, urot_u, urot_v);
}

int ref_rhs_panels_d(int64_t nnodes, const float* nx, const float* ny, int64_t np,
                     const uint32_t* idx, int nunk, const double* pu, const double* pv,
                     const float* norm_bc, double* rhs) {
  return rhs_panels<double>(nnodes, nx, ny, np, idx, nunk, pu, pv, norm_bc, rhs);
}

// finalize_vels (src/ElementBase.h:215-234) followed by one Ralston RK2 position update
// (src/ElementBase.h:299-327 with weights 1/4, 3/4 from src/Convection.h:442)
int ref_points_move2_d(int64_t n, double* x, double* y, const double* u1, const double* v1,
                       const double* u2, const double* v2, double dt, double w1, double w2) {
  std::vector<double> s(n, 1.0), r(n, 0.01);
  Points<double> p0 = make_points<double>(n, x, y, s.data(), r.data(), false);
  Points<double> p1 = make_points<double>(n, x, y, s.data(), r.data(), false);
  for (int64_t i = 0; i < n; ++i) {
    p0.get_vel()[0][i] = u1[i]; p0.get_vel()[1][i] = v1[i];
    p1.get_vel()[0][i] = u2[i]; p1.get_vel()[1][i] = v2[i];
  }
  p0.move(0.0, dt, w1, p0, w2, p1);
  copy_out<double>(p0.get_pos()[0], x);
  copy_out<double>(p0.get_pos()[1], y);
  return 0;
}

int ref_points_move3_d(int64_t n, double* x, double* y, const double* u1, const double* v1,
                       const double* u2, const double* v2, const double* u3, const double* v3,
                       double dt, double w1, double w2, double w3) {
  std::vector<double> s(n, 1.0), r(n, 0.01);
  Points<double> p0 = make_points<double>(n, x, y, s.data(), r.data(), false);
  Points<double> p1 = make_points<double>(n, x, y, s.data(), r.data(), false);
  Points<double> p2 = make_points<double>(n, x, y, s.data(), r.data(), false);
  for (int64_t i = 0; i < n; ++i) {
    p0.get_vel()[0][i] = u1[i]; p0.get_vel()[1][i] = v1[i];
    p1.get_vel()[0][i] = u2[i]; p1.get_vel()[1][i] = v2[i];
    p2.get_vel()[0][i] = u3[i]; p2.get_vel()[1][i] = v3[i];
  }
  p0.move(0.0, dt, w1, p0, w2, p1, w3, p2);
  copy_out<double>(p0.get_pos()[0], x);
  copy_out<double>(p0.get_pos()[1], y);
  return 0;
}

int ref_finalize_d(int64_t n, double* u, double* v, double fsx, double fsy) {
  std::vector<double> x(n, 0.0), s(n, 1.0), r(n, 0.01);
  Points<double> p = make_points<double>(n, x.data(), x.data(), s.data(), r.data(), false);
  for (int64_t i = 0; i < n; ++i) { p.get_vel()[0][i] = u[i]; p.get_vel()[1][i] = v[i]; }
  p.finalize_vels(std::array<double, Dimensions>({fsx, fsy}));
  copy_out<double>(p.get_vel()[0], u);
  copy_out<double>(p.get_vel()[1], v);
  return 0;
}

// get_nearest_distances (src/Reflect.h:792-826) from inert targets to panels
int ref_nearest_distances_d(int64_t nnodes, const float* nx, const float* ny, int64_t np,
                            const uint32_t* idx, int64_t nt, const double* tx, const double* ty,
                            double* dist) {
  HarnessSurfaces<double> surf =
      make_surfaces<double>(nnodes, nx, ny, np, idx, reactive, fixed, nullptr);
  Points<double> targ = make_points<double>(nt, tx, ty, nullptr, nullptr, true);
  Vector<double> d = get_nearest_distances<double>(surf, targ);
  copy_out<double>(d, dist);
  return 0;
}

// the panels' unit normals as Surfaces::compute_bases stores them (src/Surfaces.h:682-725)
int ref_panel_normals_d(int64_t nnodes, const float* nx, const float* ny, int64_t np,
                        const uint32_t* idx, double* pnx, double* pny) {
  HarnessSurfaces<double> surf =
      make_surfaces<double>(nnodes, nx, ny, np, idx, reactive, fixed, nullptr);
  copy_out<double>(surf.get_norm()[0], pnx);
  copy_out<double>(surf.get_norm()[1], pny);
  return 0;
}

// reflect_panp2 (src/Reflect.h:167-311) on particles (radius 0.01, strength 1)
int ref_reflect_d(int64_t nnodes, const float* nx, const float* ny, int64_t np,
                  const uint32_t* idx, int64_t nt, double* tx, double* ty) {
  HarnessSurfaces<double> surf =
      make_surfaces<double>(nnodes, nx, ny, np, idx, reactive, fixed, nullptr);
  std::vector<double> s(nt, 1.0), r(nt, 0.01);
  Points<double> targ = make_points<double>(nt, tx, ty, s.data(), r.data(), false);
  reflect_panp2<double>(surf, targ);
  copy_out<double>(targ.get_pos()[0], tx);
  copy_out<double>(targ.get_pos()[1], ty);
  return 0;
}

// clear_inner_panp2 (src/Reflect.h:487-727). tr == NULL makes the targets field points (no
// radius); method 0 on field points below the layer never returns in the reference, so the
// caller must not ask for it. Returns the removed circulation in *circ.
int ref_clear_inner_d(int method, int64_t nnodes, const float* nx, const float* ny, int64_t np,
                      const uint32_t* idx, double cutoff_mult, double ips, int64_t nt,
                      double* tx, double* ty, double* ts, const double* tr, double* circ) {
  HarnessSurfaces<double> surf =
      make_surfaces<double>(nnodes, nx, ny, np, idx, reactive, fixed, nullptr);
  std::vector<double> s(nt, 0.0);
  if (ts) s.assign(ts, ts + nt);
  Points<double> targ = make_points<double>(nt, tx, ty, tr ? s.data() : nullptr, tr, tr == nullptr);
  *circ = clear_inner_panp2<double>(method, surf, targ, cutoff_mult, ips);
  copy_out<double>(targ.get_pos()[0], tx);
  copy_out<double>(targ.get_pos()[1], ty);
  if (tr && ts) copy_out<double>(targ.get_str(), ts);
  return 0;
}

// init_cut_tables<double>(0.1) (src/Reflect.h:333-394); returns the row count
int ref_cut_table_d(int cap, double* pos, double* smult, double* dshift) {
  std::vector<std::tuple<double, double, double>> ct = init_cut_tables<double>(0.1);
  if ((int)ct.size() > cap) return -(int)ct.size();
  for (size_t i = 0; i < ct.size(); ++i) {
    pos[i] = std::get<0>(ct[i]);
    smult[i] = std::get<1>(ct[i]);
    dshift[i] = std::get<2>(ct[i]);
  }
  return (int)ct.size();
}

// points_affect_points_vorticity<S,A> (src/InfluenceVort.h:47-126), x86 environment; targets
// are particles when tr is given, field points otherwise (only their positions are read)
int ref_points_vorticity_d(int64_t ns, const double* sx, const double* sy, const double* ss,
                           const double* sr, int64_t nt, const double* tx, const double* ty,
                           double* tw) {
  Points<double> src = make_points<double>(ns, sx, sy, ss, sr, false);
  std::vector<double> z(nt, 0.0), r(nt, 0.01);
  Points<double> targ = make_points<double>(nt, tx, ty, z.data(), r.data(), false);
  targ.zero_vels();
  Vector<double>& w = targ.get_vort();
  std::fill(w.begin(), w.end(), 0.0);
  points_affect_points_vorticity<double, double>(src, targ, ExecEnv(true, direct, cpu_x86));
  copy_out<double>(targ.get_vort(), tw);
  return 0;
}

int ref_points_vorticity_f(int64_t ns, const float* sx, const float* sy, const float* ss,
                           const float* sr, int64_t nt, const float* tx, const float* ty,
                           float* tw) {
  Points<float> src = make_points<float>(ns, sx, sy, ss, sr, false);
  std::vector<float> z(nt, 0.0f), r(nt, 0.01f);
  Points<float> targ = make_points<float>(nt, tx, ty, z.data(), r.data(), false);
  targ.zero_vels();
  Vector<float>& w = targ.get_vort();
  std::fill(w.begin(), w.end(), 0.0f);
  points_affect_points_vorticity<float, double>(src, targ, ExecEnv(true, direct, cpu_x86));
  copy_out<float>(targ.get_vort(), tw);
  return 0;
}

// Points<float>::write_vtk (src/Points.h:556-740) run inside directory `dir`; particles (with
// strengths and radii) unless inert != 0; w (vorticity) may be NULL. The file name is copied to
// `name` (capacity cap). Float storage, as the application (STORE=float) writes its files.
int ref_write_points_vtk_f(const char* dir, int64_t n, const float* x, const float* y,
                           const float* s, const float* r, const float* u, const float* v,
                           const float* w, int inert, int64_t index, int64_t frameno,
                           double time, char* name, int64_t cap) {
  Points<float> p = make_points<float>(n, x, y, inert ? nullptr : s, inert ? nullptr : r,
                                       inert != 0);
  p.zero_vels();
  for (int64_t i = 0; i < n; ++i) { p.get_vel()[0][i] = u[i]; p.get_vel()[1][i] = v[i]; }
  if (w) {
    Vector<float>& pw = p.get_vort();
    for (int64_t i = 0; i < n; ++i) pw[i] = w[i];
  }
  char cwd[4096];
  if (!getcwd(cwd, sizeof(cwd)) || chdir(dir) != 0) return -1;
  const std::string fn = p.write_vtk((size_t)index, (size_t)frameno, time);
  if (chdir(cwd) != 0) return -1;
  std::snprintf(name, (size_t)cap, "%s", fn.c_str());
  return 0;
}

// Surfaces<float>::write_vtk (src/Surfaces.h:1238-1379) of reactive panels with vortex and
// source sheet strengths, inside directory `dir`
int ref_write_panels_vtk_f(const char* dir, int64_t nnodes, const float* nx, const float* ny,
                           int64_t np, const uint32_t* idx, const float* vs, const float* ss,
                           int64_t index, int64_t frameno, double time, char* name, int64_t cap) {
  HarnessSurfaces<float> surf = make_surfaces<float>(nnodes, nx, ny, np, idx, reactive, fixed,
                                                     nullptr);
  for (int64_t j = 0; j < np; ++j) {
    surf.get_str()[j] = vs[j];
    surf.get_src_str()[j] = ss[j];
  }
  char cwd[4096];
  if (!getcwd(cwd, sizeof(cwd)) || chdir(dir) != 0) return -1;
  const std::string fn = surf.write_vtk((size_t)index, (size_t)frameno, time);
  if (chdir(cwd) != 0) return -1;
  std::snprintf(name, (size_t)cap, "%s", fn.c_str());
  return 0;
}

int ref_bem_block_d(int64_t snn, const float* snx, const float* sny, int64_t snp,
                    const uint32_t* sidx, int skind, int64_t tnn, const float* tnx,
                    const float* tny, int64_t tnp, const uint32_t* tidx, int tkind, int nunk,
                    double* coeffs, int64_t* nrows, int64_t* ncols, double* urot_u,
                    double* urot_v) {
  return bem_block<double, double>(snn, snx, sny, snp, sidx, skind, tnn, tnx, tny, tnp, tidx,
                                   tkind, nunk, coeffs, nrows, ncols, urot_u, urot_v);
}

// A body moving by the tinyexpr expressions xexpr(t), yexpr(t) and rexpr(t) (src/Body.cpp; an
// empty string keeps that coordinate still), its body-bound Surfaces transformed to time t
// (src/ElementBase.h:250-273, src/Surfaces.h:728-752) and the body's motion added with factor f to
// zeroed panel velocities (src/Surfaces.h:492-523). Out: nodes x, y (nnodes); tangents tx, ty
// and panel velocities pu, pv (np); body = pos(t) x, y, vel(t) x, y, orient(t), rotvel(t).
int ref_body_motion_d(int64_t nnodes, const float* nx, const float* ny, int64_t np,
                      const uint32_t* idx, const char* xexpr, const char* yexpr,
                      const char* rexpr, double t, double f, double* ox, double* oy,
                      double* otx, double* oty, double* opu, double* opv, double* body) {
  auto b = std::make_shared<Body>();
  b->set_name("mover");
  if (xexpr && *xexpr) b->set_pos(0, std::string(xexpr));
  if (yexpr && *yexpr) b->set_pos(1, std::string(yexpr));
  if (rexpr && *rexpr) b->set_rot(std::string(rexpr));
  HarnessSurfaces<double> s = make_surfaces<double>(nnodes, nx, ny, np, idx, reactive, bodybound, b);
  s.transform(t);
  s.zero_vels();
  s.add_body_motion(f, t);
  copy_out(s.get_pos()[0], ox);
  copy_out(s.get_pos()[1], oy);
  copy_out(s.get_tang()[0], otx);
  copy_out(s.get_tang()[1], oty);
  copy_out(s.get_vel()[0], opu);
  copy_out(s.get_vel()[1], opv);
  const auto p = b->get_pos(t), v = b->get_vel(t);
  body[0] = p[0];
  body[1] = p[1];
  body[2] = v[0];
  body[3] = v[1];
  body[4] = b->get_orient(t);
  body[5] = b->get_rotvel(t);
  return 0;
}

}  // extern "C"
