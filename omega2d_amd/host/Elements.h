// Elements.h - SoA element collections for the HIP path: particles/tracers (Points) and
// straight two-node panels (Surfaces). Host mirror of the members of src/Points.h,
// src/Surfaces.h and src/ElementBase.h that the direct Biot-Savart path reads and writes;
// drawing, diffusion and feature bookkeeping are not part of it.
#pragma once

#include <algorithm>
#include <array>
#include <cassert>
#include <cmath>
#include <cstdint>
#include <functional>
#include <limits>
#include <memory>
#include <numeric>
#include <optional>
#include <string>
#include <vector>

#include "VtkXmlWriter.h"

namespace omega2d {

constexpr size_t Dimensions = 2;
using Int = uint32_t;

enum elem_t { active = 1, reactive = 2, inert = 3, hybrid = 4 };
enum move_t { lagrangian = 1, bodybound = 2, fixed = 3 };

template <class S>
using Vector = std::vector<S>;

// A rigid body (src/Body.cpp): its name (the "ground" body is never augmented when it bounds
// internal flow and never moves the boundary condition) and its prescribed motion. The
// reference parses position and orientation as expressions of t; here they are callables, and
// an empty one keeps that coordinate at its constant value. Velocities are the same two-point
// differences with step 1e-6 (src/Body.cpp:165-224); transform(t) stores position, velocity,
// orientation (wrapped by remainder(., 2 pi)) and rotation rate for the no-argument getters.
struct Body {
  using Vec = std::array<double, Dimensions>;
  static constexpr double kDT = 1.e-6;

  std::string name = "ground";
  std::array<std::function<double(double)>, Dimensions> pos_func;
  std::function<double(double)> apos_func;
  Vec pos = {0.0, 0.0}, vel = {0.0, 0.0};
  double apos = 0.0, avel = 0.0;

  const std::string& get_name() const { return name; }
  Vec get_pos() const { return pos; }
  Vec get_pos(double t) const {
    Vec p;
    for (size_t i = 0; i < Dimensions; ++i) p[i] = pos_func[i] ? pos_func[i](t) : pos[i];
    return p;
  }
  Vec get_vel() const { return vel; }
  Vec get_vel(double t) const {
    Vec v;
    for (size_t i = 0; i < Dimensions; ++i)
      v[i] = pos_func[i] ? (pos_func[i](t + kDT) - pos_func[i](t - kDT)) / (2.0 * kDT) : vel[i];
    return v;
  }
  double get_orient() const { return apos; }
  double get_orient(double t) const {
    return apos_func ? std::remainder(apos_func(t), 2.0 * M_PI) : apos;
  }
  double get_rotvel() const { return avel; }
  double get_rotvel(double t) const {
    return apos_func ? (apos_func(t + kDT) - apos_func(t - kDT)) / (2.0 * kDT) : avel;
  }
  void transform(double t) {
    pos = get_pos(t);
    vel = get_vel(t);
    apos = get_orient(t);
    avel = get_rotvel(t);
  }
  // src/Body.cpp:237-265: translation or rotation of one body relative to the other between two
  // times, beyond 4 machine epsilons
  bool relative_motion_vs(const Body& o, double last, double current) const {
    const double eps4 = 4.0 * std::numeric_limits<double>::epsilon();
    const Vec a0 = get_pos(last), b0 = o.get_pos(last), a1 = get_pos(current),
              b1 = o.get_pos(current);
    bool motion = false;
    for (size_t i = 0; i < Dimensions; ++i)
      if (std::abs(a1[i] - a0[i] - b1[i] + b0[i]) > eps4) motion = true;
    const double rel = get_orient(current) - get_orient(last) - o.get_orient(current) +
                       o.get_orient(last);
    if (std::abs(rel) > eps4) motion = true;
    return motion;
  }
};

// ---------------------------------------------------------------------------------- Points

template <class S>
class Points {
 public:
  // particles (strengths and radii) or, with str == nullptr, inert tracers
  Points(size_t n, const S* x, const S* y, const S* str, const S* rad, elem_t e = active,
         move_t m = lagrangian)
      : E(str ? e : inert), M(m), n(n) {
    for (size_t d = 0; d < Dimensions; ++d) {
      xs[d].assign(n, S(0));
      u[d].assign(n, S(0));
    }
    for (size_t i = 0; i < n; ++i) {
      xs[0][i] = x[i];
      xs[1][i] = y[i];
    }
    if (str) {
      s.assign(str, str + n);
      r.assign(rad, rad + n);
    }
  }

  size_t get_n() const { return n; }
  bool is_inert() const { return E == inert; }
  move_t get_movet() const { return M; }
  const std::array<Vector<S>, Dimensions>& get_pos() const { return xs; }
  std::array<Vector<S>, Dimensions>& get_pos() { return xs; }
  const Vector<S>& get_str() const { return s; }
  Vector<S>& get_str() { return s; }
  const Vector<S>& get_rad() const { return r; }
  const std::array<Vector<S>, Dimensions>& get_vel() const { return u; }
  std::array<Vector<S>, Dimensions>& get_vel() { return u; }
  Vector<S>& get_vort() {
    if (!w) w = Vector<S>(n, S(0));
    return *w;
  }
  Vector<S>& get_shear() {
    if (!e) e = Vector<S>(n, S(0));
    return *e;
  }
  bool has_vort() const { return (bool)w; }
  bool has_shear() const { return (bool)e; }

  // src/Points.h:556-740: one .vtu file (part_NN_FFFFF.vtu, fldpt_ for inert points) in `dir`
  // (the working directory when empty); returns the file name
  std::string write_vtk(size_t index, size_t frameno, double time,
                        const std::string& dir = "") const {
    const bool particles = E != inert;
    const std::string name = vtu_name(particles ? "part_" : "fldpt_", index, frameno);
    VtuFile f(dir.empty() ? name : dir + "/" + name);
    vtu_time(f, time);
    f.open("Piece", {{"NumberOfPoints", std::to_string(n)}, {"NumberOfCells", std::to_string(n)}});
    f.open("Points");
    f.data_array({{"NumberOfComponents", "3"}, {"Name", "position"}, {"type", "Float32"}},
                 xyz_float(xs[0], xs[1]));
    f.close();
    f.open("Cells");
    std::vector<int32_t> conn(n), off(n);
    std::iota(conn.begin(), conn.end(), 0);
    std::iota(off.begin(), off.end(), 1);
    f.data_array({{"Name", "connectivity"}, {"type", "Int32"}}, conn);
    f.data_array({{"Name", "offsets"}, {"type", "Int32"}}, off);
    f.data_array({{"Name", "types"}, {"type", "UInt8"}}, std::vector<uint8_t>(n, 1));
    f.close();
    std::string scalars;
    if (particles) scalars += "circulation,radius,";
    if (w) scalars += "vorticity,";
    VtkAttrs pd = {{"Vectors", "velocity"}};
    if (!scalars.empty()) {
      scalars.pop_back();
      pd.emplace_back("Scalars", scalars);
    }
    f.open("PointData", pd);
    if (particles) {
      f.data_array({{"Name", "circulation"}, {"type", "Float32"}}, as_float(s));
      f.data_array({{"Name", "radius"}, {"type", "Float32"}}, as_float(r));
    }
    if (w) f.data_array({{"Name", "vorticity"}, {"type", "Float32"}}, as_float(*w));
    if (e) f.data_array({{"Name", "shearrate"}, {"type", "Float32"}}, as_float(*e));
    f.data_array({{"NumberOfComponents", "3"}, {"Name", "velocity"}, {"type", "Float32"}},
                 xyz_float(u[0], u[1]));
    f.finish();
    return name;
  }

  // src/ElementBase.h:200-212
  void zero_vels() {
    for (size_t d = 0; d < Dimensions; ++d) std::fill(u[d].begin(), u[d].end(), S(0));
    if (w) std::fill(w->begin(), w->end(), S(0));
    if (e) std::fill(e->begin(), e->end(), S(0));
  }

  // src/ElementBase.h:215-234: u = fs + u/(2 pi) in double, stored to S
  void finalize_vels(const std::array<double, Dimensions>& fs) {
    const double factor = 0.5 / M_PI;
    for (size_t d = 0; d < Dimensions; ++d)
      for (size_t i = 0; i < n; ++i) u[d][i] = S(fs[d] + double(u[d][i]) * factor);
    if (w)
      for (auto& v : *w) v = S(double(v) * factor);
    if (e)
      for (auto& v : *e) v = S(double(v) * factor);
  }

  std::string to_string() const {
    std::string m = E == active ? " Active" : (E == reactive ? " Reactive" : " Inert");
    m += M == lagrangian ? " Lagrangian" : (M == bodybound ? " Body-fixed" : " Fixed");
    return m + " Points";
  }

 private:
  elem_t E;
  move_t M;
  size_t n;
  std::array<Vector<S>, Dimensions> xs;
  std::array<Vector<S>, Dimensions> u;
  Vector<S> s, r;
  std::optional<Vector<S>> w, e;
};

// -------------------------------------------------------------------------------- Surfaces

template <class S>
class Surfaces {
 public:
  // nodes (nx, ny) are taken as float like the reference's ElementPacket<float>; idx holds two
  // node indices per panel; normal_bc is the per-panel normal-velocity boundary condition of a
  // reactive surface (nullptr: zero). The reference always solves for vortex and source
  // strengths (src/Surfaces.h:69); set_vortex_only() selects one unknown per panel.
  Surfaces(size_t nnodes, const float* nx, const float* ny, size_t npanels, const Int* idx_in,
           elem_t e, move_t m, std::shared_ptr<Body> body, const float* normal_bc = nullptr)
      : E(e), M(m), B(std::move(body)), n(nnodes), np(npanels) {
    for (size_t d = 0; d < Dimensions; ++d) x[d].assign(nnodes, S(0));
    for (size_t i = 0; i < nnodes; ++i) {
      x[0][i] = S(nx[i]);
      x[1][i] = S(ny[i]);
    }
    ux = x;
    idx.assign(idx_in, idx_in + 2 * np);
    for (Int k : idx) assert(k < nnodes && "panel node index out of range");
    compute_bases();
    ps[0] = Vector<S>(np, S(0));
    ps[1] = Vector<S>(np, S(0));
    if (E == reactive) {
      bc[0] = Vector<S>(np, S(0));
      bc[1] = Vector<S>(np, S(0));
      if (normal_bc)
        for (size_t i = 0; i < np; ++i) (*bc[1])[i] = S(normal_bc[i]);
    }
    for (size_t d = 0; d < Dimensions; ++d) pu[d].assign(np, S(0));
    if (B && M == bodybound) set_geom_center();
    tc = utc;
  }

  size_t get_n() const { return n; }
  size_t get_npanels() const { return np; }
  bool is_inert() const { return E == inert; }
  move_t get_movet() const { return M; }
  S get_vol() const { return vol; }
  const std::array<Vector<S>, Dimensions>& get_pos() const { return x; }
  const std::vector<Int>& get_idx() const { return idx; }
  const std::array<Vector<S>, Dimensions>& get_tang() const { return b[0]; }
  const std::array<Vector<S>, Dimensions>& get_norm() const { return b[1]; }
  const Vector<S>& get_area() const { return area; }
  const std::array<Vector<S>, Dimensions>& get_vel() const { return pu; }
  std::array<Vector<S>, Dimensions>& get_vel() { return pu; }
  const Vector<S>& get_str() const { return *ps[0]; }
  Vector<S>& get_str() { return *ps[0]; }
  bool have_src_str() const { return (bool)ps[1]; }
  const Vector<S>& get_src_str() const { return *ps[1]; }
  Vector<S>& get_src_str() { return *ps[1]; }
  const Vector<S>& get_norm_bcs() const { return *bc[1]; }
  const std::shared_ptr<Body>& get_body_ptr() const { return B; }

  void set_vortex_only() { source_str_is_unknown = false; }
  Int num_unknowns_per_panel() const { return source_str_is_unknown ? 2 : 1; }
  void set_first_row(Int i) { istart = i; }
  Int get_first_row() const { return istart; }
  Int get_num_rows() const { return Int(np * num_unknowns_per_panel() + (is_augmented() ? 1 : 0)); }
  Int get_next_row() const { return istart + get_num_rows(); }

  // src/Surfaces.h:304-325
  bool is_augmented() const {
    bool aug = true;
    if (B) {
      if (B->get_name() == "ground" && vol < S(0)) aug = false;
    } else {
      aug = false;
    }
    if (E != reactive) aug = false;
    return aug;
  }

  // src/Surfaces.h:1238-1379: one .vtu file (panel_NN_FFFFF.vtu) of the nodes, the panels as
  // line cells and, unless inert, the vortex and source sheet strengths per panel
  std::string write_vtk(size_t index, size_t frameno, double time,
                        const std::string& dir = "") const {
    const std::string name = vtu_name("panel_", index, frameno);
    VtuFile f(dir.empty() ? name : dir + "/" + name);
    vtu_time(f, time);
    f.open("Piece", {{"NumberOfPoints", std::to_string(n)}, {"NumberOfCells", std::to_string(np)}});
    f.open("Points");
    f.data_array({{"NumberOfComponents", "3"}, {"Name", "position"}, {"type", "Float32"}},
                 xyz_float(x[0], x[1]));
    f.close();
    f.open("Cells");
    std::vector<int32_t> conn(idx.begin(), idx.end()), off(np);
    for (size_t i = 0; i < np; ++i) off[i] = (int32_t)(2 * (i + 1));
    f.data_array({{"Name", "connectivity"}, {"type", "Int32"}}, conn);
    f.data_array({{"Name", "offsets"}, {"type", "Int32"}}, off);
    f.data_array({{"Name", "types"}, {"type", "UInt8"}}, std::vector<uint8_t>(np, 3));
    f.close();
    if (E != inert) {
      f.open("CellData", {{"Scalars", "vortex sheet strength,source sheet strength"}});
      f.data_array({{"Name", "vortex sheet strength"}, {"type", "Float32"}}, as_float(*ps[0]));
      f.data_array({{"Name", "source sheet strength"}, {"type", "Float32"}}, as_float(*ps[1]));
      f.close();
    }
    f.finish();
    return name;
  }

  // src/Surfaces.h:266-301: pop the solved rotation rate, de-interleave [vortex, source]
  void set_str(Vector<S> in) {
    if (is_augmented()) {
      solved_omega = in.back();
      omega_error = solved_omega - S(B->get_rotvel());
      in.pop_back();
    }
    assert(in.size() == np * num_unknowns_per_panel());
    if (source_str_is_unknown) {
      for (size_t i = 0; i < np; ++i) {
        (*ps[0])[i] = in[2 * i];
        (*ps[1])[i] = in[2 * i + 1];
      }
    } else {
      *ps[0] = in;
    }
  }

  void zero_strengths() {
    for (auto& p : ps)
      if (p) std::fill(p->begin(), p->end(), S(0));
  }

  // src/Surfaces.h:570-620, on the untransformed (here: unmoved) coordinates
  void add_rot_strengths_base(S rotvel) {
    for (size_t i = 0; i < np; ++i) {
      const size_t j = idx[2 * i], jp1 = idx[2 * i + 1];
      const S dx = S(0.5) * (ux[0][j] + ux[0][jp1]) - utc[0];
      const S dy = S(0.5) * (ux[1][j] + ux[1][jp1]) - utc[1];
      const S ui = -rotvel * dy;
      const S vi = rotvel * dx;
      S px = ux[0][jp1] - ux[0][j];
      S py = ux[1][jp1] - ux[1][j];
      const S oo = S(1.0) / area[i];
      px *= oo;
      py *= oo;
      (*ps[0])[i] += S(-1.0) * (ui * px + vi * py);
      (*ps[1])[i] += S(-1.0) * (ui * py - vi * px);
    }
  }
  void add_unit_rot_strengths() { add_rot_strengths_base(S(1)); }
  void add_solved_rot_strengths(S factor) {
    if (is_augmented()) add_rot_strengths_base(factor * solved_omega);
    else if (B) add_rot_strengths_base(factor * S(B->get_rotvel()));
  }

  void zero_vels() {
    for (size_t d = 0; d < Dimensions; ++d) std::fill(pu[d].begin(), pu[d].end(), S(0));
  }

  // src/ElementBase.h:250-273 and src/Surfaces.h:728-752: a body-bound surface follows its body
  // (nodes rotated about the body origin and translated, from the untransformed nodes); the
  // bases are recomputed and the geometric centre is carried along
  void transform(double time) {
    const bool moves = B && M == bodybound;
    S st = S(0), ct = S(1);
    std::array<double, Dimensions> p = {0.0, 0.0};
    if (moves) {
      B->transform(time);
      p = B->get_pos();
      const double th = B->get_orient();
      st = S(std::sin(th));
      ct = S(std::cos(th));
      for (size_t i = 0; i < n; ++i) {
        x[0][i] = S(p[0]) + ux[0][i] * ct - ux[1][i] * st;
        x[1][i] = S(p[1]) + ux[0][i] * st + ux[1][i] * ct;
      }
    }
    compute_bases();
    if (moves) {
      tc[0] = S(p[0]) + utc[0] * ct - utc[1] * st;
      tc[1] = S(p[1]) + utc[0] * st + utc[1] * ct;
    } else {
      tc = utc;
    }
  }

  // src/Surfaces.h:492-523: factor times the body's velocity at each panel centre (translation
  // plus rotation about the transformed geometric centre) added to the panel velocities; the
  // reference rounds the body's velocities to float here
  void add_body_motion(S factor, double time) {
    if (!B || B->get_name() == "ground") return;
    assert(vol > S(0) && "a moving body needs a positive area");
    const std::array<double, Dimensions> v = B->get_vel(time);
    const double w = B->get_rotvel(time);
    for (size_t i = 0; i < np; ++i) {
      for (size_t d = 0; d < Dimensions; ++d) pu[d][i] += factor * S(float(v[d]));
      const Int id0 = idx[2 * i], id1 = idx[2 * i + 1];
      const S xc = S(0.5) * (x[0][id1] + x[0][id0]);
      const S yc = S(0.5) * (x[1][id1] + x[1][id0]);
      pu[0][i] -= factor * S(float(w)) * (yc - tc[1]);
      pu[1][i] += factor * S(float(w)) * (xc - tc[0]);
    }
  }

  // src/Surfaces.h:764-774
  void finalize_vels(const std::array<double, Dimensions>& fs) {
    const double factor = 0.5 / M_PI;
    for (size_t d = 0; d < Dimensions; ++d)
      for (size_t i = 0; i < np; ++i) pu[d][i] = S(fs[d] + double(pu[d][i]) * factor);
  }

  // src/Surfaces.h:830-878: points just off the panels, rounded to float as the reference
  // stores them; strengths (sheet strength + vortex BC) times panel length
  void represent_as_particles(S offset, S ips, std::vector<S>& px, std::vector<S>& py,
                              std::vector<S>& pstr) const {
    px.clear();
    py.clear();
    pstr.clear();
    const S thisips = ips < S(0) ? area[0] : ips;
    for (size_t i = 0; i < np; ++i) {
      const Int id0 = idx[2 * i], id1 = idx[2 * i + 1];
      float this_str = float((*ps[0])[i]);
      if (E == reactive) this_str = float(S(this_str) + (*bc[0])[i]);
      this_str = float(S(this_str) * area[i]);
      const size_t nshed = 1 + size_t(std::max(0, int(area[i] / (S(1.2) * thisips))));
      for (size_t j = 0; j < nshed; ++j) {
        const S wgt = S(1 + 2 * j) / S(2 * nshed);
        px.push_back(S(float(wgt * x[0][id1] + (S(1.0) - wgt) * x[0][id0] + offset * b[1][0][i])));
        py.push_back(S(float(wgt * x[1][id1] + (S(1.0) - wgt) * x[1][id0] + offset * b[1][1][i])));
        pstr.push_back(S(this_str / float(nshed)));
      }
    }
  }

  S get_last_body_circ() const { return S(2.0) * vol * this_omega; }
  S get_last_body_circ_error() const { return S(2.0 * double(vol) * double(omega_error)); }
  S get_reabsorbed() const { return reabsorbed_gamma; }
  // src/Surfaces.h:1053-1055: circulation taken from the particles by the clearing pass
  void add_to_reabsorbed(S circ) { reabsorbed_gamma += circ; }
  S get_solved_omega() const { return solved_omega; }

  std::string to_string() const {
    std::string m = E == active ? " Active" : (E == reactive ? " Reactive" : " Inert");
    m += M == lagrangian ? " Lagrangian" : (M == bodybound ? " Body-fixed" : " Fixed");
    return m + " Surfaces";
  }

 private:
  // src/Surfaces.h:682-725
  void compute_bases() {
    for (size_t i = 0; i < Dimensions; ++i)
      for (size_t j = 0; j < Dimensions; ++j) b[i][j].assign(np, S(0));
    area.assign(np, S(0));
    for (size_t i = 0; i < np; ++i) {
      const size_t a = idx[2 * i], c = idx[2 * i + 1];
      S t0 = x[0][c] - x[0][a], t1 = x[1][c] - x[1][a];
      const S base = std::sqrt(t0 * t0 + t1 * t1);
      t0 *= (S(1.0) / base);
      t1 *= (S(1.0) / base);
      area[i] = base;
      b[0][0][i] = t0;
      b[0][1][i] = t1;
      b[1][0][i] = -t1;
      b[1][1][i] = t0;
    }
  }

  // src/Surfaces.h:623-677: signed area and centroid from the fan of triangles at the origin
  void set_geom_center() {
    double asum = 0.0, xsum = 0.0, ysum = 0.0;
    for (size_t i = 0; i < np; ++i) {
      const size_t j = idx[2 * i], jp1 = idx[2 * i + 1];
      const double x0 = ux[0][j], y0 = ux[1][j], x1 = ux[0][jp1], y1 = ux[1][jp1];
      const double xc = (0.0 + x0 + x1) / 3.0, yc = (0.0 + y0 + y1) / 3.0;
      const double px = x1 - x0, py = y1 - y0;
      const double a = std::sqrt(std::pow(x0, 2) + std::pow(y0, 2));
      const double bb = std::sqrt(std::pow(px, 2) + std::pow(py, 2));
      const double c = std::sqrt(std::pow(x1, 2) + std::pow(y1, 2));
      const double hs = 0.5 * (a + bb + c);
      const double asq = hs * (hs - a) * (hs - bb) * (hs - c);
      double ar = asq > 0.0 ? std::sqrt(asq) : 0.0;
      if (y0 * px - x0 * py < 0.0) ar = -ar;
      asum += ar;
      xsum += xc * ar;
      ysum += yc * ar;
    }
    vol = S(asum);
    utc[0] = S(xsum / asum);
    utc[1] = S(ysum / asum);
  }

  elem_t E;
  move_t M;
  std::shared_ptr<Body> B;
  size_t n, np;
  std::array<Vector<S>, Dimensions> x, ux;
  std::vector<Int> idx;
  std::array<std::array<Vector<S>, Dimensions>, Dimensions> b;
  Vector<S> area;
  std::array<std::optional<Vector<S>>, 2> ps, bc;
  std::array<Vector<S>, Dimensions> pu;
  bool source_str_is_unknown = true;
  Int istart = 0;
  S vol = S(-1.0);
  std::array<S, Dimensions> utc = {S(0), S(0)}, tc = {S(0), S(0)};
  S solved_omega = S(0), omega_error = S(0), this_omega = S(0), reabsorbed_gamma = S(0);
};

}  // namespace omega2d
