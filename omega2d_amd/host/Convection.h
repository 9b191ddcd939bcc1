// Convection.h - velocity evaluation and Runge-Kutta advection of the gpu_hip environment: host
// mirror of src/Convection.h:172-598 (find_vels, find_derivs, advect with Euler, Ralston RK2
// and Ralston RK3). Positions advance in fp64 on the device (o2d_hip_rk_move_d); after every
// move the Lagrangian points are pushed out of the bodies' inner layer (clear_inner_layer,
// method 1, src/Reflect.h) exactly where the reference does it.
#pragma once

#include <array>
#include <stdexcept>
#include <variant>
#include <vector>

#include "BEM.h"
#include "InfluenceVort.h"
#include "Reflect.h"

namespace omega2d {

// x += dt * sum_k w[k] * u_k (src/ElementBase.h:276-359), stage velocities taken from `stages`
// (stages[0] is p itself). With two or three stages the reference also leaves the combined
// velocity u = sum_k w[k] u_k in the element (src/ElementBase.h:308-312, 340-345); so does this.
// Positions advance in fp64 on the device and are rounded to S once.
template <class S>
void move_points(Points<S>& p, double dt, const std::vector<double>& w,
                 const std::vector<const Points<S>*>& stages, int32_t device) {
  if (p.get_movet() != lagrangian) return;
  const size_t n = p.get_n();
  const size_t ns = stages.size();
  std::vector<DoubleView<S>> us, vs;
  us.reserve(ns);
  vs.reserve(ns);
  std::vector<const double*> up, vp;
  for (const Points<S>* q : stages) {
    us.emplace_back(q->get_vel()[0]);
    vs.emplace_back(q->get_vel()[1]);
    up.push_back(us.back().data());
    vp.push_back(vs.back().data());
  }
  o2d_env e = {O2D_CORE_V2, O2D_PLAIN_FP64, O2D_VELONLY, device, nullptr};
  if constexpr (std::is_same_v<S, double>) {
    o2d_check(o2d_hip_rk_move_d(&e, (int64_t)n, dt, (int32_t)ns, w.data(), up.data(), vp.data(),
                                p.get_pos()[0].data(), p.get_pos()[1].data()),
              "o2d_hip_rk_move");
  } else {
    auto x = as_double(p.get_pos()[0]), y = as_double(p.get_pos()[1]);
    o2d_check(o2d_hip_rk_move_d(&e, (int64_t)n, dt, (int32_t)ns, w.data(), up.data(), vp.data(),
                                x.data(), y.data()),
              "o2d_hip_rk_move");
    for (size_t i = 0; i < n; ++i) {
      p.get_pos()[0][i] = S(x[i]);
      p.get_pos()[1][i] = S(y[i]);
    }
  }
  if (ns >= 2) {
    // the combination the device used for the positions, in the same order of operations;
    // stages[0] may be p, so every stage value is read before p's velocity is overwritten
    auto& pu = p.get_vel();
    for (size_t i = 0; i < n; ++i) {
      double cu = w[0] * up[0][i] + w[1] * up[1][i];
      double cv = w[0] * vp[0][i] + w[1] * vp[1][i];
      if (ns == 3) {
        cu = cu + w[2] * up[2][i];
        cv = cv + w[2] * vp[2][i];
      }
      pu[0][i] = S(cu);
      pu[1][i] = S(cv);
    }
  }
}

template <class S, class A, class I>
class Convection {
 public:
  explicit Convection(const ExecEnv& env, int order = 2) : env(env), order(order) {}

  // src/Convection.h:172-223
  void find_vels(const std::array<double, Dimensions>& fs, std::vector<Collection<S>>& vort,
                 std::vector<Collection<S>>& bdry, std::vector<Collection<S>>& targets,
                 results_t results = velonly, bool force = false) {
    InfluenceVisitor<S, A> visitor = {ResultsType(results), env};
    for (auto& b : bdry)
      if (auto* s = std::get_if<Surfaces<S>>(&b)) s->add_solved_rot_strengths(S(1.0));
    for (auto& t : targets) {
      const move_t mt = std::visit([](auto& el) { return el.get_movet(); }, t);
      if (!(force || mt == lagrangian)) continue;
      std::visit([](auto& el) { el.zero_vels(); }, t);
      for (auto& s : vort) std::visit(visitor, std::as_const(s), t);
      for (auto& s : bdry) std::visit(visitor, std::as_const(s), t);
      std::visit([&](auto& el) { el.finalize_vels(fs); }, t);
    }
    for (auto& b : bdry)
      if (auto* s = std::get_if<Surfaces<S>>(&b)) s->add_solved_rot_strengths(S(-1.0));
  }

  // src/Convection.h:134-166: vorticity (and zeroed velocities) on every Points target from
  // the Points among the particles and the boundaries, finalized without a freestream
  void find_vort(std::vector<Collection<S>>& vort, std::vector<Collection<S>>& bdry,
                 std::vector<Collection<S>>& targets) {
    for (auto& t : targets) {
      auto* pt = std::get_if<Points<S>>(&t);
      if (!pt) continue;
      pt->zero_vels();
      auto& w = pt->get_vort();
      std::fill(w.begin(), w.end(), S(0));
      for (auto* list : {&vort, &bdry})
        for (auto& s : *list)
          if (auto* ps = std::get_if<Points<S>>(&s))
            points_affect_points_vorticity<S, A>(*ps, *pt, env);
      pt->finalize_vels({0.0, 0.0});
    }
  }

  // src/Convection.h:229-243
  void find_derivs(double time, const std::array<double, Dimensions>& fs, BEM<S, I>& bem,
                   std::vector<Collection<S>>& bdry, std::vector<Collection<S>>& vort,
                   std::vector<Collection<S>>& fldpt) {
    solve_bem<S, A, I>(time, fs, vort, bdry, bem, env);
    find_vels(fs, vort, bdry, vort);
    find_vels(fs, vort, bdry, fldpt);
  }

  // src/Convection.h:249-268; ips is the nominal particle spacing that sets the inner layer
  void advect(double time, double dt, const std::array<double, Dimensions>& fs, S ips,
              std::vector<Collection<S>>& vort, std::vector<Collection<S>>& bdry,
              std::vector<Collection<S>>& fldpt, BEM<S, I>& bem) {
    if (order == 1) advect_1st(time, dt, fs, ips, vort, bdry, fldpt, bem);
    else if (order == 2) advect_2nd_ralston(time, dt, fs, ips, vort, bdry, fldpt, bem);
    else if (order == 3) advect_3rd(time, dt, fs, ips, vort, bdry, fldpt, bem);
    else throw std::runtime_error("convection order must be 1, 2 or 3");
  }

  // src/Convection.h:274-303
  void advect_1st(double time, double dt, const std::array<double, Dimensions>& fs, S ips,
                  std::vector<Collection<S>>& vort, std::vector<Collection<S>>& bdry,
                  std::vector<Collection<S>>& fldpt, BEM<S, I>& bem) {
    find_derivs(time, fs, bem, bdry, vort, fldpt);
    for (auto* list : {&vort, &fldpt})
      for (auto& c : *list)
        if (auto* p = std::get_if<Points<S>>(&c)) move_points<S>(*p, dt, {1.0}, {p}, dev());
    clear(bdry, vort, ips);
    clear(bdry, fldpt, ips);
  }

  // src/Convection.h:390-466: stage at t + 2dt/3, u = u1/4 + 3 u2/4
  void advect_2nd_ralston(double time, double dt, const std::array<double, Dimensions>& fs,
                          S ips, std::vector<Collection<S>>& vort,
                          std::vector<Collection<S>>& bdry, std::vector<Collection<S>>& fldpt,
                          BEM<S, I>& bem) {
    const double twothirds = 2.0 / 3.0;
    find_derivs(time, fs, bem, bdry, vort, fldpt);
    std::vector<Collection<S>> iv = vort, ifl = fldpt;
    for (auto* list : {&iv, &ifl}) {
      for (auto& c : *list)
        if (auto* p = std::get_if<Points<S>>(&c)) move_points<S>(*p, twothirds * dt, {1.0}, {p}, dev());
      clear(bdry, *list, ips);
    }
    find_derivs(time + twothirds * dt, fs, bem, bdry, iv, ifl);
    combine(vort, {&iv}, dt, {0.25, 0.75});
    combine(fldpt, {&ifl}, dt, {0.25, 0.75});
    clear(bdry, vort, ips);
    clear(bdry, fldpt, ips);
  }

  // src/Convection.h:472-598: stages at t, t + dt/2, t + 3dt/4 (the third from x0 with u1),
  // weights 2/9, 3/9, 4/9
  void advect_3rd(double time, double dt, const std::array<double, Dimensions>& fs, S ips,
                  std::vector<Collection<S>>& vort, std::vector<Collection<S>>& bdry,
                  std::vector<Collection<S>>& fldpt, BEM<S, I>& bem) {
    find_derivs(time, fs, bem, bdry, vort, fldpt);
    std::vector<Collection<S>> v1 = vort, f1 = fldpt;
    for (auto* list : {&v1, &f1}) {
      for (auto& c : *list)
        if (auto* p = std::get_if<Points<S>>(&c)) move_points<S>(*p, 0.5 * dt, {1.0}, {p}, dev());
      clear(bdry, *list, ips);
    }
    find_derivs(time + 0.5 * dt, fs, bem, bdry, v1, f1);
    std::vector<Collection<S>> v2 = vort, f2 = fldpt;
    for (auto pr : {std::make_pair(&v2, &v1), std::make_pair(&f2, &f1)}) {
      for (size_t i = 0; i < pr.first->size(); ++i) {
        auto* p2 = std::get_if<Points<S>>(&(*pr.first)[i]);
        auto* p1 = std::get_if<Points<S>>(&(*pr.second)[i]);
        if (p2 && p1) move_points<S>(*p2, 0.75 * dt, {1.0}, {p1}, dev());
      }
      clear(bdry, *pr.first, ips);
    }
    find_derivs(time + 0.75 * dt, fs, bem, bdry, v2, f2);
    combine(vort, {&v1, &v2}, dt, {2.0 / 9.0, 3.0 / 9.0, 4.0 / 9.0});
    combine(fldpt, {&f1, &f2}, dt, {2.0 / 9.0, 3.0 / 9.0, 4.0 / 9.0});
    clear(bdry, vort, ips);
    clear(bdry, fldpt, ips);
  }

 private:
  int32_t dev() const { return env.get_device(); }

  // the reference's push after every move: method 1, layer 0.5/sqrt(2 pi) spacings thick
  void clear(std::vector<Collection<S>>& bdry, std::vector<Collection<S>>& pts, S ips) {
    clear_inner_layer<S>(1, bdry, pts, S(0.5 / std::sqrt(2.0 * M_PI)), ips, env);
  }

  void combine(std::vector<Collection<S>>& base, std::vector<std::vector<Collection<S>>*> more,
               double dt, const std::vector<double>& w) {
    for (size_t i = 0; i < base.size(); ++i) {
      auto* p0 = std::get_if<Points<S>>(&base[i]);
      if (!p0) continue;
      std::vector<const Points<S>*> st = {p0};
      for (auto* m : more) st.push_back(std::get_if<Points<S>>(&(*m)[i]));
      move_points<S>(*p0, dt, w, st, dev());
    }
  }

  ExecEnv env;
  int order;
};

}  // namespace omega2d
