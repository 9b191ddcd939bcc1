// Influence.h - the *_affect_* operators of the gpu_hip execution environment: host mirror of
// src/Influence.h:56-901 (same signatures, accumulate semantics and per-call log line), each
// forwarding to the C ABI in include/omega2d_hip.h. Only the gpu_hip accelerator is
// implemented here; the CPU environments stay in the reference.
#pragma once

#include <chrono>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <variant>
#include <vector>

#include "Elements.h"
#include "ExecEnv.h"
#include "omega2d_hip.h"

namespace omega2d {

using Clock = std::chrono::steady_clock;

inline void o2d_check(int rc, const char* what) {
  if (rc != O2D_OK)
    throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) +
                             "): " + o2d_hip_last_error_string());
}

inline void require_gpu_hip(const ExecEnv& env) {
  if (env.get_instrs() != gpu_hip || !env.is_internal())
    throw std::runtime_error("this build implements the gpu_hip environment only (got" +
                             env.to_string() + ")");
}

inline void log_rate(const char* name, Clock::time_point start, double flops) {
  const double s = std::chrono::duration<double>(Clock::now() - start).count();
  std::printf("    %s: [%.4f] seconds at %.3f GFlop/s\n", name, (float)s,
              (float)(1.e-9 * flops / s));
}

template <class S>
std::vector<double> as_double(const std::vector<S>& v) {
  return std::vector<double>(v.begin(), v.end());
}

// A read-only fp64 view of a storage array for the _d entry points: the array itself when the
// storage is double, else a converted copy.
template <class S>
class DoubleView {
 public:
  explicit DoubleView(const std::vector<S>& v) {
    if constexpr (std::is_same_v<S, double>) {
      p = v.data();
    } else {
      copy.assign(v.begin(), v.end());
      p = copy.data();
    }
  }
  const double* data() const { return p; }

 private:
  std::vector<double> copy;
  const double* p = nullptr;
};

// An fp64 in/out array for the _d entry points that update in place: the storage array itself
// when it is double, else a converted copy that commit() rounds back into the storage.
template <class S>
class DoubleInOut {
 public:
  explicit DoubleInOut(std::vector<S>& v) : v(v) {
    if constexpr (!std::is_same_v<S, double>) copy.assign(v.begin(), v.end());
  }
  double* data() {
    if constexpr (std::is_same_v<S, double>) return v.data();
    else return copy.data();
  }
  void commit() {
    if constexpr (!std::is_same_v<S, double>)
      for (size_t i = 0; i < v.size(); ++i) v[i] = S(copy[i]);
  }

 private:
  std::vector<S>& v;
  std::vector<double> copy;
};

// Accumulating fp64 outputs for the _d entry points: the storage arrays themselves when they
// are double (the entry points add into them), else zeroed fp64 arrays that commit() adds into
// the storage, each sum rounded once to S.
template <class S>
class DoubleAccum {
 public:
  DoubleAccum(std::vector<S>& u, std::vector<S>& v, size_t n) : u(u), v(v), n(n) {
    if constexpr (!std::is_same_v<S, double>) {
      tu.assign(n, 0.0);
      tv.assign(n, 0.0);
    }
  }
  double* du() {
    if constexpr (std::is_same_v<S, double>) return u.data();
    else return tu.data();
  }
  double* dv() {
    if constexpr (std::is_same_v<S, double>) return v.data();
    else return tv.data();
  }
  void commit() {
    if constexpr (!std::is_same_v<S, double>) {
      for (size_t i = 0; i < n; ++i) {
        u[i] = S(double(u[i]) + tu[i]);
        v[i] = S(double(v[i]) + tv[i]);
      }
    }
  }

 private:
  std::vector<S>& u;
  std::vector<S>& v;
  size_t n;
  std::vector<double> tu, tv;
};

// ---------------------------------------------------------------------- points on points

template <class S, class A>
void points_affect_points(const Points<S>& src, Points<S>& targ, const ResultsType& restype,
                          const ExecEnv& env) {
  require_gpu_hip(env);
  if (restype.compute_psi() || restype.compute_grad())
    throw std::runtime_error("points cannot compute streamfunction or velocity gradients");
  const auto start = Clock::now();
  const o2d_env e = env.c_env(restype.get_type());
  const auto& sx = src.get_pos();
  const auto& tx = targ.get_pos();
  auto& tu = targ.get_vel();
  S* tw = restype.compute_vort() ? targ.get_vort().data() : nullptr;
  S* te = restype.compute_shear() ? targ.get_shear().data() : nullptr;
  const S* tr = targ.is_inert() ? nullptr : targ.get_rad().data();
  if (env.get_summation() == barneshut && restype.get_type() == velonly) {
    // the treecode (fp64, double arrays), summation_t barneshut; vorticity and shear outputs
    // stay on the direct sum below
    const DoubleView<S> sxd(sx[0]), syd(sx[1]), ssd(src.get_str()), srd(src.get_rad());
    const DoubleView<S> txd(tx[0]), tyd(tx[1]), trd(targ.get_rad());
    DoubleAccum<S> acc(tu[0], tu[1], targ.get_n());
    o2d_check(o2d_hip_points_on_points_tree_d(&e, &env.get_tree(), (int64_t)src.get_n(),
                                              sxd.data(), syd.data(), ssd.data(), srd.data(),
                                              (int64_t)targ.get_n(), txd.data(), tyd.data(),
                                              tr ? trd.data() : nullptr, acc.du(), acc.dv()),
              "o2d_hip_points_on_points_tree");
    acc.commit();
    // the log reports the rate of the direct sum the treecode stands in for
    const int per = o2d_hip_flops_per_pair(env.get_core(), targ.is_inert() ? 1 : 0, 0);
    log_rate("points_affect_points", start,
             (double)targ.get_n() * (2.0 + per * (double)src.get_n()));
    return;
  }
  int rc;
  if constexpr (std::is_same_v<S, float>) {
    rc = o2d_hip_points_on_points_f(&e, (int64_t)src.get_n(), sx[0].data(), sx[1].data(),
                                    src.get_str().data(), src.get_rad().data(),
                                    (int64_t)targ.get_n(), tx[0].data(), tx[1].data(), tr,
                                    tu[0].data(), tu[1].data(), tw, te);
  } else {
    static_assert(std::is_same_v<S, double>, "storage type must be float or double");
    rc = o2d_hip_points_on_points_d(&e, (int64_t)src.get_n(), sx[0].data(), sx[1].data(),
                                    src.get_str().data(), src.get_rad().data(),
                                    (int64_t)targ.get_n(), tx[0].data(), tx[1].data(), tr,
                                    tu[0].data(), tu[1].data(), tw, te);
  }
  o2d_check(rc, "o2d_hip_points_on_points");
  const int per = o2d_hip_flops_per_pair(env.get_core(), targ.is_inert() ? 1 : 0, 0);
  log_rate("points_affect_points", start,
           (double)targ.get_n() * (2.0 + per * (double)src.get_n()));
}

// ---------------------------------------------------------------------- panels on points

template <class S, class A>
void panels_affect_points(const Surfaces<S>& src, Points<S>& targ, const ResultsType& restype,
                          const ExecEnv& env) {
  require_gpu_hip(env);
  // src/Influence.h:427-428: only streamfunction and gradients are refused; for the other
  // result types panels add velocity and leave vorticity and shear untouched
  if (restype.compute_psi() || restype.compute_grad())
    throw std::runtime_error("surfaces cannot compute streamfunction or velocity gradients");
  const auto start = Clock::now();
  const o2d_env e = env.c_env(velonly);
  const DoubleView<S> nx(src.get_pos()[0]), ny(src.get_pos()[1]);
  const DoubleView<S> vs(src.get_str()), ss(src.get_src_str());
  const DoubleView<S> tx(targ.get_pos()[0]), ty(targ.get_pos()[1]);
  DoubleAccum<S> acc(targ.get_vel()[0], targ.get_vel()[1], targ.get_n());
  o2d_check(o2d_hip_panels_on_points_d(&e, (int64_t)src.get_n(), nx.data(), ny.data(),
                                       (int64_t)src.get_npanels(), src.get_idx().data(),
                                       vs.data(), src.have_src_str() ? ss.data() : nullptr,
                                       (int64_t)targ.get_n(), tx.data(), ty.data(), acc.du(),
                                       acc.dv()),
            "o2d_hip_panels_on_points");
  acc.commit();
  log_rate("panels_affect_points", start,
           (double)targ.get_n() *
               (2.0 + o2d_hip_flops_per_pair(env.get_core(), 1, src.have_src_str() ? 2 : 1) *
                          (double)src.get_npanels()));
}

// ---------------------------------------------------------------------- points on panels

template <class S, class A>
void points_affect_panels(const Points<S>& src, Surfaces<S>& targ, const ResultsType& restype,
                          const ExecEnv& env) {
  require_gpu_hip(env);
  // src/Influence.h:638-639, as for panels_affect_points: velocity for every other result type
  if (restype.compute_psi() || restype.compute_grad())
    throw std::runtime_error("points cannot compute streamfunction or velocity gradients");
  const auto start = Clock::now();
  const o2d_env e = env.c_env(velonly);
  const DoubleView<S> sx(src.get_pos()[0]), sy(src.get_pos()[1]), ss(src.get_str());
  const DoubleView<S> nx(targ.get_pos()[0]), ny(targ.get_pos()[1]);
  DoubleAccum<S> acc(targ.get_vel()[0], targ.get_vel()[1], targ.get_npanels());
  o2d_check(o2d_hip_points_on_panels_d(&e, (int64_t)src.get_n(), sx.data(), sy.data(), ss.data(),
                                       (int64_t)targ.get_n(), nx.data(), ny.data(),
                                       (int64_t)targ.get_npanels(), targ.get_idx().data(),
                                       acc.du(), acc.dv()),
            "o2d_hip_points_on_panels");
  acc.commit();
  log_rate("points_affect_panels", start,
           (double)targ.get_npanels() * (11.0 + 39.0 * (double)src.get_n()));
}

// ---------------------------------------------------------------------- panels on panels

// src/Influence.h:774-795: panels_affect_points at points just off the target panels, added
// (first npanels values) into the target panel velocities
template <class S, class A>
void panels_affect_panels(const Surfaces<S>& src, Surfaces<S>& targ, const ResultsType& restype,
                          const ExecEnv& env) {
  std::vector<S> px, py, pstr;
  targ.represent_as_particles(S(0.0001), S(-1.0), px, py, pstr);
  std::vector<S> rad(px.size(), S(0.0001));
  Points<S> pts(px.size(), px.data(), py.data(), pstr.data(), rad.data(), active, lagrangian);
  panels_affect_points<S, A>(src, pts, restype, env);
  auto& tov = targ.get_vel();
  const auto& fromv = pts.get_vel();
  for (size_t d = 0; d < Dimensions; ++d)
    for (size_t i = 0; i < targ.get_npanels() && i < pts.get_n(); ++i) tov[d][i] += fromv[d][i];
}

// ---------------------------------------------------------------------- dispatch

template <class S>
using Collection = std::variant<Points<S>, Surfaces<S>>;

template <class S, class A>
struct InfluenceVisitor {
  void operator()(const Points<S>& s, Points<S>& t) { points_affect_points<S, A>(s, t, restype, env); }
  void operator()(const Surfaces<S>& s, Points<S>& t) { panels_affect_points<S, A>(s, t, restype, env); }
  void operator()(const Points<S>& s, Surfaces<S>& t) { points_affect_panels<S, A>(s, t, restype, env); }
  void operator()(const Surfaces<S>& s, Surfaces<S>& t) { panels_affect_panels<S, A>(s, t, restype, env); }
  ResultsType restype;
  ExecEnv env;
};

}  // namespace omega2d
