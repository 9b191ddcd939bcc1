// Reflect.h - particle/panel proximity operations of the gpu_hip environment: host mirror of
// src/Reflect.h (reflect_panp2 :167-311, reflect_interior :317-338, clear_inner_panp2
// :487-727, clear_inner_layer :733-762, get_nearest_distances :768-826), same signatures and
// per-call log lines, each forwarding to the C ABI in include/omega2d_hip.h. The geometry is
// computed in fp64 on the device whatever the storage type S.
#pragma once

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <variant>
#include <vector>

#include "Influence.h"

namespace omega2d {

namespace detail {

template <class S>
struct PanelArgs {
  DoubleView<S> nx, ny, pnx, pny;
  const std::vector<Int>* idx;
  explicit PanelArgs(const Surfaces<S>& s)
      : nx(s.get_pos()[0]), ny(s.get_pos()[1]), pnx(s.get_norm()[0]), pny(s.get_norm()[1]),
        idx(&s.get_idx()) {}
};

template <class S>
void store(std::vector<S>& dst, const std::vector<double>& src) {
  for (size_t i = 0; i < dst.size(); ++i) dst[i] = S(src[i]);
}

}  // namespace detail

// src/Reflect.h:167-311: move targets that lie behind their nearest element to the front
template <class S>
void reflect_panp2(const Surfaces<S>& src, Points<S>& targ, const ExecEnv& env) {
  require_gpu_hip(env);
  const auto start = Clock::now();
  std::printf("  Reflecting %zu particles from near %zu panels\n", targ.get_n(), src.get_npanels());
  const detail::PanelArgs<S> pa(src);
  DoubleInOut<S> x(targ.get_pos()[0]), y(targ.get_pos()[1]);
  const o2d_env e = env.c_env(O2D_VELONLY);
  int64_t nref = 0;
  o2d_check(o2d_hip_reflect_d(&e, (int64_t)src.get_n(), pa.nx.data(), pa.ny.data(),
                              (int64_t)src.get_npanels(), pa.idx->data(), pa.pnx.data(),
                              pa.pny.data(), (int64_t)targ.get_n(), x.data(), y.data(), &nref),
            "o2d_hip_reflect");
  x.commit();
  y.commit();
  std::printf("    reflected %lld particles\n", (long long)nref);
  log_rate("reflect_panp2", start,
           (double)targ.get_n() * (62.0 + 27.0 * (double)src.get_npanels()));
}

// src/Reflect.h:317-338
template <class S>
void reflect_interior(const std::vector<Collection<S>>& bdry, std::vector<Collection<S>>& vort,
                      const ExecEnv& env) {
  for (auto& t : vort)
    if (auto* pts = std::get_if<Points<S>>(&t))
      for (auto& s : bdry)
        if (auto* surf = std::get_if<Surfaces<S>>(&s)) reflect_panp2<S>(*surf, *pts, env);
}

// src/Reflect.h:487-727: returns the circulation removed (method 0); method 1 pushes targets
// out to the layer cutoff_mult * ips above the nearest element
template <class S>
S clear_inner_panp2(int method, const Surfaces<S>& src, Points<S>& targ, S cutoff_mult, S ips,
                    const ExecEnv& env) {
  require_gpu_hip(env);
  const auto start = Clock::now();
  std::printf("  Clearing %zu particles from near %zu panels\n", targ.get_n(), src.get_npanels());
  const detail::PanelArgs<S> pa(src);
  DoubleInOut<S> x(targ.get_pos()[0]), y(targ.get_pos()[1]);
  const bool fldpts = targ.is_inert();
  // field points carry no strength or radius (their arrays are empty)
  DoubleInOut<S> s(targ.get_str());
  const DoubleView<S> r(targ.get_rad());
  const o2d_env e = env.c_env(O2D_VELONLY);
  double circ = 0.0;
  int64_t ncrop = 0;
  o2d_check(o2d_hip_clear_inner_d(&e, method, (double)cutoff_mult, (double)ips,
                                  (int64_t)src.get_n(), pa.nx.data(), pa.ny.data(),
                                  (int64_t)src.get_npanels(), pa.idx->data(), pa.pnx.data(),
                                  pa.pny.data(), (int64_t)targ.get_n(), x.data(), y.data(),
                                  fldpts ? nullptr : s.data(), fldpts ? nullptr : r.data(),
                                  &circ, &ncrop),
            "o2d_hip_clear_inner");
  x.commit();
  y.commit();
  if (!fldpts && method == 0) s.commit();
  std::printf("    %s %lld particles\n", method == 0 ? "cropped" : "pushed", (long long)ncrop);
  log_rate("clear_inner_panp2", start,
           (double)(targ.get_n() + ncrop) * (62.0 + 27.0 * (double)src.get_npanels()));
  return S(circ);
}

// src/Reflect.h:733-762: only Lagrangian points are cleared; the boundary keeps account of the
// circulation it reabsorbed
template <class S>
void clear_inner_layer(int method, std::vector<Collection<S>>& bdry,
                       std::vector<Collection<S>>& vort, S cutoff_factor, S ips,
                       const ExecEnv& env) {
  for (auto& t : vort) {
    auto* pts = std::get_if<Points<S>>(&t);
    if (!pts || pts->get_movet() != lagrangian) continue;
    for (auto& b : bdry)
      if (auto* surf = std::get_if<Surfaces<S>>(&b)) {
        const S lost = clear_inner_panp2<S>(method, *surf, *pts, cutoff_factor, ips, env);
        surf->add_to_reabsorbed(lost);
      }
  }
}

// src/Reflect.h:768-826
template <class S>
Vector<S> get_nearest_distances(const Surfaces<S>& src, const Points<S>& targ,
                                const ExecEnv& env) {
  require_gpu_hip(env);
  const auto start = Clock::now();
  const detail::PanelArgs<S> pa(src);
  const DoubleView<S> x(targ.get_pos()[0]), y(targ.get_pos()[1]);
  std::vector<double> d(targ.get_n());
  const o2d_env e = env.c_env(O2D_VELONLY);
  o2d_check(o2d_hip_nearest_distances_d(&e, (int64_t)src.get_n(), pa.nx.data(), pa.ny.data(),
                                        (int64_t)src.get_npanels(), pa.idx->data(),
                                        (int64_t)targ.get_n(), x.data(), y.data(), d.data()),
            "o2d_hip_nearest_distances");
  Vector<S> out(targ.get_n());
  detail::store(out, d);
  log_rate("get_nearest_distances", start,
           (double)targ.get_n() * (1.0 + 27.0 * (double)src.get_npanels()));
  return out;
}

}  // namespace omega2d
