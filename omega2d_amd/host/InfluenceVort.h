// InfluenceVort.h - Gaussian vorticity of particles at points for the gpu_hip environment: host
// mirror of points_affect_points_vorticity (src/InfluenceVort.h:47-126), same signature,
// accumulate semantics and log line, forwarding to o2d_hip_points_vorticity_d (fp64).
#pragma once

#include <vector>

#include "Influence.h"

namespace omega2d {

template <class S, class A>
void points_affect_points_vorticity(const Points<S>& src, Points<S>& targ, const ExecEnv& env) {
  require_gpu_hip(env);
  std::printf("    in ptptvort with%s\n", env.to_string().c_str());
  const auto start = Clock::now();
  const auto sx = as_double(src.get_pos()[0]), sy = as_double(src.get_pos()[1]);
  const auto ss = as_double(src.get_str()), sr = as_double(src.get_rad());
  const auto tx = as_double(targ.get_pos()[0]), ty = as_double(targ.get_pos()[1]);
  auto& w = targ.get_vort();
  std::vector<double> tw(w.begin(), w.end());
  const o2d_env e = env.c_env(O2D_VELONLY);
  o2d_check(o2d_hip_points_vorticity_d(&e, (int64_t)src.get_n(), sx.data(), sy.data(), ss.data(),
                                       sr.data(), (int64_t)targ.get_n(), tx.data(), ty.data(),
                                       tw.data()),
            "o2d_hip_points_vorticity");
  for (size_t i = 0; i < w.size(); ++i) w[i] = S(tw[i]);
  log_rate("points_affect_points_vorticity", start,
           (double)targ.get_n() * (1.0 + 13.0 * (double)src.get_n()));
}

}  // namespace omega2d
