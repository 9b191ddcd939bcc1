// ref_dropin.cpp - the reference's own particle velocity path, built with EXTERNAL_VEL_SOLVE and
// linked against libomega2d_hip.so, so that src/Influence.h:74-98 hands its sums to the HIP
// library's external_vel_solver_f_ / external_vel_solver_tr_f_.
//
// TEST INFRASTRUCTURE ONLY. Compiled by oracle/Makefile against the reference headers where they
// lie (src/Collection.h, src/Influence.h, src/Points.h), never copied. It exists to show the
// drop-in with no change to the reference's source: a non-internal ExecEnv
// (src/ExecEnv.h:46-52) routes points_affect_points to the hook, and the steps around the calls
// are those of Convection::find_vels (src/Convection.h:185-213): zero the targets, dispatch every
// source collection through InfluenceVisitor (src/Influence.h:885-900), then finalize_vels.
// tests/test_gpu_dropin.py calls it and compares with the oracle.

#include "Collection.h"
#include "Influence.h"

#include <array>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

Points<float> dropin_points(int64_t n, const float* x, const float* y, const float* str,
                            const float* rad) {
  std::vector<float> xs(2 * n);
  for (int64_t i = 0; i < n; ++i) { xs[2 * i] = x[i]; xs[2 * i + 1] = y[i]; }
  std::vector<float> vals;
  const bool inert_pts = (rad == nullptr);
  if (!inert_pts) vals.assign(str, str + n);
  ElementPacket<float> pk(xs, std::vector<Int>(), vals, (size_t)n, 0);
  Points<float> p(pk, inert_pts ? inert : active, lagrangian, nullptr, 0.01f);
  if (!inert_pts) {
    Vector<float>& r = p.get_rad();
    for (int64_t i = 0; i < n; ++i) r[i] = rad[i];
  }
  return p;
}

}  // namespace

extern "C" {

// Velocities of nsets source collections at one target collection, the way find_vels computes
// them with an external environment. Set k has ns[k] particles at sx[k], sy[k] with strengths
// ss[k] and radii sr[k]. tr == NULL makes the targets inert field points (external_vel_solver_f_),
// otherwise particles with radii tr and unit strength (external_vel_solver_tr_f_). finalize != 0
// applies finalize_vels with freestream (fsx, fsy); finalize == 0 returns the unscaled sums.
// Returns 0, or -1 on bad arguments.
int dropin_find_vels_f(int nsets, const int64_t* ns, const float* const* sx,
                       const float* const* sy, const float* const* ss, const float* const* sr,
                       int64_t nt, const float* tx, const float* ty, const float* tr,
                       int finalize, double fsx, double fsy, float* tu, float* tv) {
  if (nsets < 0 || nt < 0 || !tu || !tv) return -1;
  std::vector<Collection> vort;
  for (int k = 0; k < nsets; ++k) vort.emplace_back(dropin_points(ns[k], sx[k], sy[k], ss[k], sr[k]));
  std::vector<float> ones(nt, 1.0f);
  std::vector<Collection> targs;
  targs.emplace_back(dropin_points(nt, tx, ty, ones.data(), tr));

  InfluenceVisitor<float, double> visitor = {ResultsType(velonly), ExecEnv(false, direct, cpu_x86)};
  for (auto& targ : targs) {
    std::visit([=](auto& elem) { elem.zero_vels(); }, targ);
    for (auto& src : vort) std::visit(visitor, src, targ);
    if (finalize)
      std::visit([=](auto& elem) { elem.finalize_vels(std::array<double, Dimensions>({fsx, fsy})); },
                 targ);
  }
  const Points<float>& out = std::get<Points<float>>(targs[0]);
  std::memcpy(tu, out.get_vel()[0].data(), nt * sizeof(float));
  std::memcpy(tv, out.get_vel()[1].data(), nt * sizeof(float));
  return 0;
}

}  // extern "C"
