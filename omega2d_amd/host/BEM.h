// BEM.h - boundary-element system of the gpu_hip environment: host mirror of src/BEM.h
// (BEM<S,I>: set_block/set_rhs/solve/get_str), src/Coefficients.h (panels_on_panels_coeff of a
// body on itself or on another body), src/RHS.h (vels_to_rhs_panels) and src/BEMHelper.h
// (solve_bem over any number of bodies).
// The matrix is assembled and LU-solved in fp64 on the device; the factors are kept until the
// panels change, like the reference keeps its solver (src/BEM.h:182-195).
#pragma once

#include <algorithm>
#include <cstdio>
#include <stdexcept>
#include <variant>
#include <vector>

#include "Influence.h"

namespace omega2d {

template <class S, class I>
class BEM {
 public:
  bool is_A_current() const { return A_is_current; }
  void just_made_A() { A_is_current = true; }
  void panels_changed() {
    A_is_current = false;
    solver_initialized = false;
  }
  // the time of the previous solve_bem call (a function-static in src/BEMHelper.h:54; kept per
  // system here so that two systems do not share it)
  double last_time() const { return last_t; }
  void set_last_time(double t) { last_t = t; }
  // blocks assembled since construction (src/BEM.h has no counter; for tests of the rebuild
  // policy)
  size_t blocks_built() const { return nblocks; }
  void set_device(int32_t d) { device = d; }

  // column-major block copy into A, growing A as needed (src/BEM.h:116-136)
  void set_block(size_t rstart, size_t nrows, size_t cstart, size_t ncols,
                 const std::vector<double>& in) {
    ++nblocks;
    resize(std::max(nr, rstart + nrows), std::max(nc, cstart + ncols));
    size_t k = 0;
    for (size_t j = 0; j < ncols; ++j)
      for (size_t i = 0; i < nrows; ++i) A[(cstart + j) * nr + rstart + i] = in[k++];
  }

  void set_rhs(size_t cstart, size_t n, const std::vector<S>& in) {
    if (b.size() < cstart + n) b.resize(cstart + n, 0.0);
    for (size_t i = 0; i < n; ++i) b[cstart + i] = double(in[i]);
  }

  void solve() {
    if (nr != nc || nr != b.size()) throw std::runtime_error("BEM system is not square");
    o2d_env e = {O2D_CORE_V2, O2D_PLAIN_FP64, O2D_VELONLY, device, nullptr};
    if (!solver_initialized) {
      o2d_check(o2d_hip_bem_factor_d(&e, (int64_t)nr, A.data(), (int64_t)nr), "bem_factor");
      solver_initialized = true;
    }
    x.assign(nr, 0.0);
    o2d_check(o2d_hip_bem_solve_d(&e, (int64_t)nr, b.data(), x.data()), "bem_solve");
  }

  std::vector<S> get_str(size_t start, size_t n) const {
    return std::vector<S>(x.begin() + start, x.begin() + start + n);
  }
  const std::vector<double>& matrix() const { return A; }
  const std::vector<double>& rhs() const { return b; }
  size_t rows() const { return nr; }

 private:
  void resize(size_t r, size_t c) {
    if (r == nr && c == nc) return;
    std::vector<double> nA(r * c, 0.0);
    for (size_t j = 0; j < std::min(c, nc); ++j)
      for (size_t i = 0; i < std::min(r, nr); ++i) nA[j * r + i] = A[j * nr + i];
    A.swap(nA);
    nr = r;
    nc = c;
  }

  std::vector<double> A, b, x;
  size_t nr = 0, nc = 0;
  bool A_is_current = false, solver_initialized = false;
  double last_t = -99.9;
  size_t nblocks = 0;
  int32_t device = 0;
};

// one block of the influence matrix (src/Coefficients.h:155-573), assembled on the device: a
// body on itself, or a source body on a different target body. For an augmented source the
// target's panel velocities must hold the source's unit-rotation velocity (solve_bem sets them).
template <class S>
std::vector<double> panels_on_panels_coeff(const Surfaces<S>& src, Surfaces<S>& targ,
                                           const ExecEnv& env) {
  const int32_t nunk = (int32_t)targ.num_unknowns_per_panel();
  const o2d_env e = env.c_env(velonly);
  const DoubleView<S> tx(targ.get_pos()[0]), ty(targ.get_pos()[1]);
  const DoubleView<S> uu(targ.get_vel()[0]), uv(targ.get_vel()[1]);
  if (&src == &targ) {
    const bool aug = targ.is_augmented();
    const int64_t n = (int64_t)targ.get_num_rows();
    std::vector<double> coeffs((size_t)(n * n), 0.0);
    o2d_check(o2d_hip_bem_assemble_d(&e, (int64_t)targ.get_n(), tx.data(), ty.data(),
                                     (int64_t)targ.get_npanels(), targ.get_idx().data(), nunk,
                                     aug ? 1 : 0, uu.data(), uv.data(), double(targ.get_vol()),
                                     coeffs.data(), n),
              "o2d_hip_bem_assemble");
    return coeffs;
  }
  if ((int32_t)src.num_unknowns_per_panel() != nunk)
    throw std::runtime_error("BEM blocks between bodies need the same unknowns per panel");
  const int64_t nr = (int64_t)targ.get_num_rows(), nc = (int64_t)src.get_num_rows();
  std::vector<double> coeffs((size_t)(nr * nc), 0.0);
  const DoubleView<S> sx(src.get_pos()[0]), sy(src.get_pos()[1]);
  o2d_check(o2d_hip_bem_assemble_pair_d(&e, (int64_t)src.get_n(), sx.data(), sy.data(),
                                        (int64_t)src.get_npanels(), src.get_idx().data(),
                                        (int64_t)targ.get_n(), tx.data(), ty.data(),
                                        (int64_t)targ.get_npanels(), targ.get_idx().data(), nunk,
                                        src.is_augmented() ? 1 : 0, targ.is_augmented() ? 1 : 0,
                                        uu.data(), uv.data(), coeffs.data(), nr),
            "o2d_hip_bem_assemble_pair");
  return coeffs;
}

// src/RHS.h:54-115
template <class S>
std::vector<S> vels_to_rhs_panels(const Surfaces<S>& targ) {
  const auto& tu = targ.get_vel();
  const auto& tt = targ.get_tang();
  const auto& tn = targ.get_norm();
  const size_t np = targ.get_npanels();
  std::vector<S> rhs(np * targ.num_unknowns_per_panel());
  if (targ.num_unknowns_per_panel() == 1) {
    for (size_t i = 0; i < np; ++i) rhs[i] = -(tu[0][i] * tt[0][i] + tu[1][i] * tt[1][i]);
  } else {
    const auto& nbc = targ.get_norm_bcs();
    for (size_t i = 0; i < np; ++i) {
      rhs[2 * i] = -(tu[0][i] * tt[0][i] + tu[1][i] * tt[1][i]);
      rhs[2 * i + 1] = -(tu[0][i] * tn[0][i] + tu[1][i] * tn[1][i]);
      rhs[2 * i + 1] += nbc[i];
    }
  }
  return rhs;
}

// src/BEMHelper.h:43-260 for boundaries that are Surfaces
template <class S, class A, class I>
void solve_bem(double time, const std::array<double, Dimensions>& fs,
               std::vector<Collection<S>>& vort, std::vector<Collection<S>>& bdry,
               BEM<S, I>& bem, const ExecEnv& env) {
  if (bdry.empty()) return;
  if (!bem.is_A_current()) {
    I row = 0;
    for (auto& t : bdry) {
      std::visit([&](auto& el) {
        if constexpr (std::is_same_v<std::decay_t<decltype(el)>, Surfaces<S>>) {
          el.set_first_row(row);
          row = el.get_next_row();
        }
      }, t);
    }
  }
  InfluenceVisitor<S, A> iv = {ResultsType(velonly), env};

  for (auto& t : bdry) {
    auto* surf = std::get_if<Surfaces<S>>(&t);
    if (!surf) continue;
    surf->transform(time);  // prescribed motion of the parent body
    surf->zero_vels();
    for (auto& s : vort) std::visit(iv, std::as_const(s), t);
    surf->finalize_vels(fs);
    surf->add_body_motion(S(-1.0), time);
    std::vector<S> rhs = vels_to_rhs_panels<S>(*surf);
    if (surf->is_augmented()) {
      const S tot = surf->get_last_body_circ() + surf->get_reabsorbed() +
                    surf->get_last_body_circ_error();
      rhs.push_back(tot);
    }
    bem.set_rhs(surf->get_first_row(), surf->get_num_rows(), rhs);
  }

  // src/BEMHelper.h:147-219: after a reset every block is built; at a new time only the blocks
  // between bodies in relative motion are rebuilt
  const bool rebuild_all = !bem.is_A_current();
  const bool check_motion = !rebuild_all && bem.last_time() != time;
  if (rebuild_all || check_motion) {
    const double last = bem.last_time();
    bool rebuilt = false;
    for (auto& t : bdry) {
      auto* targ = std::get_if<Surfaces<S>>(&t);
      if (!targ) continue;
      for (auto& sc : bdry) {
        auto* src = std::get_if<Surfaces<S>>(&sc);
        if (!src) continue;
        bool rebuild = rebuild_all;
        if (check_motion && targ->get_body_ptr() && src->get_body_ptr())
          rebuild = targ->get_body_ptr()->relative_motion_vs(*src->get_body_ptr(), last, time);
        if (!rebuild) continue;
        rebuilt = true;
        if (src->is_augmented()) {
          // velocity on the target's panels from the source's unit-rotation strengths
          targ->zero_vels();
          src->zero_strengths();
          src->add_unit_rot_strengths();
          panels_affect_panels<S, A>(*src, *targ, ResultsType(velonly), env);
          src->zero_strengths();
          targ->finalize_vels({0.0, 0.0});
        }
        const auto coeffs = panels_on_panels_coeff<S>(*src, *targ, env);
        bem.set_block(targ->get_first_row(), targ->get_num_rows(), src->get_first_row(),
                      src->get_num_rows(), coeffs);
      }
    }
    // the reference re-initialises its solver whenever the time changes; the factors here are
    // redone only when a block changed, which gives the same solution
    if (rebuilt) bem.panels_changed();
    bem.just_made_A();
  }

  bem.solve();
  for (auto& t : bdry) {
    auto* surf = std::get_if<Surfaces<S>>(&t);
    if (!surf) continue;
    surf->set_str(bem.get_str(surf->get_first_row(), surf->get_num_rows()));
  }
  bem.set_last_time(time);
}

}  // namespace omega2d
