// ExecEnv.h - execution environment with the gpu_hip accelerator (host mirror of
// src/ExecEnv.h:26-108 and src/ResultsType.h:26-93; same names and meanings, plus the run-time
// core/precision choices the HIP environment takes).
#pragma once

#include <cstdint>
#include <string>

#include "omega2d_hip.h"

namespace omega2d {

enum summation_t { direct = 1, barneshut = 2, vic = 3, fmm = 4 };

// cpu_x86/cpu_vc/gpu_opengl/gpu_cuda keep the reference's values; gpu_hip is the new one
enum accel_t { cpu_x86 = 1, cpu_vc = 2, gpu_opengl = 3, gpu_cuda = 4, gpu_hip = 5 };

class ExecEnv {
 public:
  ExecEnv(bool internal, summation_t summ, accel_t accel)
      : m_internal(internal), m_summ(summ), m_accel(accel) {}
  ExecEnv() : ExecEnv(true, direct, gpu_hip) {}

  void use_internal() { m_internal = true; }
  void use_external() { m_internal = false; }
  void set_internal(bool v) { m_internal = v; }
  bool is_internal() const { return m_internal; }
  void set_summation(summation_t s) { m_summ = s; }
  summation_t get_summation() const { return m_summ; }
  // treecode parameters (summation_t barneshut): opening angle, expansion order, core tolerance
  void set_tree(double theta, int32_t order, double core_tol = 1e-12) {
    m_tree = o2d_tree_opts{theta, order, 0, core_tol};
  }
  const o2d_tree_opts& get_tree() const { return m_tree; }
  void set_instrs(accel_t a) { m_accel = a; }
  accel_t get_instrs() const { return m_accel; }

  // gpu_hip settings: the regularised core (the reference fixes it at compile time,
  // src/CoreFunc.h:35-39), the precision form and the device ordinal
  void set_core(int32_t c) { m_core = c; }
  int32_t get_core() const { return m_core; }
  void set_precision(int32_t p) { m_prec = p; }
  int32_t get_precision() const { return m_prec; }
  void set_device(int32_t d) { m_device = d; }
  int32_t get_device() const { return m_device; }

  o2d_env c_env(int32_t results) const {
    o2d_env e;
    e.core = m_core;
    e.precision = m_prec;
    e.results = results;
    e.device = m_device;
    e.stream = nullptr;
    return e;
  }

  std::string to_string() const {
    if (!m_internal) return " external solver";
    std::string s;
    if (m_accel == cpu_x86) s += " native";
    else if (m_accel == cpu_vc) s += " Vc-accelerated";
    else if (m_accel == gpu_hip) s += " HIP (gfx950)";
    else s += " unknown acceleration";
    if (m_summ == direct) s += " direct sums";
    else if (m_summ == barneshut) s += " treecode";
    else s += " unknown algorithm";
    return s;
  }

 private:
  bool m_internal;
  summation_t m_summ;
  accel_t m_accel;
  int32_t m_core = O2D_CORE_V2;
  int32_t m_prec = O2D_PLAIN_FP64;
  int32_t m_device = 0;
  o2d_tree_opts m_tree = {0.6, 28, 0, 1e-12};
};

enum results_t { velonly = 1, velandgrad = 2, psionly = 3, velandvort = 4, velandshear = 5 };

class ResultsType {
 public:
  explicit ResultsType(results_t t = velonly) : m_t(t) {}
  bool compute_psi() const { return m_t == psionly; }
  bool compute_vel() const { return m_t != psionly; }
  bool compute_grad() const { return m_t == velandgrad; }
  bool compute_vort() const { return m_t == velandvort; }
  bool compute_shear() const { return m_t == velandshear; }
  results_t get_type() const { return m_t; }

 private:
  results_t m_t;
};

}  // namespace omega2d
