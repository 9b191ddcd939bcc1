// host_driver.cpp - drives the C++ host mirror (omega2d_amd/host) the way the reference's
// Simulation drives Convection: reads a binary case written by tests/test_gpu_host_driver.py,
// runs it through the gpu_hip environment, and writes the results back for the test to check
// against the fp64 oracle.
//
//   o2d_host_driver <mode> <in.bin> <out.bin>
//   mode: rk       one advection step of free particles (order, core, storage from the input)
//         rktree   the same step with the treecode summation (summation_t barneshut)
//         bem      BEM solve on a reactive circle + velocities of the wake particles
//         bemstep  one advection step of the wake with the BEM solved at every stage and the
//                  inner layer cleared after every move
//         vort     Convection::find_vort of particles on themselves
//         bem2     BEM solve on two bodies (a ground circle and a second, named body) plus the
//                  velocities of the wake particles
//         vtk      write particle, field-point and panel .vtu files into the directory given
//                  in place of <out.bin> (no GPU work)
//
// Input: int64 header [mode-specific counts], then float64 arrays (node coordinates as float32).
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "Convection.h"

using namespace omega2d;

namespace {

struct Reader {
  std::ifstream f;
  explicit Reader(const char* p) : f(p, std::ios::binary) {
    if (!f) throw std::runtime_error(std::string("cannot open ") + p);
  }
  int64_t i64() {
    int64_t v;
    f.read(reinterpret_cast<char*>(&v), 8);
    return v;
  }
  double f64() {
    double v;
    f.read(reinterpret_cast<char*>(&v), 8);
    return v;
  }
  template <class T>
  std::vector<T> vec(size_t n) {
    std::vector<T> v(n);
    f.read(reinterpret_cast<char*>(v.data()), sizeof(T) * n);
    if (!f) throw std::runtime_error("short input file");
    return v;
  }
};

struct Writer {
  std::ofstream f;
  explicit Writer(const char* p) : f(p, std::ios::binary) {}
  template <class T>
  void vec(const std::vector<T>& v) {
    std::vector<double> d(v.begin(), v.end());
    f.write(reinterpret_cast<const char*>(d.data()), 8 * d.size());
  }
  void f64(double v) { f.write(reinterpret_cast<const char*>(&v), 8); }
};

// header: n, order, core, storage(0 double / 1 float), precision; then dt, x, y, s, r
template <class S>
void run_rk(Reader& in, Writer& out, int64_t n, int order, int core, int prec,
            summation_t summ = direct) {
  const double dt = in.f64();
  const auto x = in.vec<double>(n), y = in.vec<double>(n), s = in.vec<double>(n),
             r = in.vec<double>(n);
  std::vector<S> xs(x.begin(), x.end()), ys(y.begin(), y.end()), ss(s.begin(), s.end()),
      rs(r.begin(), r.end());
  std::vector<Collection<S>> vort, bdry, fldpt;
  vort.emplace_back(Points<S>(n, xs.data(), ys.data(), ss.data(), rs.data()));
  ExecEnv env(true, summ, gpu_hip);
  env.set_core(core);
  env.set_precision(prec);
  env.set_tree(0.4, 30);
  Convection<S, double, Int> conv(env, order);
  BEM<S, Int> bem;
  conv.advect(0.0, dt, {0.0, 0.0}, S(0.1), vort, bdry, fldpt, bem);
  const auto& p = std::get<Points<S>>(vort[0]);
  out.vec(p.get_pos()[0]);
  out.vec(p.get_pos()[1]);
  // the velocity the step leaves in the particles (the stage combination for RK2 and RK3)
  out.vec(p.get_vel()[0]);
  out.vec(p.get_vel()[1]);
}

// header: npan, nnodes, nwake, nunk, core, order; then fsx, fsy, dt, ips, nodes x/y
// (float32), idx (uint32 pairs), wake x, y, s, r
void run_bem(Reader& in, Writer& out, const std::string& mode) {
  const int64_t npan = in.i64(), nnodes = in.i64(), nw = in.i64(), nunk = in.i64(),
                core = in.i64(), order = in.i64();
  const double fsx = in.f64(), fsy = in.f64(), dt = in.f64(), ips = in.f64();
  const auto nx = in.vec<float>(nnodes), ny = in.vec<float>(nnodes);
  const auto idx = in.vec<uint32_t>(2 * npan);
  auto wx = in.vec<double>(nw), wy = in.vec<double>(nw), ws = in.vec<double>(nw),
       wr = in.vec<double>(nw);
  auto ground = std::make_shared<Body>();
  Surfaces<double> circle(nnodes, nx.data(), ny.data(), npan, idx.data(), reactive, bodybound,
                          ground);
  if (nunk == 1) circle.set_vortex_only();
  std::vector<Collection<double>> vort, bdry, fldpt;
  vort.emplace_back(Points<double>(nw, wx.data(), wy.data(), ws.data(), wr.data()));
  bdry.emplace_back(circle);
  ExecEnv env(true, direct, gpu_hip);
  env.set_core((int32_t)core);
  Convection<double, double, Int> conv(env, (int)order);
  BEM<double, Int> bem;
  const std::array<double, Dimensions> fs = {fsx, fsy};
  if (mode == "bem") {
    conv.find_derivs(0.0, fs, bem, bdry, vort, fldpt);
    const auto& c = std::get<Surfaces<double>>(bdry[0]);
    out.vec(c.get_str());
    out.vec(c.get_src_str());
    out.f64(c.get_solved_omega());
    out.vec(bem.rhs());
    const auto& p = std::get<Points<double>>(vort[0]);
    out.vec(p.get_vel()[0]);
    out.vec(p.get_vel()[1]);
  } else if (mode == "bemvort") {
    // velocity and vorticity of the flow after the solve (src/Simulation.cpp:412), the wake
    // summed in the split form
    env.set_precision(O2D_SPLIT_FP32);
    Convection<double, double, Int> sconv(env, (int)order);
    sconv.find_derivs(0.0, fs, bem, bdry, vort, fldpt);
    sconv.find_vels(fs, vort, bdry, vort, velandvort, true);
    auto& p = std::get<Points<double>>(vort[0]);
    out.vec(p.get_vel()[0]);
    out.vec(p.get_vel()[1]);
    out.vec(p.get_vort());
  } else {
    conv.advect(0.0, dt, fs, ips, vort, bdry, fldpt, bem);
    const auto& p = std::get<Points<double>>(vort[0]);
    out.vec(p.get_pos()[0]);
    out.vec(p.get_pos()[1]);
    out.vec(p.get_vel()[0]);
    out.vec(p.get_vel()[1]);
  }
}

// Wall time of whole advection steps (src/Convection.h:249-598: every stage solves the BEM,
// sums panels and particles on the wake, moves it and clears the inner layer), after three
// untimed steps (the matrix is built and factored once; the second solve with the same factors
// builds the inverse). header: npan, nnodes, nwake, nunk, core, order, nsteps, precision; then
// as for run_bem. Output: milliseconds per step, then the wake positions.
void run_bem_time(Reader& in, Writer& out) {
  const int64_t npan = in.i64(), nnodes = in.i64(), nw = in.i64(), nunk = in.i64(),
                core = in.i64(), order = in.i64(), nsteps = in.i64(), prec = in.i64();
  const double fsx = in.f64(), fsy = in.f64(), dt = in.f64(), ips = in.f64();
  const auto nx = in.vec<float>(nnodes), ny = in.vec<float>(nnodes);
  const auto idx = in.vec<uint32_t>(2 * npan);
  auto wx = in.vec<double>(nw), wy = in.vec<double>(nw), ws = in.vec<double>(nw),
       wr = in.vec<double>(nw);
  auto ground = std::make_shared<Body>();
  Surfaces<double> circle(nnodes, nx.data(), ny.data(), npan, idx.data(), reactive, bodybound,
                          ground);
  if (nunk == 1) circle.set_vortex_only();
  std::vector<Collection<double>> vort, bdry, fldpt;
  vort.emplace_back(Points<double>(nw, wx.data(), wy.data(), ws.data(), wr.data()));
  bdry.emplace_back(circle);
  ExecEnv env(true, direct, gpu_hip);
  env.set_core((int32_t)core);
  env.set_precision((int32_t)prec);
  Convection<double, double, Int> conv(env, (int)order);
  BEM<double, Int> bem;
  const std::array<double, Dimensions> fs = {fsx, fsy};
  double t = 0.0;
  for (int k = 0; k < 3; ++k, t += dt) conv.advect(t, dt, fs, ips, vort, bdry, fldpt, bem);
  const auto t0 = std::chrono::steady_clock::now();
  for (int64_t k = 0; k < nsteps; ++k, t += dt) conv.advect(t, dt, fs, ips, vort, bdry, fldpt, bem);
  const auto t1 = std::chrono::steady_clock::now();
  out.f64(std::chrono::duration<double, std::milli>(t1 - t0).count() / (double)nsteps);
  const auto& p = std::get<Points<double>>(vort[0]);
  out.vec(p.get_pos()[0]);
  out.vec(p.get_pos()[1]);
}

// header: npanA, nnodesA, npanB, nnodesB, nwake, nunk, core; then fsx, fsy, nodes of A
// (float32 x, y), idxA, nodes of B, idxB, wake x, y, s, r
void run_bem2(Reader& in, Writer& out) {
  const int64_t npa = in.i64(), nna = in.i64(), npb = in.i64(), nnb = in.i64(), nw = in.i64(),
                nunk = in.i64(), core = in.i64();
  const double fsx = in.f64(), fsy = in.f64();
  const auto ax = in.vec<float>(nna), ay = in.vec<float>(nna);
  const auto ai = in.vec<uint32_t>(2 * npa);
  const auto bx = in.vec<float>(nnb), by = in.vec<float>(nnb);
  const auto bi = in.vec<uint32_t>(2 * npb);
  auto wx = in.vec<double>(nw), wy = in.vec<double>(nw), ws = in.vec<double>(nw),
       wr = in.vec<double>(nw);
  auto ground = std::make_shared<Body>();
  auto other = std::make_shared<Body>();
  other->name = "cylinder";
  Surfaces<double> sa(nna, ax.data(), ay.data(), npa, ai.data(), reactive, bodybound, ground);
  Surfaces<double> sb(nnb, bx.data(), by.data(), npb, bi.data(), reactive, bodybound, other);
  if (nunk == 1) {
    sa.set_vortex_only();
    sb.set_vortex_only();
  }
  std::vector<Collection<double>> vort, bdry, fldpt;
  vort.emplace_back(Points<double>(nw, wx.data(), wy.data(), ws.data(), wr.data()));
  bdry.emplace_back(sa);
  bdry.emplace_back(sb);
  ExecEnv env(true, direct, gpu_hip);
  env.set_core((int32_t)core);
  Convection<double, double, Int> conv(env, 2);
  BEM<double, Int> bem;
  conv.find_derivs(0.0, {fsx, fsy}, bem, bdry, vort, fldpt);
  for (auto& c : bdry) {
    const auto& surf = std::get<Surfaces<double>>(c);
    out.vec(surf.get_str());
    out.vec(surf.get_src_str());
    out.f64(surf.get_solved_omega());
  }
  const auto& p = std::get<Points<double>>(vort[0]);
  out.vec(p.get_vel()[0]);
  out.vec(p.get_vel()[1]);
}

// the prescribed motions the tests give bodies (the reference parses the same expressions)
void set_motion(Body& b, int64_t kind) {
  if (kind == 1) {  // "0.3*sin(2*t)", "0.1*t*t", "0.5*t"
    b.pos_func[0] = [](double t) { return 0.3 * std::sin(2 * t); };
    b.pos_func[1] = [](double t) { return 0.1 * t * t; };
    b.apos_func = [](double t) { return 0.5 * t; };
  } else if (kind == 2) {  // "1.5+0.1*sin(t)", "", "0.3*t"
    b.pos_func[0] = [](double t) { return 1.5 + 0.1 * std::sin(t); };
    b.apos_func = [](double t) { return 0.3 * t; };
  } else if (kind == 3) {  // "", "", "7*t" (rotation only, wraps past 2 pi)
    b.apos_func = [](double t) { return 7 * t; };
  } else if (kind == 4) {  // "t", "", "" (unit speed along x)
    b.pos_func[0] = [](double t) { return t; };
  }
}

// header: npan, nnodes, nunk; then t, nodes x/y (float32), idx. A cylinder moving at unit speed
// along x through still fluid, solved at time t, and the same cylinder held still where the
// first one is at t in a freestream (-1, 0): the two boundary conditions are the same, so are
// the solutions. Out: strengths, source strengths and solved rotation rate of each.
void run_bem_galilean(Reader& in, Writer& out) {
  const int64_t np = in.i64(), nn = in.i64(), nunk = in.i64();
  const double t = in.f64();
  const auto nx = in.vec<float>(nn), ny = in.vec<float>(nn);
  const auto idx = in.vec<uint32_t>(2 * np);
  ExecEnv env(true, direct, gpu_hip);
  env.set_core(O2D_CORE_V2);
  for (int still = 0; still < 2; ++still) {
    auto b = std::make_shared<Body>();
    b->name = "cylinder";
    if (still) b->pos = {t, 0.0};
    else set_motion(*b, 4);
    Surfaces<double> surf(nn, nx.data(), ny.data(), np, idx.data(), reactive, bodybound, b);
    if (nunk == 1) surf.set_vortex_only();
    std::vector<Collection<double>> vort, bdry;
    bdry.emplace_back(surf);
    BEM<double, Int> bem;
    const std::array<double, Dimensions> fs = {still ? -1.0 : 0.0, 0.0};
    solve_bem<double, double, Int>(t, fs, vort, bdry, bem, env);
    const auto& s2 = std::get<Surfaces<double>>(bdry[0]);
    out.vec(s2.get_str());
    out.vec(s2.get_src_str());
    out.f64(s2.get_solved_omega());
  }
}

// header: nnodes, npan, kind; then t, factor, nodes x/y (float32), idx. Body motion only (no
// device calls): transformed nodes, tangents, panel velocities and the body state.
void run_motion(Reader& in, Writer& out) {
  const int64_t nn = in.i64(), np = in.i64(), kind = in.i64();
  const double t = in.f64(), factor = in.f64();
  const auto nx = in.vec<float>(nn), ny = in.vec<float>(nn);
  const auto idx = in.vec<uint32_t>(2 * np);
  auto b = std::make_shared<Body>();
  b->name = "mover";
  set_motion(*b, kind);
  Surfaces<double> s(nn, nx.data(), ny.data(), np, idx.data(), reactive, bodybound, b);
  s.transform(t);
  s.zero_vels();
  s.add_body_motion(factor, t);
  out.vec(s.get_pos()[0]);
  out.vec(s.get_pos()[1]);
  out.vec(s.get_tang()[0]);
  out.vec(s.get_tang()[1]);
  out.vec(s.get_vel()[0]);
  out.vec(s.get_vel()[1]);
  const auto p = b->get_pos(t), v = b->get_vel(t);
  out.vec(std::vector<double>{p[0], p[1], v[0], v[1], b->get_orient(t), b->get_rotvel(t)});
}

// header as bem2 (npanA, nnodesA, npanB, nnodesB, nwake, nunk, core) plus t0, t1. Body A is the
// ground, body B moves by motion 2. One system is solved at t0, t1 and t1 again (only the blocks
// in relative motion rebuilt at t1); a second one at t0, then reset, t1 (every block rebuilt)
// and t1 again, so both carry the same circulation history. Out: blocks built after each solve
// of the first system, the relative difference of the two matrices, then the strengths and
// solved rotation rates of both bodies for each system.
void run_bem_move(Reader& in, Writer& out) {
  const int64_t npa = in.i64(), nna = in.i64(), npb = in.i64(), nnb = in.i64(), nw = in.i64(),
                nunk = in.i64(), core = in.i64();
  const double t0 = in.f64(), t1 = in.f64();
  const auto ax = in.vec<float>(nna), ay = in.vec<float>(nna);
  const auto ai = in.vec<uint32_t>(2 * npa);
  const auto bx = in.vec<float>(nnb), by = in.vec<float>(nnb);
  const auto bi = in.vec<uint32_t>(2 * npb);
  auto wx = in.vec<double>(nw), wy = in.vec<double>(nw), ws = in.vec<double>(nw),
       wr = in.vec<double>(nw);
  ExecEnv env(true, direct, gpu_hip);
  env.set_core((int32_t)core);
  auto make = [&](std::vector<Collection<double>>& bdry) {
    auto ground = std::make_shared<Body>();
    auto mover = std::make_shared<Body>();
    mover->name = "cylinder";
    set_motion(*mover, 2);
    Surfaces<double> sa(nna, ax.data(), ay.data(), npa, ai.data(), reactive, bodybound, ground);
    Surfaces<double> sb(nnb, bx.data(), by.data(), npb, bi.data(), reactive, bodybound, mover);
    if (nunk == 1) {
      sa.set_vortex_only();
      sb.set_vortex_only();
    }
    bdry.emplace_back(sa);
    bdry.emplace_back(sb);
  };
  std::vector<Collection<double>> vort;
  vort.emplace_back(Points<double>(nw, wx.data(), wy.data(), ws.data(), wr.data()));
  auto dump = [&](std::vector<Collection<double>>& bdry) {
    for (auto& c : bdry) {
      const auto& surf = std::get<Surfaces<double>>(c);
      out.vec(surf.get_str());
      out.vec(surf.get_src_str());
      out.f64(surf.get_solved_omega());
    }
  };
  const std::array<double, Dimensions> fs = {1.0, 0.0};
  std::vector<Collection<double>> bdry, fresh;
  make(bdry);
  BEM<double, Int> bem;
  solve_bem<double, double, Int>(t0, fs, vort, bdry, bem, env);
  const double b0 = (double)bem.blocks_built();
  solve_bem<double, double, Int>(t1, fs, vort, bdry, bem, env);
  const double b1 = (double)bem.blocks_built();
  solve_bem<double, double, Int>(t1, fs, vort, bdry, bem, env);  // same time: nothing rebuilt
  const double b2 = (double)bem.blocks_built();
  make(fresh);
  BEM<double, Int> bem2;
  solve_bem<double, double, Int>(t0, fs, vort, fresh, bem2, env);
  bem2.panels_changed();
  solve_bem<double, double, Int>(t1, fs, vort, fresh, bem2, env);
  solve_bem<double, double, Int>(t1, fs, vort, fresh, bem2, env);
  const auto& m1 = bem.matrix();
  const auto& m2 = bem2.matrix();
  double dd = 0.0, nn = 0.0;
  for (size_t i = 0; i < m1.size() && i < m2.size(); ++i) {
    dd += (m1[i] - m2[i]) * (m1[i] - m2[i]);
    nn += m2[i] * m2[i];
  }
  out.vec(std::vector<double>{b0, b1, b2, std::sqrt(dd / nn)});
  dump(bdry);
  dump(fresh);
}

// header: n; then x, y, s, r
void run_vort(Reader& in, Writer& out, int64_t n) {
  const auto x = in.vec<double>(n), y = in.vec<double>(n), s = in.vec<double>(n),
             r = in.vec<double>(n);
  std::vector<Collection<double>> vort, bdry, fldpt;
  vort.emplace_back(Points<double>(n, x.data(), y.data(), s.data(), r.data()));
  ExecEnv env(true, direct, gpu_hip);
  Convection<double, double, Int> conv(env, 2);
  conv.find_vort(vort, bdry, vort);
  out.vec(std::get<Points<double>>(vort[0]).get_vort());
}

// header: n, nnodes, npan, index, frameno, pindex, pframeno; then time, ptime, x, y, s, r, u,
// v, w (particles), nodes x/y (float32), idx (uint32 pairs), vortex and source strengths
void run_vtk(Reader& in, const std::string& dir) {
  const int64_t n = in.i64(), nnodes = in.i64(), npan = in.i64(), index = in.i64(),
                frameno = in.i64(), pindex = in.i64(), pframeno = in.i64();
  const double time = in.f64(), ptime = in.f64();
  const auto x = in.vec<double>(n), y = in.vec<double>(n), s = in.vec<double>(n),
             r = in.vec<double>(n), u = in.vec<double>(n), v = in.vec<double>(n),
             w = in.vec<double>(n);
  const auto nx = in.vec<float>(nnodes), ny = in.vec<float>(nnodes);
  const auto idx = in.vec<uint32_t>(2 * npan);
  const auto vs = in.vec<double>(npan), ss = in.vec<double>(npan);
  Points<double> part(n, x.data(), y.data(), s.data(), r.data());
  Points<double> fld(n, x.data(), y.data(), nullptr, nullptr);
  for (auto* p : {&part, &fld}) {
    p->get_vel()[0] = u;
    p->get_vel()[1] = v;
  }
  part.get_vort() = w;
  part.write_vtk((size_t)index, (size_t)frameno, time, dir);
  fld.write_vtk((size_t)index + 1, (size_t)frameno, time, dir);
  Surfaces<double> surf(nnodes, nx.data(), ny.data(), npan, idx.data(), reactive, fixed, nullptr);
  surf.get_src_str() = ss;
  surf.get_str() = vs;
  surf.write_vtk((size_t)pindex, (size_t)pframeno, ptime, dir);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 4) {
    std::cerr << "usage: o2d_host_driver <rk|rktree|bem|bemstep|bemvort|bemtime|bem2|bemmove|bemgalilean|vort|motion|vtk> <in.bin> <out.bin|dir>\n";
    return 2;
  }
  try {
    const std::string mode = argv[1];
    Reader in(argv[2]);
    if (mode == "vtk") {
      run_vtk(in, argv[3]);
      return 0;
    }
    if (mode == "motion") {
      Writer mo(argv[3]);
      run_motion(in, mo);
      return mo.f ? 0 : 1;
    }
    Writer out(argv[3]);
    if (mode == "rk" || mode == "rktree") {
      const int64_t n = in.i64(), order = in.i64(), core = in.i64(), storage = in.i64(),
                    prec = in.i64();
      const summation_t summ = mode == "rk" ? direct : barneshut;
      if (storage == 1) run_rk<float>(in, out, n, (int)order, (int)core, (int)prec, summ);
      else run_rk<double>(in, out, n, (int)order, (int)core, (int)prec, summ);
    } else if (mode == "bem" || mode == "bemstep" || mode == "bemvort") {
      run_bem(in, out, mode);
    } else if (mode == "vort") {
      run_vort(in, out, in.i64());
    } else if (mode == "bem2") {
      run_bem2(in, out);
    } else if (mode == "bemmove") {
      run_bem_move(in, out);
    } else if (mode == "bemgalilean") {
      run_bem_galilean(in, out);
    } else if (mode == "bemtime") {
      run_bem_time(in, out);

    } else {
      std::cerr << "unknown mode " << mode << "\n";
      return 2;
    }
    if (!out.f) throw std::runtime_error("write failed");
  } catch (const std::exception& e) {
    std::cerr << "host_driver: " << e.what() << "\n";
    return 1;
  }
  o2d_hip_release();
  return 0;
}
