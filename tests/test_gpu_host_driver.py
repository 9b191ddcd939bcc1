"""End-to-end check of the C++ host mirror (omega2d_amd/host: ExecEnv gpu_hip, Influence
operators, BEM/solve_bem, Convection) through tests/cpp/host_driver.cpp against a pipeline of
the fp64 oracle that restates src/Convection.h and src/BEMHelper.h step by step."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from helpers import circle_panels, rel_l2
from omega2d_amd import build as B
from omega2d_amd import workloads as W

pytestmark = pytest.mark.gpu

CORE_ID = {"rm": 1, "gauss": 2, "wl": 3, "v2": 4, "v3": 5}


def run_driver(tmp_path, mode, header, payload, nout):
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.asarray(header, dtype=np.int64).tobytes())
        for a in payload:
            f.write(np.ascontiguousarray(a).tobytes())
    assert os.path.exists(B.HOST_TEST_BIN), "host driver not built"
    r = subprocess.run([B.HOST_TEST_BIN, mode, str(inp), str(out)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    data = np.fromfile(out, dtype=np.float64)
    assert data.size == nout
    return data


def oracle_vel(core, x, y, s, r, fs=(0.0, 0.0)):
    u, v = O.points_on_points(core, x, y, s, r, x, y, r)
    return O.finalize(u, v, fs)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_rk_step_free_particles(tmp_path, order):
    w = W.c1_disc(n=3000)
    x, y, s, r = w["x"], w["y"], w["s"], w["r"]
    dt = 0.01
    n = len(x)
    got = run_driver(tmp_path, "rk", [n, order, CORE_ID["rm"], 0, 1], [np.float64(dt), x, y, s, r],
                     4 * n)
    # the Ralston schemes of src/Convection.h with oracle velocities; after two or three stages
    # the particles hold the stage combination (src/ElementBase.h:308-312, 340-345), after one
    # the stage velocity itself
    u0 = oracle_vel("rm", x, y, s, r)
    if order == 1:
        ex, ey = O.rk_combine(x, y, [u0[0]], [u0[1]], [1.0], dt)
        eu, ev = u0
    elif order == 2:
        x1, y1 = O.rk_combine(x, y, [u0[0]], [u0[1]], [1.0], 2.0 / 3.0 * dt)
        u1 = oracle_vel("rm", x1, y1, s, r)
        ex, ey = O.rk_combine(x, y, [u0[0], u1[0]], [u0[1], u1[1]], [0.25, 0.75], dt)
        eu, ev = 0.25 * u0[0] + 0.75 * u1[0], 0.25 * u0[1] + 0.75 * u1[1]
    else:
        x1, y1 = O.rk_combine(x, y, [u0[0]], [u0[1]], [1.0], 0.5 * dt)
        u1 = oracle_vel("rm", x1, y1, s, r)
        x2, y2 = O.rk_combine(x, y, [u1[0]], [u1[1]], [1.0], 0.75 * dt)
        u2 = oracle_vel("rm", x2, y2, s, r)
        ex, ey = O.rk_combine(x, y, [u0[0], u1[0], u2[0]], [u0[1], u1[1], u2[1]],
                              [2 / 9, 3 / 9, 4 / 9], dt)
        eu = 2 / 9 * u0[0] + 3 / 9 * u1[0] + 4 / 9 * u2[0]
        ev = 2 / 9 * u0[1] + 3 / 9 * u1[1] + 4 / 9 * u2[1]
    gx, gy, gu, gv = got[:n], got[n:2 * n], got[2 * n:3 * n], got[3 * n:]
    assert np.max(np.abs(gx - ex)) <= 1e-14 and np.max(np.abs(gy - ey)) <= 1e-14
    assert rel_l2(gu, eu) <= 1e-12 and rel_l2(gv, ev) <= 1e-12


def test_rk2_step_with_treecode_summation(tmp_path):
    """Convection with ExecEnv summation barneshut: the host mirror's points_affect_points takes
    the treecode (theta 0.4, order 30); one Ralston RK2 step against the oracle's direct one."""
    rng = np.random.default_rng(9)
    n = 6000
    x, y = rng.normal(0, 0.5, n), rng.normal(0, 0.5, n)
    s, r = rng.normal(0, 1.0 / n, n), np.full(n, 2e-3)
    dt = 0.01
    got = run_driver(tmp_path, "rktree", [n, 2, CORE_ID["gauss"], 0, 1],
                     [np.float64(dt), x, y, s, r], 4 * n)
    u0 = oracle_vel("gauss", x, y, s, r)
    x1, y1 = O.rk_combine(x, y, [u0[0]], [u0[1]], [1.0], 2.0 / 3.0 * dt)
    u1 = oracle_vel("gauss", x1, y1, s, r)
    ex, ey = O.rk_combine(x, y, [u0[0], u1[0]], [u0[1], u1[1]], [0.25, 0.75], dt)
    scale = dt * np.max(np.hypot(u0[0], u0[1]))
    assert np.max(np.abs(got[:n] - ex)) <= 1e-9 * scale
    assert np.max(np.abs(got[n:2 * n] - ey)) <= 1e-9 * scale


def test_rk_step_float_storage(tmp_path):
    # the reference application's STORE=float through the _f entry point
    w = W.c1_disc(n=2000)
    x, y, s, r = (w[k].astype(np.float32).astype(np.float64) for k in ("x", "y", "s", "r"))
    n = len(x)
    got = run_driver(tmp_path, "rk", [n, 2, CORE_ID["rm"], 1, 1], [np.float64(0.01), x, y, s, r],
                     4 * n)
    u0 = oracle_vel("rm", x, y, s, r)
    x1, y1 = O.rk_combine(x, y, [u0[0]], [u0[1]], [1.0], 2.0 / 3.0 * 0.01)
    u1 = oracle_vel("rm", x1, y1, s, r)
    ex, ey = O.rk_combine(x, y, [u0[0], u1[0]], [u0[1], u1[1]], [0.25, 0.75], 0.01)
    # positions stored as float: agreement to float rounding of the positions
    assert np.max(np.abs(got[:n] - ex)) <= 1e-7 and np.max(np.abs(got[n:2 * n] - ey)) <= 1e-7
    # and the combined velocity, stored as float
    eu, ev = 0.25 * u0[0] + 0.75 * u1[0], 0.25 * u0[1] + 0.75 * u1[1]
    assert rel_l2(got[2 * n:3 * n], eu) <= 1e-6 and rel_l2(got[3 * n:], ev) <= 1e-6


class PyBem:
    """solve_bem + find_vels (src/BEMHelper.h:43-260, src/Convection.h:172-243) for one ground
    circle, restated on the oracle."""

    def __init__(self, nx, ny, idx, nunk, core, fs):
        self.nx, self.ny, self.idx = nx.astype(np.float64), ny.astype(np.float64), idx
        self.nunk, self.core, self.fs = nunk, core, fs
        self.np = len(idx) // 2
        self.tx, self.ty, self.nnx, self.nny, self.area = O.panel_geometry(self.nx, self.ny, idx)
        self.vol, self.cx, self.cy = O.body_volume(self.nx, self.ny, idx)
        self.rvs, self.rss = O.unit_rot_strengths(self.nx, self.ny, idx, self.cx, self.cy)
        self.A = None
        self.omega_error = 0.0
        self.omega = 0.0

    def assemble(self):
        a, b = self.idx[0::2], self.idx[1::2]
        mx = (0.5 * self.nx[b] + 0.5 * self.nx[a] + 1e-4 * self.nnx).astype(np.float32)
        my = (0.5 * self.ny[b] + 0.5 * self.ny[a] + 1e-4 * self.nny).astype(np.float32)
        uu, uv = O.panels_on_points(self.nx, self.ny, self.idx, self.rvs, self.rss, mx, my)
        uu, uv = O.finalize(uu, uv)
        self.A = O.bem_assemble(self.nx, self.ny, self.idx, self.nunk, True, uu, uv, self.vol)

    def solve(self, wx, wy, ws):
        pu, pv = O.points_on_panels(wx, wy, ws, self.nx, self.ny, self.idx)
        pu, pv = O.finalize(pu, pv, self.fs)
        rhs = O.vels_to_rhs(pu, pv, self.tx, self.ty, self.nunk, np.zeros(self.np))
        rhs = np.concatenate([rhs, [2.0 * self.vol * self.omega_error]])
        if self.A is None:
            self.assemble()
        sol = np.linalg.solve(self.A, rhs)
        self.omega = sol[-1]
        self.omega_error = self.omega
        sol = sol[:-1]
        if self.nunk == 2:
            self.vs, self.ss = sol[0::2].copy(), sol[1::2].copy()
        else:
            self.vs, self.ss = sol.copy(), np.zeros(self.np)
        return rhs

    def wake_vel(self, wx, wy, ws, wr):
        u, v = O.points_on_points(self.core, wx, wy, ws, wr, wx, wy, wr)
        pu, pv = O.panels_on_points(self.nx, self.ny, self.idx, self.vs + self.omega * self.rvs,
                                    self.ss + self.omega * self.rss, wx, wy)
        return O.finalize(u + pu, v + pv, self.fs)


def bem_case(nunk, npan=128, nw=1500):
    w = W.c3_circle_wake(npan=npan, nwake=nw)
    nx, ny, idx = circle_panels(npan)
    return w, nx, ny, idx


@pytest.mark.parametrize("nunk", [1, 2])
def test_bem_solve_and_wake_velocity(tmp_path, nunk):
    w, nx, ny, idx = bem_case(nunk)
    nw = len(w["x"])
    header = [len(idx) // 2, len(nx), nw, nunk, CORE_ID["v2"], 2]
    payload = [np.array([1.0, 0.0, 0.01, 0.1]), nx, ny, idx, w["x"], w["y"], w["s"], w["r"]]
    npan = len(idx) // 2
    nrow = npan * nunk + 1
    got = run_driver(tmp_path, "bem", header, payload, 2 * npan + 1 + nrow + 2 * nw)
    gvs, gss, gom = got[:npan], got[npan:2 * npan], got[2 * npan]
    grhs = got[2 * npan + 1:2 * npan + 1 + nrow]
    gu, gv = got[2 * npan + 1 + nrow:2 * npan + 1 + nrow + nw], got[2 * npan + 1 + nrow + nw:]
    ref = PyBem(nx, ny, idx, nunk, "v2", (1.0, 0.0))
    rhs = ref.solve(w["x"], w["y"], w["s"])
    assert rel_l2(grhs, rhs) <= 1e-12
    assert rel_l2(gvs, ref.vs) <= 1e-10
    if nunk == 2:
        assert rel_l2(gss, ref.ss) <= 1e-10
    assert abs(gom - ref.omega) <= 1e-10
    eu, ev = ref.wake_vel(w["x"], w["y"], w["s"], w["r"])
    assert rel_l2(gu, eu) <= 1e-10 and rel_l2(gv, ev) <= 1e-10
    # physics: the sheet on the circle in uniform flow is close to 2|sin theta|
    if nunk == 1:
        mx = 0.5 * (nx[idx[0::2]].astype(np.float64) + nx[idx[1::2]])
        my = 0.5 * (ny[idx[0::2]].astype(np.float64) + ny[idx[1::2]])
        th = np.arctan2(my, mx)
        assert np.max(np.abs(np.abs(gvs) - 2 * np.abs(np.sin(th)))) < 0.1


def test_bem_velocity_and_vorticity_split(tmp_path):
    """find_vels with velandvort after the solve (src/Simulation.cpp:412), the wake in the split
    form: the panels add velocity only and leave the vorticity to the particles
    (src/Influence.h:422-614)."""
    w, nx, ny, idx = bem_case(2)
    nw = len(w["x"])
    header = [len(idx) // 2, len(nx), nw, 2, CORE_ID["v2"], 2]
    payload = [np.array([1.0, 0.0, 0.01, 0.1]), nx, ny, idx, w["x"], w["y"], w["s"], w["r"]]
    got = run_driver(tmp_path, "bemvort", header, payload, 3 * nw)
    gu, gv, gw = got[:nw], got[nw:2 * nw], got[2 * nw:]
    ref = PyBem(nx, ny, idx, 2, "v2", (1.0, 0.0))
    ref.solve(w["x"], w["y"], w["s"])
    eu, ev = ref.wake_vel(w["x"], w["y"], w["s"], w["r"])
    assert rel_l2(gu, eu) <= 1e-6 and rel_l2(gv, ev) <= 1e-6
    _, _, ew = O.points_on_points("v2", w["x"], w["y"], w["s"], w["r"], w["x"], w["y"], w["r"],
                                  restype="vort")
    assert rel_l2(gw, ew * (0.5 / np.pi)) <= 1e-6


def test_bem_rk2_step_with_wake(tmp_path):
    """Ralston RK2 with the BEM solved at both stages and the inner layer cleared after each
    move (method 1, src/Convection.h:413,464), as src/Convection.h:390-466 does it."""
    w, nx, ny, idx = bem_case(2, npan=96, nw=1000)
    nw = len(w["x"])
    dt, ips = 0.01, 0.5  # a layer 0.1 thick: wake particles start from radius 0.52
    header = [len(idx) // 2, len(nx), nw, 2, CORE_ID["v2"], 2]
    payload = [np.array([1.0, 0.0, dt, ips]), nx, ny, idx, w["x"], w["y"], w["s"], w["r"]]
    got = run_driver(tmp_path, "bemstep", header, payload, 4 * nw)
    ref = PyBem(nx, ny, idx, 2, "v2", (1.0, 0.0))
    x, y, s, r = w["x"], w["y"], w["s"], w["r"]
    n64x, n64y = nx.astype(np.float64), ny.astype(np.float64)
    _, _, pnx, pny, _ = O.panel_geometry(n64x, n64y, idx)
    pushed = []

    def clear(px, py):
        c = O.clear_inner(1, n64x, n64y, idx, pnx, pny, 0.5 / np.sqrt(2.0 * np.pi), ips, px, py,
                          s, r)
        pushed.append(c["num_cropped"])
        return c["x"], c["y"]

    ref.solve(x, y, s)
    u0 = ref.wake_vel(x, y, s, r)
    x1, y1 = clear(*O.rk_combine(x, y, [u0[0]], [u0[1]], [1.0], 2.0 / 3.0 * dt))
    ref.solve(x1, y1, s)
    u1 = ref.wake_vel(x1, y1, s, r)
    ex, ey = clear(*O.rk_combine(x, y, [u0[0], u1[0]], [u0[1], u1[1]], [0.25, 0.75], dt))
    assert min(pushed) > 0
    assert np.max(np.abs(got[:nw] - ex)) <= 1e-12 and np.max(np.abs(got[nw:2 * nw] - ey)) <= 1e-12
    # the particles hold the stage combination 0.25 u0 + 0.75 u1 (src/ElementBase.h:308-312)
    assert rel_l2(got[2 * nw:3 * nw], 0.25 * u0[0] + 0.75 * u1[0]) <= 1e-10
    assert rel_l2(got[3 * nw:], 0.25 * u0[1] + 0.75 * u1[1]) <= 1e-10


def test_find_vort_self(tmp_path):
    """Convection::find_vort (src/Convection.h:134-166): Gaussian vorticity of the particles on
    themselves, finalized (1/2pi), against the oracle."""
    rng = np.random.default_rng(8)
    n = 3000
    x, y = rng.normal(0, 0.5, n), rng.normal(0, 0.5, n)
    s, r = rng.normal(0, 1e-3, n), rng.uniform(0.02, 0.05, n)
    got = run_driver(tmp_path, "vort", [n], [x, y, s, r], n)
    ref = O.points_vorticity(x, y, s, r, x, y) * (0.5 / np.pi)
    assert rel_l2(got, ref) <= 1e-12


class MultiBem:
    """solve_bem + find_vels for several reactive bodies (src/BEMHelper.h:43-260), restated on
    the oracle: self blocks, blocks between bodies (with the source body's unit-rotation
    velocity on the target's collocation points), one augmented row and column per body."""

    def __init__(self, bodies, nunk, core, fs):
        self.b = []
        for nx, ny, idx in bodies:
            nx, ny = nx.astype(np.float64), ny.astype(np.float64)
            tx, ty, nnx, nny, area = O.panel_geometry(nx, ny, idx)
            vol, cx, cy = O.body_volume(nx, ny, idx)
            rvs, rss = O.unit_rot_strengths(nx, ny, idx, cx, cy)
            a, b = idx[0::2], idx[1::2]
            mx = (0.5 * nx[b] + 0.5 * nx[a] + 1e-4 * nnx).astype(np.float32)
            my = (0.5 * ny[b] + 0.5 * ny[a] + 1e-4 * nny).astype(np.float32)
            self.b.append(dict(nx=nx, ny=ny, idx=idx, tx=tx, ty=ty, vol=vol, rvs=rvs, rss=rss,
                               mx=mx, my=my, np=len(idx) // 2))
        self.nunk, self.core, self.fs = nunk, core, fs

    def rows(self, k):
        return self.b[k]["np"] * self.nunk + 1

    def solve(self, wx, wy, ws):
        n = sum(self.rows(k) for k in range(len(self.b)))
        A = np.zeros((n, n))
        rhs = np.zeros(n)
        r0 = np.cumsum([0] + [self.rows(k) for k in range(len(self.b))])
        for t, T in enumerate(self.b):
            pu, pv = O.points_on_panels(wx, wy, ws, T["nx"], T["ny"], T["idx"])
            pu, pv = O.finalize(pu, pv, self.fs)
            rhs[r0[t]:r0[t + 1] - 1] = O.vels_to_rhs(pu, pv, T["tx"], T["ty"], self.nunk,
                                                     np.zeros(T["np"]))
            for s_, S in enumerate(self.b):
                uu, uv = O.panels_on_points(S["nx"], S["ny"], S["idx"], S["rvs"], S["rss"],
                                            T["mx"], T["my"])
                uu, uv = O.finalize(uu, uv)
                if s_ == t:
                    blk = O.bem_assemble(T["nx"], T["ny"], T["idx"], self.nunk, True, uu, uv,
                                         T["vol"])
                else:
                    blk = O.bem_assemble_pair(S["nx"], S["ny"], S["idx"], T["nx"], T["ny"],
                                              T["idx"], self.nunk, True, True, uu, uv)
                A[r0[t]:r0[t + 1], r0[s_]:r0[s_ + 1]] = blk
        sol = np.linalg.solve(A, rhs)
        for k, B in enumerate(self.b):
            part = sol[r0[k]:r0[k + 1]]
            B["omega"] = part[-1]
            if self.nunk == 2:
                B["vs"], B["ss"] = part[:-1][0::2].copy(), part[:-1][1::2].copy()
            else:
                B["vs"], B["ss"] = part[:-1].copy(), np.zeros(B["np"])

    def wake_vel(self, wx, wy, ws, wr):
        u, v = O.points_on_points(self.core, wx, wy, ws, wr, wx, wy, wr)
        for B in self.b:
            pu, pv = O.panels_on_points(B["nx"], B["ny"], B["idx"], B["vs"] + B["omega"] * B["rvs"],
                                        B["ss"] + B["omega"] * B["rss"], wx, wy)
            u, v = u + pu, v + pv
        return O.finalize(u, v, self.fs)


@pytest.mark.parametrize("nunk", [1, 2])
def test_bem_two_bodies_and_wake_velocity(tmp_path, nunk):
    """A ground circle and a second cylinder: the host mirror's solve_bem assembles the four
    blocks (two self, two between the bodies) and solves; strengths, rotation rates and the
    wake velocities against the oracle restatement."""
    ax, ay, ai = circle_panels(96)
    bx, by, bi = circle_panels(48, diam=0.6, cx=1.6, cy=0.3)
    rng = np.random.default_rng(21)
    nw = 800
    r = np.sqrt(rng.uniform(0.6 ** 2, 2.5 ** 2, nw))
    th = rng.uniform(0, 2 * np.pi, nw)
    wx, wy = r * np.cos(th), r * np.sin(th)
    far = np.hypot(wx - 1.6, wy - 0.3) > 0.4
    wx, wy = wx[far], wy[far]
    nw = len(wx)
    ws, wr = rng.normal(0, 2e-4, nw), np.full(nw, 0.005)
    header = [len(ai) // 2, len(ax), len(bi) // 2, len(bx), nw, nunk, CORE_ID["v2"]]
    payload = [np.array([1.0, 0.0]), ax, ay, ai, bx, by, bi, wx, wy, ws, wr]
    na, nb = len(ai) // 2, len(bi) // 2
    got = run_driver(tmp_path, "bem2", header, payload, 2 * na + 1 + 2 * nb + 1 + 2 * nw)
    ref = MultiBem([(ax, ay, ai), (bx, by, bi)], nunk, "v2", (1.0, 0.0))
    ref.solve(wx, wy, ws)
    o = 0
    for B in ref.b:
        gvs, gss, gom = got[o:o + B["np"]], got[o + B["np"]:o + 2 * B["np"]], got[o + 2 * B["np"]]
        o += 2 * B["np"] + 1
        assert rel_l2(gvs, B["vs"]) <= 1e-10
        if nunk == 2:
            assert rel_l2(gss, B["ss"]) <= 1e-10
        assert abs(gom - B["omega"]) <= 1e-10
    eu, ev = ref.wake_vel(wx, wy, ws, wr)
    assert rel_l2(got[o:o + nw], eu) <= 1e-10 and rel_l2(got[o + nw:], ev) <= 1e-10



@pytest.mark.parametrize("nunk", [1, 2])
def test_moving_body_rebuilds_only_the_blocks_in_relative_motion(tmp_path, nunk):
    """A ground circle and a cylinder that translates and rotates (motion 2 of the driver:
    x = 1.5 + 0.1 sin t, theta = 0.3 t). solve_bem at t0 builds all four blocks; at t1 only the
    two blocks between the bodies are rebuilt (src/BEMHelper.h:147-199, src/Body.cpp:237-265),
    the self blocks being unchanged by a rigid motion; a second call at t1 rebuilds nothing.
    The matrix and strengths then match a system with the same history whose matrix was rebuilt
    in full at t1."""
    ax, ay, ai = circle_panels(96)
    bx, by, bi = circle_panels(48, diam=0.6, cx=0.0, cy=0.0)
    rng = np.random.default_rng(22)
    nw = 600
    r = np.sqrt(rng.uniform(0.6 ** 2, 2.5 ** 2, nw))
    th = rng.uniform(0, 2 * np.pi, nw)
    wx, wy = r * np.cos(th), r * np.sin(th)
    keep = np.hypot(wx - 1.5, wy) > 0.5
    wx, wy = wx[keep], wy[keep]
    nw = len(wx)
    ws, wr = rng.normal(0, 2e-4, nw), np.full(nw, 0.005)
    header = [len(ai) // 2, len(ax), len(bi) // 2, len(bx), nw, nunk, CORE_ID["v2"]]
    payload = [np.array([0.0, 0.05]), ax, ay, ai, bx, by, bi, wx, wy, ws, wr]
    na, nb = len(ai) // 2, len(bi) // 2
    per = 2 * na + 1 + 2 * nb + 1
    got = run_driver(tmp_path, "bemmove", header, payload, 4 + 2 * per)
    assert got[:3].tolist() == [4.0, 6.0, 6.0]
    # the kept self blocks were built from the geometry at t0. A rigid motion leaves them
    # unchanged except for rounding: the unit-rotation column is evaluated at collocation points
    # rounded to float (src/Surfaces.h:830-878), which moves with the body, so the two matrices
    # agree to float precision there and to fp64 rounding elsewhere
    assert got[3] <= 1e-7
    moved, fresh = got[4:4 + per], got[4 + per:]
    assert rel_l2(moved, fresh) <= 1e-6


@pytest.mark.parametrize("nunk", [1, 2])
def test_moving_body_boundary_condition_is_galilean(tmp_path, nunk):
    """A cylinder moving at unit speed through still fluid (its motion enters the right-hand side
    through add_body_motion, src/Surfaces.h:492-523) and the same cylinder held still in a
    freestream (-1, 0) have the same boundary condition: the strengths agree, and the vortex
    sheet follows the potential-flow slip 2|sin theta| of a circle."""
    nx, ny, idx = circle_panels(128)
    npan = len(idx) // 2
    t = 0.3
    got = run_driver(tmp_path, "bemgalilean", [npan, len(nx), nunk], [np.array([t]), nx, ny, idx],
                     2 * (2 * npan + 1))
    moving, still = got[:2 * npan + 1], got[2 * npan + 1:]
    assert rel_l2(moving, still) <= 1e-8
    if nunk == 1:
        mx = 0.5 * (nx[idx[0::2]].astype(np.float64) + nx[idx[1::2]])
        my = 0.5 * (ny[idx[0::2]].astype(np.float64) + ny[idx[1::2]])
        th = np.arctan2(my, mx)
        assert np.max(np.abs(np.abs(moving[:npan]) - 2 * np.abs(np.sin(th)))) < 0.1
