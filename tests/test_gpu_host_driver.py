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
    got = run_driver(tmp_path, "rk", [len(x), order, CORE_ID["rm"], 0, 1], [np.float64(dt), x, y, s, r],
                     2 * len(x))
    # the Ralston schemes of src/Convection.h with oracle velocities
    u0 = oracle_vel("rm", x, y, s, r)
    if order == 1:
        ex, ey = O.rk_combine(x, y, [u0[0]], [u0[1]], [1.0], dt)
    elif order == 2:
        x1, y1 = O.rk_combine(x, y, [u0[0]], [u0[1]], [1.0], 2.0 / 3.0 * dt)
        u1 = oracle_vel("rm", x1, y1, s, r)
        ex, ey = O.rk_combine(x, y, [u0[0], u1[0]], [u0[1], u1[1]], [0.25, 0.75], dt)
    else:
        x1, y1 = O.rk_combine(x, y, [u0[0]], [u0[1]], [1.0], 0.5 * dt)
        u1 = oracle_vel("rm", x1, y1, s, r)
        x2, y2 = O.rk_combine(x, y, [u1[0]], [u1[1]], [1.0], 0.75 * dt)
        u2 = oracle_vel("rm", x2, y2, s, r)
        ex, ey = O.rk_combine(x, y, [u0[0], u1[0], u2[0]], [u0[1], u1[1], u2[1]],
                              [2 / 9, 3 / 9, 4 / 9], dt)
    gx, gy = got[: len(x)], got[len(x):]
    assert np.max(np.abs(gx - ex)) <= 1e-14 and np.max(np.abs(gy - ey)) <= 1e-14


def test_rk2_step_with_treecode_summation(tmp_path):
    """Convection with ExecEnv summation barneshut: the host mirror's points_affect_points takes
    the treecode (theta 0.4, order 30); one Ralston RK2 step against the oracle's direct one."""
    rng = np.random.default_rng(9)
    n = 6000
    x, y = rng.normal(0, 0.5, n), rng.normal(0, 0.5, n)
    s, r = rng.normal(0, 1.0 / n, n), np.full(n, 2e-3)
    dt = 0.01
    got = run_driver(tmp_path, "rktree", [n, 2, CORE_ID["gauss"], 0, 1],
                     [np.float64(dt), x, y, s, r], 2 * n)
    u0 = oracle_vel("gauss", x, y, s, r)
    x1, y1 = O.rk_combine(x, y, [u0[0]], [u0[1]], [1.0], 2.0 / 3.0 * dt)
    u1 = oracle_vel("gauss", x1, y1, s, r)
    ex, ey = O.rk_combine(x, y, [u0[0], u1[0]], [u0[1], u1[1]], [0.25, 0.75], dt)
    scale = dt * np.max(np.hypot(u0[0], u0[1]))
    assert np.max(np.abs(got[:n] - ex)) <= 1e-9 * scale
    assert np.max(np.abs(got[n:] - ey)) <= 1e-9 * scale


def test_rk_step_float_storage(tmp_path):
    # the reference application's STORE=float through the _f entry point
    w = W.c1_disc(n=2000)
    x, y, s, r = (w[k].astype(np.float32).astype(np.float64) for k in ("x", "y", "s", "r"))
    got = run_driver(tmp_path, "rk", [len(x), 2, CORE_ID["rm"], 1, 1], [np.float64(0.01), x, y, s, r],
                     2 * len(x))
    u0 = oracle_vel("rm", x, y, s, r)
    x1, y1 = O.rk_combine(x, y, [u0[0]], [u0[1]], [1.0], 2.0 / 3.0 * 0.01)
    u1 = oracle_vel("rm", x1, y1, s, r)
    ex, ey = O.rk_combine(x, y, [u0[0], u1[0]], [u0[1], u1[1]], [0.25, 0.75], 0.01)
    # positions stored as float: agreement to float rounding of the positions
    assert np.max(np.abs(got[: len(x)] - ex)) <= 1e-7 and np.max(np.abs(got[len(x):] - ey)) <= 1e-7


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


def test_bem_rk2_step_with_wake(tmp_path):
    """Ralston RK2 with the BEM solved at both stages and the inner layer cleared after each
    move (method 1, src/Convection.h:413,464), as src/Convection.h:390-466 does it."""
    w, nx, ny, idx = bem_case(2, npan=96, nw=1000)
    nw = len(w["x"])
    dt, ips = 0.01, 0.5  # a layer 0.1 thick: wake particles start from radius 0.52
    header = [len(idx) // 2, len(nx), nw, 2, CORE_ID["v2"], 2]
    payload = [np.array([1.0, 0.0, dt, ips]), nx, ny, idx, w["x"], w["y"], w["s"], w["r"]]
    got = run_driver(tmp_path, "bemstep", header, payload, 2 * nw)
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
    assert np.max(np.abs(got[:nw] - ex)) <= 1e-12 and np.max(np.abs(got[nw:] - ey)) <= 1e-12


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

