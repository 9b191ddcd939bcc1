"""Parity at the full sizes of the five benchmark configurations, through properties that do not
depend on size, plus oracle checks on target (or panel) subsets of the full sums:

- C2 (1M Gaussian particles + 65,536 tracers): split against fp64 over every target; fp64
  against the oracle on a target subset; exact linearity in the strengths; momentum balance of
  the self-interaction (the 0b kernel is symmetric, so sum_i G_i u_i = 0).
- C1 (10,000-particle disc, RM): one Ralston RK2 step, split and fp64, against the oracle's
  step at full size.
- C5 (4M particles, RM): split against fp64 on a 16,384-target subsample (the RK3 step itself
  is checked against the oracle at 20,000 particles in test_gpu_device_step.py).
- C3 (2,048 panels + 200,000 particles): the 2049^2 and 4097^2 systems against LAPACK, the
  2049^2 matrix against the oracle; panel sums against the oracle on panel/target subsets; the
  wake's V2 self-interaction, split against fp64 on every particle and fp64 against the oracle
  on a subset.
- C4 (250,000 particles, 4,096 probes): the strength adjoint is the transposed sum,
  <dJ/dG, G> = <gu, u> + <gv, v>.
"""
import numpy as np
import pytest

import oracle as O
from helpers import rel_l2
from omega2d_amd import workloads as W

pytestmark = pytest.mark.gpu


def _cu(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _np(t):
    return t.cpu().numpy()


def _err2(got, ref):
    gu, gv = got
    ru, rv = ref
    return float(np.sqrt((np.sum((gu - ru) ** 2) + np.sum((gv - rv) ** 2))
                         / (np.sum(ru ** 2) + np.sum(rv ** 2))))


@pytest.fixture(scope="module")
def c2():
    w = W.c2_blobs()
    return w, {k: _cu(w[k]) for k in ("x", "y", "s", "r", "tx", "ty")}


def test_c2_split_against_fp64_on_every_target(c2):
    import omega2d_amd.device as dv
    w, d = c2
    out = {}
    for prec in ("split", "fp64"):
        pu, pv = dv.points_on_points(d["x"], d["y"], d["s"], d["r"], d["x"], d["y"], d["r"],
                                     core="gauss", precision=prec)
        tu, tv = dv.points_on_points(d["x"], d["y"], d["s"], d["r"], d["tx"], d["ty"], None,
                                     core="gauss", precision=prec)
        out[prec] = (np.concatenate([_np(pu), _np(tu)]), np.concatenate([_np(pv), _np(tv)]))
    err = _err2(out["split"], out["fp64"])
    print("C2 split vs fp64 over 1,114,112 targets:", err)
    assert err <= 1e-6
    # the fp64 form against the oracle on 96 particle targets and 32 tracers
    n = len(w["x"])
    pi = np.linspace(0, n - 1, 96).astype(np.int64)
    ti = np.linspace(0, len(w["tx"]) - 1, 32).astype(np.int64)
    ru, rv = O.points_on_points("gauss", w["x"], w["y"], w["s"], w["r"], w["x"][pi], w["y"][pi],
                                w["r"][pi])
    su, sv = O.points_on_points("gauss", w["x"], w["y"], w["s"], w["r"], w["tx"][ti],
                                w["ty"][ti])
    fu, fv = out["fp64"]
    got = (np.concatenate([fu[pi], fu[n + ti]]), np.concatenate([fv[pi], fv[n + ti]]))
    assert _err2(got, (np.concatenate([ru, su]), np.concatenate([rv, sv]))) <= 1e-12


@pytest.mark.parametrize("prec", ["split", "fp64"])
def test_c2_linear_in_strength_exactly(c2, prec):
    """Doubling every strength doubles every velocity bit for bit: scaling by 2 is exact in
    every rounding of both forms."""
    import omega2d_amd.device as dv
    _, d = c2
    u1, v1 = dv.points_on_points(d["x"], d["y"], d["s"], d["r"], d["tx"], d["ty"], None,
                                 core="gauss", precision=prec)
    u2, v2 = dv.points_on_points(d["x"], d["y"], 2.0 * d["s"], d["r"], d["tx"], d["ty"], None,
                                 core="gauss", precision=prec)
    assert np.array_equal(_np(u2), 2.0 * _np(u1)) and np.array_equal(_np(v2), 2.0 * _np(v1))


@pytest.mark.parametrize("prec,bound", [("fp64", 1e-12), ("split", 1e-6)])
def test_c2_self_interaction_conserves_momentum(c2, prec, bound):
    """K(d, sigma_i, sigma_j) is symmetric in the two particles, so the pair terms of
    sum_i G_i (u_i, v_i) cancel: the sum vanishes against sum_i |G_i| |(u_i, v_i)|."""
    import omega2d_amd.device as dv
    w, d = c2
    u, v = dv.points_on_points(d["x"], d["y"], d["s"], d["r"], d["x"], d["y"], d["r"],
                               core="gauss", precision=prec)
    u, v, s = _np(u), _np(v), w["s"]
    scale = np.sum(np.abs(s) * np.hypot(u, v))
    print(prec, abs(np.dot(s, u)) / scale, abs(np.dot(s, v)) / scale)
    assert abs(np.dot(s, u)) <= bound * scale and abs(np.dot(s, v)) <= bound * scale


def test_c5_split_against_fp64_on_the_16k_subsample():
    import omega2d_amd.device as dv
    w = W.c5_shear_layer()
    d = {k: _cu(w[k]) for k in ("x", "y", "s", "r")}
    rng = np.random.default_rng(4)
    sub = np.sort(rng.choice(len(w["x"]), 16384, replace=False))
    tx, ty, tr = (_cu(w[k][sub]) for k in ("x", "y", "r"))
    res = {}
    for prec in ("split", "fp64"):
        u, v = dv.points_on_points(d["x"], d["y"], d["s"], d["r"], tx, ty, tr, core="rm",
                                   precision=prec)
        res[prec] = (_np(u), _np(v))
    err = _err2(res["split"], res["fp64"])
    print("C5 split vs fp64 on 16,384 targets:", err)
    assert err <= 1e-6
    # and the fp64 form against the oracle on 64 of them
    ru, rv = O.points_on_points("rm", w["x"], w["y"], w["s"], w["r"], w["x"][sub[:64]],
                                w["y"][sub[:64]], w["r"][sub[:64]])
    assert _err2((res["fp64"][0][:64], res["fp64"][1][:64]), (ru, rv)) <= 1e-12


@pytest.mark.parametrize("nunk", [1, 2])
def test_c3_full_systems_against_lapack(nunk):
    import torch
    import omega2d_amd.device as dv
    w = W.c3_circle_wake()
    npan = len(w["idx"]) // 2
    vol, _, _ = O.body_volume(w["nx"], w["ny"], w["idx"])
    rng = np.random.default_rng(nunk)
    uu, uv = rng.normal(size=npan), rng.normal(size=npan)
    n = npan * nunk + 1
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    idx = torch.from_numpy(w["idx"].astype(np.int32)).cuda()
    dv.bem_assemble(_cu(w["nx"]), _cu(w["ny"]), idx, nunk, True, _cu(uu), _cu(uv), vol, out=A)
    Ah = _np(A).T.copy()  # the device matrix is column-major
    if nunk == 1:
        ref = O.bem_assemble(w["nx"], w["ny"], w["idx"], 1, True, uu, uv, vol)
        assert rel_l2(Ah, ref) <= 1e-12 and np.max(np.abs(Ah - ref)) <= 1e-13
    dv.bem_factor(A)
    fro = np.linalg.norm(Ah)
    for k in range(5):  # the fresh factors, then the inverse path (from the second solve)
        b = rng.normal(size=n)
        x = _np(dv.bem_solve(_cu(b)))
        # two backward-stable fp64 solves agree to about cond(A) eps; cond(A) = 9.3e3 here, so
        # the bar is 1e-11 rather than 1e-12, and the residual (backward error) is checked too
        assert rel_l2(x, np.linalg.solve(Ah, b)) <= 1e-11, k
        assert np.linalg.norm(Ah @ x - b) <= 1e-13 * fro * np.linalg.norm(x), k


def test_c3_panel_sums_against_the_oracle_on_subsets():
    import torch
    import omega2d_amd.device as dv
    w = W.c3_circle_wake()
    npan = len(w["idx"]) // 2
    rng = np.random.default_rng(7)
    vs, ss = rng.normal(size=npan), rng.normal(size=npan)
    idx = torch.from_numpy(w["idx"].astype(np.int32)).cuda()
    nx, ny = _cu(w["nx"]), _cu(w["ny"])
    x, y, s = _cu(w["x"]), _cu(w["y"]), _cu(w["s"])
    # every panel on every particle; the oracle on 1,000 of the targets
    u, v = dv.panels_on_points(nx, ny, idx, _cu(vs), _cu(ss), x, y)
    t = np.linspace(0, len(w["x"]) - 1, 1000).astype(np.int64)
    ru, rv = O.panels_on_points(w["nx"], w["ny"], w["idx"], vs, ss, w["x"][t], w["y"][t])
    assert rel_l2(_np(u)[t], ru) <= 1e-12 and rel_l2(_np(v)[t], rv) <= 1e-12
    # every particle on every panel (the BEM right-hand side); the oracle on 64 panels, whose
    # sums do not depend on the other panels
    pu, pv = dv.points_on_panels(x, y, s, nx, ny, idx)
    sel = np.linspace(0, npan - 1, 64).astype(np.int64)
    sidx = w["idx"].reshape(-1, 2)[sel].ravel()
    qu, qv = O.points_on_panels(w["x"], w["y"], w["s"], w["nx"], w["ny"], sidx)
    assert rel_l2(_np(pu)[sel], qu) <= 1e-12 and rel_l2(_np(pv)[sel], qv) <= 1e-12


def test_c4_strength_adjoint_is_the_transposed_sum():
    import omega2d_amd.device as dv
    w = W.c4_adjoint()
    d = {k: _cu(w[k]) for k in ("x", "y", "s", "r", "tx", "ty")}
    u, v = dv.points_on_points(d["x"], d["y"], d["s"], d["r"], d["tx"], d["ty"], None,
                               core="gauss", precision="fp64")
    rng = np.random.default_rng(5)
    gu, gv = rng.normal(size=len(w["tx"])), rng.normal(size=len(w["tx"]))
    gx, gy, gs, gr = dv.points_on_points_adjoint(d["x"], d["y"], d["s"], d["r"], d["tx"],
                                                 d["ty"], _cu(gu), _cu(gv), core="gauss")
    lhs = float(np.dot(_np(gs), w["s"]))
    rhs = float(np.dot(gu, _np(u)) + np.dot(gv, _np(v)))
    scale = float(np.dot(np.abs(_np(gs)), np.abs(w["s"])))
    print("C4 <dJ/dG, G> vs <g, u>:", lhs, rhs)
    assert abs(lhs - rhs) <= 1e-12 * scale
    assert np.all(np.isfinite(_np(gx))) and np.all(np.isfinite(_np(gy)))


@pytest.mark.parametrize("precision,bar", [("fp64", 1e-12), ("split", 1e-6)])
def test_c1_rk2_step_at_full_size_against_the_oracle(precision, bar):
    import omega2d_amd.device_step as ds
    w = W.c1_disc()
    x0, y0, s, r, dt = w["x"], w["y"], w["s"], w["r"], w["dt"]
    x, y = _cu(x0), _cu(y0)
    st = ds.Stepper(x, y, _cu(s), _cu(r), core="rm", precision=precision)
    u, v = st.step(dt, 2)

    def vel(px, py):
        return O.finalize(*O.points_on_points("rm", px, py, s, r, px, py, r))
    u0 = vel(x0, y0)
    x1, y1 = O.rk_combine(x0, y0, [u0[0]], [u0[1]], [1.0], 2.0 / 3.0 * dt)
    u1 = vel(x1, y1)
    ex, ey = O.rk_combine(x0, y0, [u0[0], u1[0]], [u0[1], u1[1]], [0.25, 0.75], dt)
    eu, ev = 0.25 * u0[0] + 0.75 * u1[0], 0.25 * u0[1] + 0.75 * u1[1]
    assert _err2((_np(u), _np(v)), (eu, ev)) <= bar
    tol = 1e-14 if precision == "fp64" else 1e-5 * dt * np.max(np.hypot(eu, ev))
    assert np.max(np.abs(_np(x) - ex)) <= tol and np.max(np.abs(_np(y) - ey)) <= tol


def test_c3_wake_split_against_fp64_and_the_oracle():
    import omega2d_amd.device as dv
    w = W.c3_circle_wake()
    d = {k: _cu(w[k]) for k in ("x", "y", "s", "r")}
    res = {}
    for prec in ("split", "fp64"):
        u, v = dv.points_on_points(d["x"], d["y"], d["s"], d["r"], d["x"], d["y"], d["r"],
                                   core="v2", precision=prec)
        res[prec] = (_np(u), _np(v))
    err = _err2(res["split"], res["fp64"])
    print("C3 wake split vs fp64 over 200,000 particles:", err)
    assert err <= 1e-6
    t = np.linspace(0, len(w["x"]) - 1, 256).astype(np.int64)
    ru, rv = O.points_on_points("v2", w["x"], w["y"], w["s"], w["r"], w["x"][t], w["y"][t],
                                w["r"][t])
    assert _err2((res["fp64"][0][t], res["fp64"][1][t]), (ru, rv)) <= 1e-12
