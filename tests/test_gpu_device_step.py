"""The device-resident stepper (omega2d_amd/device_step.py) against the C++ host mirror and the
oracle.

The stepper makes the host mirror's calls (host/Convection.h, host/BEM.h, which restate
src/Convection.h:172-598 and src/BEMHelper.h:43-260) on device tensors, so whole steps must
agree with the mirror's to rounding, and free-particle steps with the oracle's Ralston schemes
to the bars of the precision forms.
"""
import numpy as np
import pytest
import torch

import oracle as O
from helpers import circle_panels, rel_l2
from omega2d_amd import device_step as ds
from omega2d_amd import workloads as W
from test_gpu_host_driver import CORE_ID, run_driver

pytestmark = pytest.mark.gpu


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


@pytest.mark.parametrize("nunk", [1, 2])
def test_rk2_step_with_body_matches_host_mirror(tmp_path, nunk):
    # one Ralston RK2 step of a wake around a 96-panel circle, fp64 pair sums
    w = W.c3_circle_wake(npan=96, nwake=1000)
    nx, ny, idx = circle_panels(96)
    nw = len(w["x"])
    dt, ips = 0.01, 0.5
    header = [96, len(nx), nw, nunk, CORE_ID["v2"], 2]
    payload = [np.array([1.0, 0.0, dt, ips]), nx, ny, idx, w["x"], w["y"], w["s"], w["r"]]
    got = run_driver(tmp_path, "bemstep", header, payload, 4 * nw)
    x, y, s, r = (T(w[k]) for k in ("x", "y", "s", "r"))
    body = ds.Body(nx, ny, idx, nunk)
    st = ds.Stepper(x, y, s, r, core="v2", precision="fp64", fs=(1.0, 0.0), body=body, ips=ips)
    u, v = st.step(dt, 2)
    gx, gy = x.cpu().numpy(), y.cpu().numpy()
    assert np.max(np.abs(gx - got[:nw])) <= 1e-12
    assert np.max(np.abs(gy - got[nw:2 * nw])) <= 1e-12
    assert rel_l2(u.cpu().numpy(), got[2 * nw:3 * nw]) <= 1e-12
    assert rel_l2(v.cpu().numpy(), got[3 * nw:]) <= 1e-12


@pytest.mark.parametrize("order,nunk,precision", [(2, 1, "split"), (3, 1, "fp64"),
                                                  (2, 2, "fp64")])
def test_several_steps_with_body_match_host_mirror(tmp_path, order, nunk, precision):
    # five steps (the circulation error of each solve feeds the next one's constraint row)
    w = W.c3_circle_wake(npan=128, nwake=1500)
    nx, ny, idx = circle_panels(128)
    nw = len(w["x"])
    dt, ips = 0.01, 0.05
    prec_id = 2 if precision == "split" else 1
    header = [128, len(nx), nw, nunk, CORE_ID["v2"], order, 2, prec_id]
    payload = [np.array([1.0, 0.0, dt, ips]), nx.astype(np.float32), ny.astype(np.float32),
               idx.astype(np.uint32), w["x"], w["y"], w["s"], w["r"]]
    got = run_driver(tmp_path, "bemtime", header, payload, 1 + 2 * nw)
    x, y, s, r = (T(w[k]) for k in ("x", "y", "s", "r"))
    body = ds.Body(nx, ny, idx, nunk)
    st = ds.Stepper(x, y, s, r, core="v2", precision=precision, fs=(1.0, 0.0), body=body,
                    ips=ips)
    for _ in range(5):
        st.step(dt, order)
    gx, gy = x.cpu().numpy(), y.cpu().numpy()
    assert np.max(np.abs(gx - got[1:1 + nw])) <= 1e-12
    assert np.max(np.abs(gy - got[1 + nw:])) <= 1e-12


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_free_particle_steps_match_oracle(order, precision):
    # the shear layer of the 4M-particle RK3 workload at 20,000 particles, Rosenhead-Moore
    w = W.c5_shear_layer(n=20000)
    x0, y0, s, r = w["x"], w["y"], w["s"], w["r"]
    dt = w["dt"]
    x, y = T(x0), T(y0)
    st = ds.Stepper(x, y, T(s), T(r), core="rm", precision=precision)
    u, v = st.step(dt, order)

    def vel(px, py):
        return O.finalize(*O.points_on_points("rm", px, py, s, r, px, py, r))
    u0 = vel(x0, y0)
    if order == 1:
        ex, ey = O.rk_combine(x0, y0, [u0[0]], [u0[1]], [1.0], dt)
        eu, ev = u0
    elif order == 2:
        x1, y1 = O.rk_combine(x0, y0, [u0[0]], [u0[1]], [1.0], 2.0 / 3.0 * dt)
        u1 = vel(x1, y1)
        ex, ey = O.rk_combine(x0, y0, [u0[0], u1[0]], [u0[1], u1[1]], [0.25, 0.75], dt)
        eu, ev = 0.25 * u0[0] + 0.75 * u1[0], 0.25 * u0[1] + 0.75 * u1[1]
    else:
        x1, y1 = O.rk_combine(x0, y0, [u0[0]], [u0[1]], [1.0], 0.5 * dt)
        u1 = vel(x1, y1)
        x2, y2 = O.rk_combine(x0, y0, [u1[0]], [u1[1]], [1.0], 0.75 * dt)
        u2 = vel(x2, y2)
        ex, ey = O.rk_combine(x0, y0, [u0[0], u1[0], u2[0]], [u0[1], u1[1], u2[1]],
                              [2 / 9, 3 / 9, 4 / 9], dt)
        eu = 2 / 9 * u0[0] + 3 / 9 * u1[0] + 4 / 9 * u2[0]
        ev = 2 / 9 * u0[1] + 3 / 9 * u1[1] + 4 / 9 * u2[1]
    bar = 1e-12 if precision == "fp64" else 1e-6
    assert rel_l2(u.cpu().numpy(), eu) <= bar and rel_l2(v.cpu().numpy(), ev) <= bar
    # positions: fp64 to rounding; the split form's displacement dt * u carries its velocity
    # error (bounded pointwise here by 1e-5 of the largest displacement)
    tol = 1e-14 if precision == "fp64" else 1e-5 * dt * np.max(np.hypot(eu, ev))
    assert np.max(np.abs(x.cpu().numpy() - ex)) <= tol
    assert np.max(np.abs(y.cpu().numpy() - ey)) <= tol
