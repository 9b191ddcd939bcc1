"""Adjoint harness on the GPU: torch autograd through the hand-written adjoint kernel, checked
against fp64 central finite differences (h = 1e-6) and the fp64 CPU oracle gradient, on a
2,000-particle subset of the 250k-particle / 4,096-probe adjoint workload."""
import numpy as np
import pytest

import oracle as O
from helpers import rel_l2
from omega2d_amd import workloads as W

pytestmark = pytest.mark.gpu


def _objective(core, x, y, s, r, px, py):
    u, v = O.points_on_points(core, x, y, s, r, px, py)
    u, v = O.finalize(u, v)
    return float(np.sum(u * u + v * v))


@pytest.mark.parametrize("inside", [False, True])
def test_probe_objective_gradients(inside):
    import torch
    from omega2d_amd.autograd import probe_velocity

    w = W.c4_adjoint()
    n = 2000
    x, y, s, r = (w[k][:n].copy() for k in ("x", "y", "s", "r"))
    px, py = w["tx"].copy(), w["ty"].copy()
    if inside:
        # probes inside the cloud, where the Gaussian core is not beyond its cutoff for every
        # pair, so dJ/dsigma is not identically zero (the config's probes at r = 1.5 are)
        px, py = 0.6 * px, 0.6 * py
    T = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda", requires_grad=True)
    tx, ty, ts, trr = T(x), T(y), T(s), T(r)
    tpx = torch.tensor(px, dtype=torch.float64, device="cuda")
    tpy = torch.tensor(py, dtype=torch.float64, device="cuda")
    u, v = probe_velocity(tx, ty, ts, trr, tpx, tpy, core="gauss")
    J = (u * u + v * v).sum()
    J.backward()
    grads = [a.grad.cpu().numpy() for a in (tx, ty, ts, trr)]

    # the CPU oracle gradient of the same objective: dJ/dtu = 2 u / (2 pi)^2 on unscaled sums
    ou, ov = O.points_on_points("gauss", x, y, s, r, px, py)
    c = 0.5 / np.pi
    ref = O.points_on_points_adjoint("gauss", x, y, s, r, px, py, 2 * c * c * ou, 2 * c * c * ov)
    assert abs(J.item() - _objective("gauss", x, y, s, r, px, py)) <= 1e-12 * J.item()
    for g, rg in zip(grads, ref):
        if np.linalg.norm(rg) == 0:
            assert np.all(g == 0)
        else:
            assert rel_l2(g, rg) <= 1e-11
    if not inside:
        assert np.all(grads[3] == 0)  # every pair is beyond x = 16: K = 1/d^2, no sigma dependence

    # fp64 central differences on a sample of particles and every parameter
    rng = np.random.default_rng(0)
    h = 1e-6
    for j in rng.choice(n, 6, replace=False):
        for k, arr in enumerate((x, y, s, r)):
            if k == 3 and not inside:
                continue
            p, m = arr.copy(), arr.copy()
            p[j] += h
            m[j] -= h
            args_p = [p if a is arr else a for a in (x, y, s, r)]
            args_m = [m if a is arr else a for a in (x, y, s, r)]
            fd = (_objective("gauss", *args_p, px, py) - _objective("gauss", *args_m, px, py)) / (2 * h)
            scale = max(abs(fd), 1e-3 * np.max(np.abs(grads[k])))
            assert abs(grads[k][j] - fd) <= 1e-5 * scale, (k, j, grads[k][j], fd)


@pytest.mark.parametrize("core", ["rm", "v2"])
def test_padding_probes_leave_a_source_at_the_origin_finite(core):
    # 300 probes (not a whole tile of 256): the padded slots repeat a real probe with zero
    # upstream gradient. A zero-radius source at the origin, where a padding probe placed at
    # (0, 0) would coincide with it, must keep finite gradients equal to the oracle's.
    import omega2d_amd as o2d
    rng = np.random.default_rng(3)
    x = np.concatenate([[0.0], rng.uniform(-1, 1, 99)])
    y = np.concatenate([[0.0], rng.uniform(-1, 1, 99)])
    s = rng.normal(0, 1, 100)
    r = np.concatenate([[0.0], rng.uniform(0.01, 0.05, 99)])
    th = np.linspace(0, 2 * np.pi, 300, endpoint=False)
    px, py = 2.0 * np.cos(th), 2.0 * np.sin(th)
    gu, gv = rng.normal(size=300), rng.normal(size=300)
    got = o2d.points_on_points_adjoint(x, y, s, r, px, py, gu, gv, core=core)
    ref = O.points_on_points_adjoint(core, x, y, s, r, px, py, gu, gv)
    for g, rg in zip(got, ref):
        assert np.all(np.isfinite(g))
        assert rel_l2(g, rg) <= 1e-12
