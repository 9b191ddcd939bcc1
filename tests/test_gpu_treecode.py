"""GPU treecode (summation_t barneshut, src/ExecEnv.h:26-31) against the direct sum.

The reference reserves the treecode slot but has no treecode of its own, so the direct sum is
the reference here. The small cases compare against the CPU oracle; the 1M-particle blob
workload compares against the GPU's fp64 direct sum on 4,096 targets. Bars follow the expansion
order: relative L2 <= 1e-7 at (theta 0.5, order 20) and <= 1e-10 at (theta 0.4, order 30), and
the error falls as the order grows.
"""
import numpy as np
import pytest

import oracle as O
import omega2d_amd as o2d
from helpers import rel_l2
from omega2d_amd import workloads as W

pytestmark = pytest.mark.gpu


def blobs(n, seed, sigma=1e-4, nb=8):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, (nb, 2))
    k = rng.integers(0, nb, n)
    x = c[k, 0] + 0.15 * rng.normal(size=n)
    y = c[k, 1] + 0.15 * rng.normal(size=n)
    return x, y, rng.normal(0, 1.0 / n, n), np.full(n, sigma)


def rel2(a, b):
    return rel_l2(np.concatenate(a), np.concatenate(b))


@pytest.mark.parametrize("particle_targets", [False, True])
def test_tree_matches_oracle_direct_sum(particle_targets):
    x, y, s, r = blobs(20000, 1)
    rng = np.random.default_rng(2)
    tx, ty = rng.uniform(-1.3, 1.3, 3000), rng.uniform(-1.3, 1.3, 3000)
    tr = np.full(3000, 1e-4) if particle_targets else None
    ref = O.points_on_points("gauss", x, y, s, r, tx, ty, tr)
    got = o2d.points_on_points_tree(x, y, s, r, tx, ty, tr, core="gauss", theta=0.5, order=20)
    assert rel2(got, ref) <= 1e-7
    got = o2d.points_on_points_tree(x, y, s, r, tx, ty, tr, core="gauss", theta=0.4, order=30)
    assert rel2(got, ref) <= 1e-10


def test_tree_error_falls_with_order():
    x, y, s, r = blobs(30000, 3)
    ref = o2d.points_on_points(x, y, s, r, x[:4000], y[:4000], r[:4000], core="gauss")
    errs = [rel2(o2d.points_on_points_tree(x, y, s, r, x[:4000], y[:4000], r[:4000],
                                           core="gauss", theta=0.5, order=p), ref)
            for p in (6, 12, 20, 28)]
    assert all(a > b for a, b in zip(errs, errs[1:])), errs
    assert errs[-1] <= 1e-9


@pytest.mark.parametrize("core", ["rm", "wl", "v2", "v3"])
def test_tree_other_cores_respect_core_tolerance(core):
    x, y, s, _ = blobs(12000, 4)
    r = np.full(len(x), 2e-3)
    rng = np.random.default_rng(5)
    tx, ty = rng.uniform(-1.3, 1.3, 2000), rng.uniform(-1.3, 1.3, 2000)
    ref = O.points_on_points(core, x, y, s, r, tx, ty)
    got = o2d.points_on_points_tree(x, y, s, r, tx, ty, core=core, theta=0.4, order=30,
                                    core_tol=1e-12)
    assert rel2(got, ref) <= 1e-9


@pytest.mark.parametrize("ns,nt", [(1, 1), (63, 5), (65, 300), (257, 1025), (4097, 4097)])
def test_tree_ragged_sizes(ns, nt):
    rng = np.random.default_rng(ns + nt)
    x, y = rng.uniform(-1, 1, ns), rng.uniform(-1, 1, ns)
    s, r = rng.normal(size=ns), np.full(ns, 0.01)
    tx, ty = rng.uniform(-1, 1, nt), rng.uniform(-1, 1, nt)
    ref = O.points_on_points("gauss", x, y, s, r, tx, ty)
    got = o2d.points_on_points_tree(x, y, s, r, tx, ty, core="gauss", theta=0.4, order=30)
    assert rel2(got, ref) <= 1e-10


def test_tree_accumulates_and_device_entry():
    import torch
    import omega2d_amd.device as dv
    x, y, s, r = blobs(10000, 6)
    base_u, base_v = np.ones(10000), -np.ones(10000)
    gu, gv = o2d.points_on_points_tree(x, y, s, r, x, y, r, theta=0.5, order=20,
                                       tu=base_u.copy(), tv=base_v.copy())
    hu, hv = o2d.points_on_points_tree(x, y, s, r, x, y, r, theta=0.5, order=20)
    assert np.allclose(gu - 1.0, hu, rtol=0, atol=1e-15) and np.allclose(gv + 1.0, hv, rtol=0,
                                                                        atol=1e-15)
    cu = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    du, dvv = dv.points_on_points_tree(cu(x), cu(y), cu(s), cu(r), cu(x), cu(y), cu(r),
                                       theta=0.5, order=20)
    assert np.array_equal(du.cpu().numpy(), hu) and np.array_equal(dvv.cpu().numpy(), hv)


def test_tree_argument_errors():
    x = np.zeros(4)
    for kw, code in (({"theta": 1.0}, -1), ({"order": 0}, -1), ({"order": 33}, -1)):
        with pytest.raises(o2d.O2DError) as e:
            o2d.points_on_points_tree(x, x, x, x + 0.1, x, x, **kw)
        assert e.value.code == code
    env = o2d.make_env("gauss", "fp64", "vort")
    import ctypes
    opt = o2d.TreeOpts(0.5, 20, 0, 1e-12)
    with pytest.raises(o2d.O2DError) as e:
        o2d.call("o2d_hip_points_on_points_tree_d", ctypes.byref(env), ctypes.byref(opt), 0, None,
                 None, None, None, 0, None, None, None, None, None)
    assert e.value.code == -4


def test_tree_on_the_1m_blob_workload():
    """The headline 1M-particle blob workload (Gaussian, sigma 7.8e-5): the treecode against the
    fp64 direct sum on 4,096 particle targets."""
    import torch
    import omega2d_amd.device as dv
    w = W.c2_blobs()
    cu = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    x, y, s, r = (cu(w[k]) for k in ("x", "y", "s", "r"))
    tu, tv = dv.points_on_points_tree(x, y, s, r, x, y, r, theta=0.5, order=20)
    m = 4096
    idx = torch.arange(0, x.numel(), x.numel() // m, device="cuda")[:m]
    ru, rv = dv.points_on_points(x, y, s, r, x[idx].contiguous(), y[idx].contiguous(),
                                 r[idx].contiguous(), core="gauss", precision="fp64")
    got = (tu[idx].cpu().numpy(), tv[idx].cpu().numpy())
    assert rel2(got, (ru.cpu().numpy(), rv.cpu().numpy())) <= 1e-7


def test_tree_split_walk_for_few_targets():
    """65,536 tracers are 1,024 target groups, too few waves for the chip, so each group's walk is
    split over several waves (each from a share of one level's cells) and the partial sums are
    added in a fixed order: the result is reproducible bit for bit, accumulates, and keeps the
    error bar against the fp64 direct sum."""
    import torch
    import omega2d_amd.device as dv
    w = W.c2_blobs()
    cu = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    x, y, s, r = (cu(w[k]) for k in ("x", "y", "s", "r"))
    tx, ty = cu(w["tx"]), cu(w["ty"])
    tu, tv = dv.points_on_points_tree(x, y, s, r, tx, ty, None, theta=0.6, order=28)
    u2, v2 = dv.points_on_points_tree(x, y, s, r, tx, ty, None, theta=0.6, order=28)
    assert torch.equal(tu, u2) and torch.equal(tv, v2)
    au, av = dv.points_on_points_tree(x, y, s, r, tx, ty, None, theta=0.6, order=28,
                                      tu=torch.ones_like(tu), tv=-torch.ones_like(tv))
    assert torch.allclose(au - 1.0, tu, rtol=1e-15, atol=1e-15)
    assert torch.allclose(av + 1.0, tv, rtol=1e-15, atol=1e-15)
    ru, rv = dv.points_on_points(x, y, s, r, tx, ty, None, core="gauss", precision="fp64")
    err = rel2((tu.cpu().numpy(), tv.cpu().numpy()), (ru.cpu().numpy(), rv.cpu().numpy()))
    print("split tree walk, 65,536 tracers:", err)
    assert err <= 1e-9
