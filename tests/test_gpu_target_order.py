"""The particle sums do not depend on the caller's order of sources or targets.

Both forms put sources and targets in Morton order internally and scatter results back to the
input order (src/Influence.h:56-416 loops over targets in input order and writes each target's
own slot). Permuting the inputs and permuting the results back must reproduce the unpermuted
results to the forms' bars: fp64 <= 1e-12 and split <= 1e-6 relative L2 (the summation order
within a tie of Morton keys follows the input order, so permuted runs may differ by rounding).
Every case is also checked against the oracle.
"""
import numpy as np
import pytest
import torch

import oracle as O
import omega2d_amd as o2d
import omega2d_amd.device  # noqa: F401  (binds o2d.device)
from omega2d_amd import device as dv
from helpers import rel_l2

pytestmark = pytest.mark.gpu

BAR = {"fp64": 1e-12, "split": 1e-6}


def _blobs(rng, n, nblob=8):
    """Particles in a few compact blobs (the layout where ordering matters) and varied radii."""
    c = rng.uniform(-2, 2, (nblob, 2))
    b = rng.integers(0, nblob, n)
    x = c[b, 0] + 0.1 * rng.normal(size=n)
    y = c[b, 1] + 0.1 * rng.normal(size=n)
    s = rng.normal(0, 1e-3, n)
    r = rng.uniform(0.005, 0.02, n)
    return x, y, s, r


@pytest.mark.parametrize("precision", ["fp64", "split"])
@pytest.mark.parametrize("core", ["rm", "gauss", "v2"])
@pytest.mark.parametrize("results", ["vel", "vort"])
def test_self_interaction_is_order_independent(precision, core, results):
    # 40,000 particles on themselves: several source chunks, so the chunk reduction scatters
    rng = np.random.default_rng(11)
    x, y, s, r = _blobs(rng, 40000)
    base = o2d.points_on_points(x, y, s, r, x, y, r, core=core, precision=precision,
                                results=results)
    p = rng.permutation(len(x))
    perm = o2d.points_on_points(x[p], y[p], s[p], r[p], x[p], y[p], r[p], core=core,
                                precision=precision, results=results)
    sub = np.arange(0, len(x), 97)
    ref = O.points_on_points(core, x, y, s, r, x[sub], y[sub], r[sub], results)
    for k, (a, b) in enumerate(zip(base, perm)):
        back = np.empty_like(b)
        back[p] = b
        assert rel_l2(back, a) <= BAR[precision], (k, rel_l2(back, a))
        assert rel_l2(a[sub], ref[k]) <= BAR[precision], (k, rel_l2(a[sub], ref[k]))


@pytest.mark.parametrize("precision", ["fp64", "split"])
@pytest.mark.parametrize("point_targets", [False, True])
@pytest.mark.parametrize("ns,nt", [(200, 5000), (30000, 3000)])
def test_separate_targets_are_order_independent(precision, point_targets, ns, nt):
    # (200, 5000): one source chunk, results written straight through the order;
    # (30000, 3000): chunk partials reduced and scattered through it
    rng = np.random.default_rng(ns + nt)
    sx, sy, ss, sr = _blobs(rng, ns)
    tx, ty, _, tr = _blobs(rng, nt)
    tr = None if point_targets else tr
    base = o2d.points_on_points(sx, sy, ss, sr, tx, ty, tr, core="gauss", precision=precision)
    ps, pt = rng.permutation(ns), rng.permutation(nt)
    perm = o2d.points_on_points(sx[ps], sy[ps], ss[ps], sr[ps], tx[pt], ty[pt],
                                None if tr is None else tr[pt], core="gauss",
                                precision=precision)
    ref = O.points_on_points("gauss", sx, sy, ss, sr, tx, ty, tr)
    for a, b, c in zip(base, perm, ref):
        back = np.empty_like(b)
        back[pt] = b
        assert rel_l2(back, a) <= BAR[precision]
        assert rel_l2(a, c) <= BAR[precision]


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_targets_sharing_the_sources_storage(precision):
    # targets that are a prefix view of the source arrays (same device address, fewer points)
    # must not take the sources' order, which indexes all of them
    rng = np.random.default_rng(5)
    x, y, s, r = _blobs(rng, 6000)
    T = lambda a: torch.from_numpy(a).cuda()  # noqa: E731
    gx, gy, gs, gr = T(x), T(y), T(s), T(r)
    m = 2500
    u, v = dv.points_on_points(gx, gy, gs, gr, gx[:m], gy[:m], gr[:m], core="rm",
                               precision=precision)
    torch.cuda.synchronize()
    ru, rv = O.points_on_points("rm", x, y, s, r, x[:m], y[:m], r[:m])
    assert rel_l2(u.cpu().numpy(), ru) <= BAR[precision]
    assert rel_l2(v.cpu().numpy(), rv) <= BAR[precision]


def test_accumulates_in_input_order():
    # outputs are accumulated (+=) into each target's own input slot
    rng = np.random.default_rng(9)
    x, y, s, r = _blobs(rng, 3000)
    u0 = rng.normal(size=3000)
    v0 = rng.normal(size=3000)
    for precision in ("fp64", "split"):
        tu, tv = u0.copy(), v0.copy()
        o2d.points_on_points(x, y, s, r, x, y, r, core="v2", precision=precision, tu=tu, tv=tv)
        ru, rv = O.points_on_points("v2", x, y, s, r, x, y, r)
        assert rel_l2(tu - u0, ru) <= BAR[precision]
        assert rel_l2(tv - v0, rv) <= BAR[precision]
