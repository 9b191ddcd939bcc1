"""GPU parity of the Gaussian vorticity sum (src/InfluenceVort.h:47-126) against the oracle:
relative L2 <= 1e-12 in fp64, at ragged sizes, with few targets (chunked sources), with tiny
blobs whose far pairs underflow, and through the device entry point."""
import numpy as np
import pytest

import oracle as O
import omega2d_amd as o2d
from helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ns,nt", [(1, 1), (257, 513), (3000, 7), (9000, 1100), (20000, 3001)])
def test_vorticity_matches_oracle(ns, nt):
    rng = np.random.default_rng(ns + nt)
    sx, sy = rng.uniform(-1, 1, ns), rng.uniform(-1, 1, ns)
    ss, sr = rng.normal(0, 1, ns), rng.uniform(0.02, 0.2, ns)
    tx, ty = rng.uniform(-1.2, 1.2, nt), rng.uniform(-1.2, 1.2, nt)
    got = o2d.points_vorticity(sx, sy, ss, sr, tx, ty)
    ref = O.points_vorticity(sx, sy, ss, sr, tx, ty)
    assert rel_l2(got, ref) <= 1e-12


def test_vorticity_tiny_blobs_and_self():
    """Blob radii far below the spacing: almost every exponential underflows and is skipped."""
    rng = np.random.default_rng(5)
    n = 30000
    x, y = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    s, r = rng.normal(0, 1, n), np.full(n, 2e-3)
    got = o2d.points_vorticity(x, y, s, r, x, y)
    ref = O.points_vorticity(x, y, s, r, x, y)
    assert rel_l2(got, ref) <= 1e-12


def test_vorticity_accumulates_and_device_entry():
    import torch
    import omega2d_amd.device as dv
    rng = np.random.default_rng(6)
    ns, nt = 5000, 2000
    sx, sy = rng.uniform(-1, 1, ns), rng.uniform(-1, 1, ns)
    ss, sr = rng.normal(0, 1, ns), rng.uniform(0.05, 0.1, ns)
    tx, ty = rng.uniform(-1, 1, nt), rng.uniform(-1, 1, nt)
    base = rng.normal(size=nt)
    got = o2d.points_vorticity(sx, sy, ss, sr, tx, ty, tw=base.copy())
    ref = base + O.points_vorticity(sx, sy, ss, sr, tx, ty)
    assert rel_l2(got, ref) <= 1e-12
    cu = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")  # noqa: E731
    tw = dv.points_vorticity(cu(sx), cu(sy), cu(ss), cu(sr), cu(tx), cu(ty))
    assert np.array_equal(tw.cpu().numpy(), o2d.points_vorticity(sx, sy, ss, sr, tx, ty))
    empty = np.zeros(0)
    assert np.all(o2d.points_vorticity(empty, empty, empty, empty, tx, ty) == 0.0)


def _wake(rng, n, r):
    """particles along a few thin sheets, as a wake sheds them, with small blobs"""
    k = rng.integers(0, 4, n)
    t = rng.uniform(0, 6, n)
    x = t + rng.normal(0, 0.01, n)
    y = 0.3 * k + 0.05 * np.sin(3 * t) + rng.normal(0, 0.01, n)
    return x, y, rng.normal(0, 1e-3, n), np.full(n, r)


def test_vorticity_pruned_tiles_per_target():
    """Sources and targets in Morton order with whole source tiles skipped beyond the underflow
    distance: every target within 1e-12 of the oracle's sum scaled by its largest term, a
    target group far from every source exactly 0, and the result independent of the input
    order of sources and targets up to rounding."""
    rng = np.random.default_rng(11)
    ns, nt = 24000, 5000
    sx, sy, ss, sr = _wake(rng, ns, 0.004)
    tx, ty = np.concatenate([sx[:nt - 600], rng.uniform(20, 21, 600)]), \
        np.concatenate([sy[:nt - 600], rng.uniform(20, 21, 600)])
    got = o2d.points_vorticity(sx, sy, ss, sr, tx, ty)
    ref = O.points_vorticity(sx, sy, ss, sr, tx, ty)
    scale = 2.0 * np.max(np.abs(ss)) / 0.004 ** 2
    assert np.max(np.abs(got - ref)) <= 1e-12 * scale
    assert np.all(got[nt - 600:] == 0.0)
    ps, pt = rng.permutation(ns), rng.permutation(nt)
    got2 = o2d.points_vorticity(sx[ps], sy[ps], ss[ps], sr[ps], tx[pt], ty[pt])
    assert np.max(np.abs(got2 - got[pt])) <= 1e-13 * scale


def test_vorticity_pruned_nan_positions():
    """A NaN target position gives NaN at that target only; a NaN source position or radius
    gives NaN at every target, as the reference's exp(NaN) does; the pruned path keeps both."""
    rng = np.random.default_rng(12)
    ns, nt = 9000, 3000
    sx, sy, ss, sr = _wake(rng, ns, 0.004)
    tx, ty = sx[:nt].copy(), sy[:nt].copy()
    ty[1234] = np.nan
    got = o2d.points_vorticity(sx, sy, ss, sr, tx, ty)
    ref = O.points_vorticity(sx, sy, ss, sr, tx, ty)
    assert np.array_equal(np.isnan(got), np.isnan(ref)) and np.sum(np.isnan(got)) == 1
    ok = ~np.isnan(ref)
    assert np.max(np.abs(got[ok] - ref[ok])) <= 1e-12 * 2.0 * np.max(np.abs(ss)) / 0.004 ** 2
    for arr in (sx, sr):
        bad = arr.copy()
        bad[777] = np.nan
        a = (bad, sy, ss, sr) if arr is sx else (sx, sy, ss, bad)
        assert np.all(np.isnan(o2d.points_vorticity(*a, tx, ty)))


def test_vorticity_exponential_single_pair():
    """One source (s = 1/2, r = 1) on targets at d^2 = x over [0, 746): each sum is the one
    term exp(-x), against numpy's exp within 3 ulp (within 3 units of 2^-1074 where the value
    is subnormal), and exactly 0 at and beyond the underflow bound."""
    x = np.concatenate([np.linspace(0.0, 745.99, 200000), [745.999999, 746.0, 800.0]])
    tx = np.sqrt(x)
    d2 = tx * tx
    got = o2d.points_vorticity([0.0], [0.0], [0.5], [1.0], tx, np.zeros_like(tx))
    ref = np.exp(-d2)
    ulp = np.spacing(ref)
    err = np.abs(got - ref)
    normal = ref >= np.finfo(np.float64).tiny
    assert np.all(err[normal] <= 3.0 * ulp[normal])
    assert np.all(err[~normal] <= 3.0 * 2.0 ** -1074)
    assert np.all(got[d2 >= 746.0] == 0.0) and got[0] == 1.0
