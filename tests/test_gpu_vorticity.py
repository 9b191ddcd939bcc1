"""GPU parity of the Gaussian vorticity sum (src/InfluenceVort.h:47-126) against the oracle:
relative L2 <= 1e-12 in fp64, at ragged sizes, with few targets (chunked sources), with tiny
blobs whose far pairs underflow, and through the device entry point."""
import numpy as np
import pytest

import oracle as O
import omega2d_amd as o2d
from helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ns,nt", [(1, 1), (257, 513), (3000, 7), (20000, 3001)])
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
