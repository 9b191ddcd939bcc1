"""GPU parity of the particle/panel proximity kernels (src/Reflect.h) against the oracle.

The kernels and the oracle both compute in IEEE double without fused multiply-adds in the
reference's operation order, so nearest distances, reflected and pushed positions, trimmed
strengths and move counts are compared bit for bit. The removed circulation is a sum over
targets in a different order: relative 1e-13.
"""
import os

import numpy as np
import pytest

import oracle as O
import omega2d_amd as o2d
from helpers import circle_panels

pytestmark = pytest.mark.gpu

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))
CM, IPS = 0.5 / np.sqrt(2.0 * np.pi), 0.05


def _circle(npan):
    nx, ny, idx = circle_panels(npan)
    nx, ny = nx.astype(np.float64), ny.astype(np.float64)
    _, _, pnx, pny, _ = O.panel_geometry(nx, ny, idx)
    return nx, ny, idx, pnx, pny


def _spur():
    g = GOLD
    return (g["rt_nx"].astype(np.float64), g["rt_ny"].astype(np.float64), g["rt_idx"],
            g["rt_pnx"], g["rt_pny"])


def _targets(rng, n, rmax=0.8):
    rho, a = rng.uniform(0.0, rmax, n), rng.uniform(0, 2 * np.pi, n)
    return rho * np.cos(a), rho * np.sin(a)


GEOMS = {"circle48": lambda: _circle(48), "circle129": lambda: _circle(129),
         "spur": _spur, "one_panel": lambda: (np.array([-0.5, 0.5]), np.array([0.0, 0.0]),
                                              np.array([0, 1], np.uint32), np.array([0.0]),
                                              np.array([1.0]))}


@pytest.mark.parametrize("geom", list(GEOMS))
@pytest.mark.parametrize("nt", [1, 255, 257, 3000])
def test_nearest_distances_bit_exact(geom, nt):
    nx, ny, idx, _, _ = GEOMS[geom]()
    tx, ty = _targets(np.random.default_rng(nt), nt)
    got = o2d.nearest_distances(nx, ny, idx, tx, ty)
    ref = O.nearest_distances(nx, ny, idx, tx, ty)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("geom", list(GEOMS))
def test_reflect_bit_exact(geom):
    nx, ny, idx, pnx, pny = GEOMS[geom]()
    tx, ty = _targets(np.random.default_rng(11), 2500)
    gx, gy, gn = o2d.reflect(nx, ny, idx, pnx, pny, tx, ty)
    rx, ry, rn = O.reflect(nx, ny, idx, pnx, pny, tx, ty)
    assert gn == rn
    assert np.array_equal(gx, rx) and np.array_equal(gy, ry)


@pytest.mark.parametrize("geom", list(GEOMS))
@pytest.mark.parametrize("kind", ["particles", "field"])
def test_clear_push_bit_exact(geom, kind):
    nx, ny, idx, pnx, pny = GEOMS[geom]()
    rng = np.random.default_rng(12)
    tx, ty = _targets(rng, 2500)
    ts, tr = (rng.normal(0, 1e-3, 2500), rng.uniform(0.015, 0.025, 2500)) \
        if kind == "particles" else (None, None)
    got = o2d.clear_inner(1, nx, ny, idx, pnx, pny, CM, IPS, tx, ty, ts, tr)
    ref = O.clear_inner(1, nx, ny, idx, pnx, pny, CM, IPS, tx, ty, ts, tr)
    assert ref["unsettled"] == 0
    assert got["num_cropped"] == ref["num_cropped"]
    assert np.array_equal(got["x"], ref["x"]) and np.array_equal(got["y"], ref["y"])
    assert got["circ_removed"] == 0.0


@pytest.mark.parametrize("shell", ["sh", "dp"])
def test_clear_trim_bit_exact(shell):
    g = GOLD
    nx, ny, idx, pnx, pny = (g["rf_nx"].astype(np.float64), g["rf_ny"].astype(np.float64),
                             g["rf_idx"], g["rf_pnx"], g["rf_pny"])
    args = (g[f"rf_c0{shell}_tx"], g[f"rf_c0{shell}_ty"], g[f"rf_c0{shell}_ts"],
            g[f"rf_c0{shell}_tr"])
    got = o2d.clear_inner(0, nx, ny, idx, pnx, pny, CM, IPS, *args)
    ref = O.clear_inner(0, nx, ny, idx, pnx, pny, CM, IPS, *args)
    assert got["num_cropped"] == ref["num_cropped"]
    for k in ("x", "y", "str"):
        assert np.array_equal(got[k], ref[k])
    assert abs(got["circ_removed"] - ref["circ_removed"]) <= 1e-13 * abs(ref["circ_removed"])


def test_clear_at_bem_wake_scale():
    """The C3 body (2048 panels) against 100,000 targets in and around it, method 1."""
    nx, ny, idx, pnx, pny = _circle(2048)
    rng = np.random.default_rng(33)
    tx, ty = _targets(rng, 100_000, rmax=1.0)
    got = o2d.clear_inner(1, nx, ny, idx, pnx, pny, CM, IPS, tx, ty)
    ref = O.clear_inner(1, nx, ny, idx, pnx, pny, CM, IPS, tx, ty)
    assert got["num_cropped"] == ref["num_cropped"] > 0
    assert np.array_equal(got["x"], ref["x"]) and np.array_equal(got["y"], ref["y"])
    d = o2d.nearest_distances(nx, ny, idx, tx, ty)
    assert np.array_equal(d, O.nearest_distances(nx, ny, idx, tx, ty))


def test_empty_inputs():
    nx, ny, idx, pnx, pny = _circle(16)
    z = np.zeros(0)
    assert o2d.nearest_distances(nx, ny, idx, z, z).shape == (0,)
    d = o2d.nearest_distances(z, z, np.zeros(0, np.uint32), np.ones(3), np.ones(3))
    assert np.all(d == np.sqrt(np.finfo(np.float64).max))
    x, y, n = o2d.reflect(z, z, np.zeros(0, np.uint32), z, z, np.ones(3), np.ones(3))
    assert n == 0 and np.all(x == 1.0)
    r = o2d.clear_inner(1, nx, ny, idx, pnx, pny, CM, IPS, z, z)
    assert r["num_cropped"] == 0 and r["circ_removed"] == 0.0


def test_host_entry_errors():
    nx, ny, idx, pnx, pny = _circle(16)
    t = np.array([0.0]), np.array([0.0])
    with pytest.raises(o2d.O2DError) as e:
        o2d.clear_inner(2, nx, ny, idx, pnx, pny, CM, IPS, *t)
    assert e.value.code == -1
    bad = idx.copy()
    bad[5] = 99
    with pytest.raises(o2d.O2DError) as e:
        o2d.nearest_distances(nx, ny, bad, *t)
    assert e.value.code == -2
    dup = idx.copy()
    dup[1] = dup[0]  # panel 0 from a node to itself
    with pytest.raises(o2d.O2DError) as e:
        o2d.reflect(nx, ny, dup, pnx, pny, *t)
    assert e.value.code == -1


def test_unsettled_targets_are_reported_after_writing_results():
    nx, ny = np.linspace(-2, 2, 9), np.zeros(9)
    idx = np.stack([np.arange(8), np.arange(1, 9)], 1).astype(np.uint32).ravel()
    pnx, pny = np.zeros(8), np.ones(8)
    tx, ty = np.array([0.1, 0.2, 0.3]), np.array([-0.1, 0.3, 0.05])
    with pytest.raises(o2d.O2DError) as e:
        o2d.clear_inner(0, nx, ny, idx, pnx, pny, CM, IPS, tx, ty)
    assert e.value.code == -7 and "2 targets" in str(e.value)
    ref = O.clear_inner(0, nx, ny, idx, pnx, pny, CM, IPS, tx, ty, max_passes=64)
    assert ref["unsettled"] == 2


def test_device_entries_match_host_entries():
    import torch
    import omega2d_amd.device as dv
    nx, ny, idx, pnx, pny = _spur()
    rng = np.random.default_rng(44)
    tx, ty = rng.uniform(-1.2, 1.2, 5000), rng.uniform(-0.5, 1.2, 5000)
    ts, tr = rng.normal(0, 1e-3, 5000), np.full(5000, 0.02)
    cu = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dt)  # noqa: E731
    dnx, dny, dpnx, dpny = cu(nx), cu(ny), cu(pnx), cu(pny)
    didx = cu(idx.astype(np.int32), torch.int32)
    dist = dv.nearest_distances(dnx, dny, didx, cu(tx), cu(ty))
    assert np.array_equal(dist.cpu().numpy(), o2d.nearest_distances(nx, ny, idx, tx, ty))
    x, y = cu(tx), cu(ty)
    c = dv.reflect(dnx, dny, didx, dpnx, dpny, x, y)
    hx, hy, hn = o2d.reflect(nx, ny, idx, pnx, pny, tx, ty)
    assert np.array_equal(x.cpu().numpy(), hx) and np.array_equal(y.cpu().numpy(), hy)
    assert c.cpu().tolist() == [hn, 0, 0]
    for method in (1, 0):
        x, y, s, r = cu(tx), cu(ty), cu(ts), cu(tr)
        circ, c = dv.clear_inner(method, dnx, dny, didx, dpnx, dpny, CM, IPS, x, y, s, r)
        h = o2d.clear_inner(method, nx, ny, idx, pnx, pny, CM, IPS, tx, ty, ts, tr)
        assert np.array_equal(x.cpu().numpy(), h["x"]) and np.array_equal(y.cpu().numpy(), h["y"])
        assert np.array_equal(s.cpu().numpy(), h["str"])
        assert c.cpu().tolist() == [h["num_cropped"], 0, 0]
        assert float(circ.cpu()[0]) == h["circ_removed"]


def test_device_entry_counts_bad_panels_without_reading_out_of_range():
    import torch
    import omega2d_amd.device as dv
    nx, ny, idx, pnx, pny = _circle(32)
    bad = idx.astype(np.int32).copy()
    bad[7] = 1000  # clamped to the last node on the device and counted
    cu = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")  # noqa: E731
    x, y = cu(np.array([0.0, 0.9])), cu(np.array([0.0, 0.0]))
    c = dv.reflect(cu(nx), cu(ny), cu(bad), cu(pnx), cu(pny), x, y)
    assert c.cpu().tolist()[2] == 1


def _node_at(want, seed):
    """A node whose squared distance from the origin, as the scan computes it, is exactly want."""
    rng = np.random.default_rng(seed)
    for _ in range(20000):
        c = rng.uniform(0.05, 0.25)
        e = np.sqrt(want - c * c)
        for k in range(-6, 7):
            ee = e + k * np.spacing(e)
            if (0.0 - c) * (0.0 - c) + (0.0 - ee) * (0.0 - ee) == want:
                return (c, ee)
    raise AssertionError("no node found")


def _one_ulp_nodes():
    """Node A = (0.3, 0) at squared distance M = 0.09 from the origin, node B at exactly
    next_up(M) and node C at next_up(next_up(M)): the distances the scan's one-ulp windows
    treat by panel order."""
    a = 0.3
    M = (0.0 - a) * (0.0 - a)
    M1 = np.nextafter(M, np.inf)
    return (a, 0.0), _node_at(M1, 5), _node_at(np.nextafter(M1, np.inf), 6), M


def _ulp_geometry(case):
    """Panels on a ring of radius 2 (far from the origin) with panels at the nearest distances
    put in at chosen indices. A near panel starts at its node and points away from the origin,
    so the origin's distance to it is the node's. The scan gives wave w panels w, w + 4, ...:
      succ_first  a panel at next_up(M) (index 5) before one at M (index 10): the sequential
                  scan keeps next_up(M) as its minimum;
      m_first     the same two the other way round;
      spread      eight panels at M, two in each wave's share;
      crowd       six panels at M in one wave's share (more than it keeps);
      both        panels at M and at next_up(M) interleaved across the waves;
      gap_first   a panel at next_up(next_up(M)) (index 5) before one at M (index 20), none
                  between them;
      three       panels at M, next_up(M) and next_up(next_up(M)) (the sequential rescan)."""
    (ax, ay), (bx, by), (cx, cy), M = _one_ulp_nodes()
    nring = 40
    ang = np.linspace(0.0, 2 * np.pi, nring, endpoint=False)
    xs, ys = list(2.0 * np.cos(ang)), list(2.0 * np.sin(ang))
    ring = [(k, (k + 1) % nring) for k in range(nring)]
    place = {"succ_first": {5: "B", 10: "A"}, "m_first": {5: "A", 10: "B"},
             "spread": {k: "A" for k in range(3, 11)},
             "crowd": {k: "A" for k in (0, 4, 8, 12, 16, 20)},
             "both": {3: "B", 4: "A", 6: "A", 9: "B", 13: "A", 14: "B"},
             "gap_first": {5: "C", 20: "A"},
             "three": {2: "C", 7: "B", 11: "A", 30: "B"}}[case]
    xs += [ax, bx, cx]
    ys += [ay, by, cy]
    nodes = {"A": nring, "B": nring + 1, "C": nring + 2}
    pans, r = [], 0
    for j in range(nring + len(place)):
        if j in place:
            n0 = nodes[place[j]]
            xs.append(1.2 * xs[n0] + 0.01 * len(pans))
            ys.append(1.2 * ys[n0])
            pans.append((n0, len(xs) - 1))
        else:
            pans.append(ring[r])
            r += 1
    nx, ny = np.array(xs), np.array(ys)
    idx = np.array(pans, dtype=np.uint32).ravel()
    _, _, pnx, pny, _ = O.panel_geometry(nx, ny, idx)
    return nx, ny, idx, pnx, pny, M


@pytest.mark.parametrize("case", ["succ_first", "m_first", "spread", "crowd", "both",
                                  "gap_first", "three"])
def test_one_ulp_windows_follow_panel_order(case):
    """Nearest distance, reflection and clearing where panels lie exactly one ulp apart in
    squared distance: the split scan must end where the sequential scan ends."""
    nx, ny, idx, pnx, pny, M = _ulp_geometry(case)
    rng = np.random.default_rng(3)
    tx = np.concatenate([[0.0, 0.0], rng.uniform(-0.5, 0.5, 300)])
    ty = np.concatenate([[0.0, 0.0], rng.uniform(-0.5, 0.5, 300)])
    ref = O.nearest_distances(nx, ny, idx, tx, ty)
    if case == "succ_first":  # the case is live: the scan's minimum is not the smallest distance
        assert ref[0] == np.sqrt(np.nextafter(M, np.inf))
    assert np.array_equal(o2d.nearest_distances(nx, ny, idx, tx, ty), ref)
    gx, gy, gn = o2d.reflect(nx, ny, idx, pnx, pny, tx, ty)
    rx, ry, rn = O.reflect(nx, ny, idx, pnx, pny, tx, ty)
    assert gn == rn and np.array_equal(gx, rx) and np.array_equal(gy, ry)
    got = o2d.clear_inner(1, nx, ny, idx, pnx, pny, CM, IPS, tx, ty)
    ref = O.clear_inner(1, nx, ny, idx, pnx, pny, CM, IPS, tx, ty)
    assert got["num_cropped"] == ref["num_cropped"]
    assert np.array_equal(got["x"], ref["x"]) and np.array_equal(got["y"], ref["y"])
