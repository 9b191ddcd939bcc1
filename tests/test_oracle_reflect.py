"""Pins the oracle's particle/panel proximity routines (src/Reflect.h) before they check the GPU.

Golden vectors come from the reference's own templates (tests/golden/reference_vectors.npz);
the mathematics pins the rest: brute-force segment distances, mirror images across a flat wall,
the cutoff layer of a flat wall, linear interpolation of the cut table by numpy, and the cut
table's Gaussian moments. CPU only.
"""
import os

import numpy as np
import pytest

import oracle as O
from helpers import circle_panels

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))
CM, IPS = 0.5 / np.sqrt(2.0 * np.pi), 0.05


def _geom(prefix):
    g = GOLD
    return (g[f"{prefix}_nx"].astype(np.float64), g[f"{prefix}_ny"].astype(np.float64),
            g[f"{prefix}_idx"], g[f"{prefix}_pnx"], g[f"{prefix}_pny"])


def brute_distance(nx, ny, idx, tx, ty):
    """Distance to the nearest segment point by clamped projection, vectorised over panels."""
    a, b = idx[0::2], idx[1::2]
    x0, y0, x1, y1 = nx[a], ny[a], nx[b], ny[b]
    bx, by = x1 - x0, y1 - y0
    ax, ay = tx[:, None] - x0[None], ty[:, None] - y0[None]
    t = np.clip((ax * bx + ay * by) / (bx * bx + by * by), 0.0, 1.0)
    dx, dy = ax - t * bx, ay - t * by
    return np.sqrt((dx * dx + dy * dy).min(axis=1))


def flat_wall(npan=8, half=2.0):
    """Panels along y = 0 from x = -half to +half, walked in +x: normals point to +y."""
    xs = np.linspace(-half, half, npan + 1)
    nx, ny = xs.astype(np.float64), np.zeros(npan + 1)
    idx = np.stack([np.arange(npan), np.arange(1, npan + 1)], 1).astype(np.uint32).ravel()
    return nx, ny, idx, np.zeros(npan), np.ones(npan)


# ----------------------------------------------------------------------------- golden vectors

@pytest.mark.parametrize("prefix", ["rf", "rt"])
def test_nearest_distances_golden(prefix):
    nx, ny, idx, _, _ = _geom(prefix)
    d = O.nearest_distances(nx, ny, idx, GOLD[f"{prefix}_tx"], GOLD[f"{prefix}_ty"])
    ref = GOLD[f"{prefix}_dist"]
    assert np.max(np.abs(d - ref) / ref) <= 1e-13


@pytest.mark.parametrize("prefix", ["rf", "rt"])
def test_reflect_golden(prefix):
    nx, ny, idx, pnx, pny = _geom(prefix)
    x, y, n = O.reflect(nx, ny, idx, pnx, pny, GOLD[f"{prefix}_tx"], GOLD[f"{prefix}_ty"])
    assert np.max(np.abs(x - GOLD[f"{prefix}_refl_x"])) <= 1e-14
    assert np.max(np.abs(y - GOLD[f"{prefix}_refl_y"])) <= 1e-14
    moved = (GOLD[f"{prefix}_refl_x"] != GOLD[f"{prefix}_tx"]) | \
            (GOLD[f"{prefix}_refl_y"] != GOLD[f"{prefix}_ty"])
    assert n == int(moved.sum()) and n > 0


def test_panel_normals_match_geometry():
    for prefix in ("rf", "rt"):
        nx, ny, idx, pnx, pny = _geom(prefix)
        _, _, ox, oy, _ = O.panel_geometry(nx, ny, idx)
        assert np.array_equal(ox, pnx) and np.array_equal(oy, pny)


@pytest.mark.parametrize("kind", ["particles", "field"])
def test_clear_push_golden(kind):
    nx, ny, idx, pnx, pny = _geom("rf")
    g = GOLD
    ts, tr = (g["rf_ts"], g["rf_tr"]) if kind == "particles" else (None, None)
    r = O.clear_inner(1, nx, ny, idx, pnx, pny, CM, IPS, g["rf_tx"], g["rf_ty"], ts, tr)
    tag = "rf_c1" if kind == "particles" else "rf_c1f"
    assert np.max(np.abs(r["x"] - g[f"{tag}_x"])) <= 1e-14
    assert np.max(np.abs(r["y"] - g[f"{tag}_y"])) <= 1e-14
    assert r["unsettled"] == 0 and r["circ_removed"] == 0.0
    if ts is not None:
        assert np.array_equal(r["str"], ts)


def test_clear_push_golden_spur():
    nx, ny, idx, pnx, pny = _geom("rt")
    r = O.clear_inner(1, nx, ny, idx, pnx, pny, CM, IPS, GOLD["rt_tx"], GOLD["rt_ty"])
    assert np.max(np.abs(r["x"] - GOLD["rt_c1_x"])) <= 1e-14
    assert np.max(np.abs(r["y"] - GOLD["rt_c1_y"])) <= 1e-14


def test_clear_trim_golden_shell():
    nx, ny, idx, pnx, pny = _geom("rf")
    g = GOLD
    r = O.clear_inner(0, nx, ny, idx, pnx, pny, CM, IPS, g["rf_c0sh_tx"], g["rf_c0sh_ty"],
                      g["rf_c0sh_ts"], g["rf_c0sh_tr"])
    assert np.max(np.abs(r["x"] - g["rf_c0sh_x"])) <= 1e-13
    assert np.max(np.abs(r["y"] - g["rf_c0sh_y"])) <= 1e-13
    assert np.max(np.abs(r["str"] - g["rf_c0sh_s"])) <= 1e-15
    ref_circ = float(g["rf_c0sh_circ"][0])
    assert abs(r["circ_removed"] - ref_circ) <= 1e-12 * abs(ref_circ)
    # the removed circulation is exactly the strength lost
    assert abs(r["circ_removed"] - (g["rf_c0sh_ts"] - r["str"]).sum()) <= 1e-15
    assert r["unsettled"] == 0 and r["num_cropped"] > 0


def test_clear_trim_deep_band():
    """A target deeper than 3.05 radii below the layer loses all strength and moves to exactly
    one radius above the layer; rounding of that height decides one more trim, in the reference
    as here. Both land between 1 and 1 + (one trim step) radii above the layer."""
    nx, ny, idx, pnx, pny = _geom("rf")
    g = GOLD
    r = O.clear_inner(0, nx, ny, idx, pnx, pny, CM, IPS, g["rf_c0dp_tx"], g["rf_c0dp_ty"],
                      g["rf_c0dp_ts"], g["rf_c0dp_tr"])
    rad = g["rf_c0dp_tr"]
    pos, _, ds = O.cut_table()
    step = float(np.interp(1.0, pos, ds))  # the largest shift taken from height 1
    for x, y, s in ((r["x"], r["y"], r["str"]), (g["rf_c0dp_x"], g["rf_c0dp_y"], g["rf_c0dp_s"])):
        h = (brute_distance(nx, ny, idx, x, y) - CM * IPS) / rad
        assert np.all(h >= 1.0 - 1e-9) and np.all(h <= 1.0 + step + 1e-9)
        assert np.all(s == 0.0)


def test_cut_table_golden():
    pos, sm, ds = O.cut_table()
    assert len(pos) == len(GOLD["rf_ct_pos"]) == 63
    assert np.array_equal(pos, GOLD["rf_ct_pos"])
    assert np.max(np.abs(sm - GOLD["rf_ct_sm"])) <= 1e-15
    assert np.max(np.abs(ds - GOLD["rf_ct_ds"])) <= 1e-14


# ----------------------------------------------------------------------------- mathematics

def test_nearest_distances_brute_force():
    rng = np.random.default_rng(3)
    nx, ny, idx = circle_panels(37)
    nx, ny = nx.astype(np.float64), ny.astype(np.float64)
    tx, ty = rng.uniform(-1, 1, 3000), rng.uniform(-1, 1, 3000)
    d = O.nearest_distances(nx, ny, idx, tx, ty)
    ref = brute_distance(nx, ny, idx, tx, ty)
    assert np.max(np.abs(d - ref) / np.maximum(ref, 1e-300)) <= 1e-12
    # the centre is an apothem away from every panel
    c = O.nearest_distances(nx, ny, idx, np.zeros(1), np.zeros(1))[0]
    assert abs(c - 0.5 * np.cos(np.pi / 37)) <= 1e-7
    # no panels: the reference's starting minimum, DBL_MAX
    e = O.nearest_distances(np.zeros(0), np.zeros(0), np.zeros(0, np.uint32), tx[:3], ty[:3])
    assert np.all(e == np.sqrt(np.finfo(np.float64).max))


def test_reflect_flat_wall_mirror():
    rng = np.random.default_rng(4)
    nx, ny, idx, pnx, pny = flat_wall()
    tx, ty = rng.uniform(-1.5, 1.5, 500), rng.uniform(-0.3, 0.3, 500)
    x, y, n = O.reflect(nx, ny, idx, pnx, pny, tx, ty)
    below = ty < 0
    assert n == int(below.sum())
    assert np.max(np.abs(y[below] + ty[below])) <= 1e-15
    assert np.max(np.abs(x[below] - tx[below])) <= 1e-15
    assert np.array_equal(x[~below], tx[~below]) and np.array_equal(y[~below], ty[~below])


def test_reflect_leaves_convex_body_empty():
    rng = np.random.default_rng(5)
    nx, ny, idx = circle_panels(64)
    nx, ny = nx.astype(np.float64), ny.astype(np.float64)
    _, _, pnx, pny, _ = O.panel_geometry(nx, ny, idx)
    rho, a = rng.uniform(0.05, 0.49, 2000), rng.uniform(0, 2 * np.pi, 2000)
    x, y, n = O.reflect(nx, ny, idx, pnx, pny, rho * np.cos(a), rho * np.sin(a))
    assert n == 2000
    assert np.min(np.hypot(x, y)) >= 0.5 * np.cos(np.pi / 64) - 1e-12


def test_clear_push_flat_wall_layer():
    rng = np.random.default_rng(6)
    nx, ny, idx, pnx, pny = flat_wall()
    tx, ty = rng.uniform(-1.5, 1.5, 800), rng.uniform(-0.2, 0.2, 800)
    r = O.clear_inner(1, nx, ny, idx, pnx, pny, CM, IPS, tx, ty)
    layer = CM * IPS
    pushed = ty - layer < -0.001 * IPS
    assert r["num_cropped"] == int(pushed.sum()) and r["unsettled"] == 0
    assert np.max(np.abs(r["y"][pushed] - layer)) <= 1e-15
    assert np.max(np.abs(r["x"][pushed] - tx[pushed])) <= 1e-15
    assert np.array_equal(r["y"][~pushed], ty[~pushed])


def test_clear_trim_flat_wall_one_pass_interpolates_table():
    rng = np.random.default_rng(7)
    nx, ny, idx, pnx, pny = flat_wall()
    n = 600
    rad = np.full(n, 0.02)
    tx, ty = rng.uniform(-1.5, 1.5, n), CM * IPS + rng.uniform(-0.07, 0.03, n)
    ts = rng.normal(0, 1, n)
    r = O.clear_inner(0, nx, ny, idx, pnx, pny, CM, IPS, tx, ty, ts, rad, max_passes=1)
    p = (ty - CM * IPS) / rad
    pos, sm, ds = O.cut_table()
    hit = p < 1.0
    inside = hit & (p >= pos[0])
    assert np.max(np.abs(r["str"][inside] - ts[inside] * np.interp(p[inside], pos, sm))) <= 1e-14
    assert np.max(np.abs(r["y"][inside] - (ty[inside] + np.interp(p[inside], pos, ds) * rad[inside]))) <= 1e-15
    deep = p < pos[0]
    assert np.all(r["str"][deep] == 0.0)
    assert np.max(np.abs(r["y"][deep] - (CM * IPS + rad[deep]))) <= 1e-15
    assert np.array_equal(r["str"][~hit], ts[~hit])
    assert abs(r["circ_removed"] - (ts - r["str"]).sum()) <= 1e-13
    assert r["num_cropped"] == int(hit.sum())


def test_cut_table_is_a_cumulative_gaussian():
    pos, sm, ds = O.cut_table()
    assert pos[0] == pytest.approx(-3.05) and pos[-1] == pytest.approx(3.15)
    assert sm[0] == 0.0 and sm[-1] == 1.0 and ds[-1] == 0.0
    assert np.all(np.diff(pos) > 0) and np.all(np.diff(sm) >= 0)
    # rows: fraction of a unit 2D Gaussian below x_i + dx/2, i.e. (1 + erf(x))/2 sampled on the
    # grid, and the shift that restores the kept part's centroid, -E[x | x < x_i]
    from scipy.special import erf
    x = pos[1:-1] - 0.05
    assert np.max(np.abs(sm[1:-1] - 0.5 * (1 + erf(x + 0.05)))) <= 2e-3
    mid = len(pos) // 2  # the row at x = 0
    assert ds[mid] == pytest.approx(1 / np.sqrt(np.pi), rel=0.06)
    assert sm[-2] == pytest.approx(1.0, abs=1e-4)


def test_clear_unsettled_field_points_are_reported():
    """Method 0 never moves field points, so one below the layer would keep the reference's
    outer loop running for ever; here it is counted as unsettled after max_passes."""
    nx, ny, idx, pnx, pny = flat_wall()
    tx, ty = np.array([0.1, 0.2, 0.3]), np.array([-0.1, 0.3, 0.05])
    r = O.clear_inner(0, nx, ny, idx, pnx, pny, CM, IPS, tx, ty, max_passes=5)
    # heights relative to the layer: -0.11 and 0.04 are below ips = 0.05, 0.29 is above
    assert r["unsettled"] == 2 and r["num_cropped"] == 10
    assert np.array_equal(r["x"], tx) and np.array_equal(r["y"], ty)
