"""Prescribed body motion in the C++ host mirror (omega2d_amd/host/Elements.h: Body, Surfaces
transform and add_body_motion) against the reference's own Body and Surfaces (src/Body.cpp,
src/ElementBase.h:250-273, src/Surfaces.h:492-523,728-752), compiled and run by
tests/golden/make_golden.py: a body-bound circle moved by expressions of t, its transformed
nodes and tangents, the body's motion added to the panel velocities, and the body's position,
velocity, wrapped orientation and rotation rate. No GPU work: the host driver's motion mode
makes no device call. The motions are the same expressions, written as C++ lambdas in the
driver."""
import os
import subprocess

import numpy as np
import pytest

from omega2d_amd import build as B

HERE = os.path.dirname(__file__)
GOLD = np.load(os.path.join(HERE, "golden", "reference_vectors.npz"))


def run_motion(tmp_path, kind, t, factor=-1.0):
    g = GOLD
    nx, ny, ni = g["bm_nx"], g["bm_ny"], g["bm_ni"]
    header = np.array([len(nx), len(ni) // 2, kind], np.int64)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(header.tobytes())
        f.write(np.array([t, factor]).tobytes())
        for a in (nx.astype(np.float32), ny.astype(np.float32), ni.astype(np.uint32)):
            f.write(np.ascontiguousarray(a).tobytes())
    r = subprocess.run([B.HOST_TEST_BIN, "motion", str(inp), str(out)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    d = np.fromfile(out, dtype=np.float64)
    nn, npn = len(nx), len(ni) // 2
    sizes = [nn, nn, npn, npn, npn, npn, 6]
    parts = np.split(d, np.cumsum(sizes)[:-1])
    return dict(zip(("x", "y", "tx", "ty", "pu", "pv", "body"), parts))


@pytest.mark.skipif(not os.path.exists(B.HOST_TEST_BIN), reason="host driver not built")
@pytest.mark.parametrize("kind", [1, 2, 3])
@pytest.mark.parametrize("ti", [0, 1, 2])
def test_transform_and_body_motion_match_reference(tmp_path, kind, ti):
    t = float(GOLD["bm_times"][ti])
    got = run_motion(tmp_path, kind, t)
    for k in ("x", "y", "tx", "ty", "pu", "pv", "body"):
        ref = GOLD[f"bm_{kind}_{ti}_{k}"]
        assert got[k].shape == ref.shape, k
        # the reference build contracts multiply-adds (x86-64-v3) and the driver does not, so a
        # value may differ in its last bits
        np.testing.assert_allclose(got[k], ref, rtol=1e-13, atol=1e-14, err_msg=k)


@pytest.mark.skipif(not os.path.exists(B.HOST_TEST_BIN), reason="host driver not built")
def test_rigid_motion_keeps_panel_lengths_and_wraps_orientation(tmp_path):
    # closed forms: a rigid motion keeps every panel's length; 7 t at t = 1.9 wraps to
    # remainder(13.3, 2 pi); the rotation-only body's panel velocity is omega x (c - centre)
    got = run_motion(tmp_path, 3, 1.9, factor=1.0)
    nx, ny, ni = GOLD["bm_nx"], GOLD["bm_ny"], GOLD["bm_ni"].reshape(-1, 2)
    x0 = nx.astype(np.float32).astype(np.float64)
    y0 = ny.astype(np.float32).astype(np.float64)
    l0 = np.hypot(x0[ni[:, 1]] - x0[ni[:, 0]], y0[ni[:, 1]] - y0[ni[:, 0]])
    l1 = np.hypot(got["x"][ni[:, 1]] - got["x"][ni[:, 0]], got["y"][ni[:, 1]] - got["y"][ni[:, 0]])
    np.testing.assert_allclose(l1, l0, rtol=1e-13)
    assert got["body"][4] == pytest.approx(np.remainder(13.3, 2 * np.pi) - (
        2 * np.pi if np.remainder(13.3, 2 * np.pi) > np.pi else 0.0), abs=1e-12)
    assert got["body"][5] == pytest.approx(7.0, rel=1e-8)
    # rotation about the transformed geometric centre: the panel velocities are perpendicular
    # to (panel centre - centre) and have magnitude omega * distance
    xc = 0.5 * (got["x"][ni[:, 0]] + got["x"][ni[:, 1]])
    yc = 0.5 * (got["y"][ni[:, 0]] + got["y"][ni[:, 1]])
    w = np.float32(got["body"][5])
    # the centre is where the velocity vanishes: u = -w (y - yc0), v = w (x - xc0)
    cy = yc + got["pu"] / w
    cx = xc - got["pv"] / w
    assert np.ptp(cx) < 1e-12 and np.ptp(cy) < 1e-12
