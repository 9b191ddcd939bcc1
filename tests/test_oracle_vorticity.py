"""Pins the oracle's Gaussian vorticity sum (src/InfluenceVort.h:47-126) before it checks the GPU:
golden vectors from the reference's own template, the closed form of one blob, the blob's
circulation by quadrature, and linearity. CPU only."""
import os

import numpy as np

import oracle as O
from helpers import rel_l2

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))


def test_vorticity_golden():
    g = GOLD
    w = O.points_vorticity(g["vo_sx"], g["vo_sy"], g["vo_ss"], g["vo_sr"], g["vo_tx"], g["vo_ty"])
    assert rel_l2(w, g["vo_w"]) <= 1e-14
    # the reference's float storage (STORE=float, double accumulation) is within float rounding
    assert rel_l2(w, g["vo_wf"]) <= 2e-6


def test_single_blob_closed_form():
    s, sig = 0.7, 0.05
    r = np.linspace(0.0, 0.3, 61)
    w = O.points_vorticity([0.3], [-0.2], [s], [sig], 0.3 + r, np.full_like(r, -0.2))
    # finalized (1/2pi) this is the Gaussian blob s/(pi sig^2) exp(-r^2/sig^2)
    ref = 2.0 * s * np.exp(-(r / sig) ** 2) / sig ** 2
    assert np.max(np.abs(w - ref)) <= 1e-13 * np.max(ref)


def test_blob_vorticity_integrates_to_circulation():
    rng = np.random.default_rng(2)
    n = 20
    sx, sy = rng.uniform(-0.2, 0.2, n), rng.uniform(-0.2, 0.2, n)
    ss, sr = rng.normal(0, 1, n), rng.uniform(0.03, 0.06, n)
    h = 0.004
    g = np.arange(-0.6, 0.6 + h / 2, h)
    gx, gy = np.meshgrid(g, g)
    w = O.points_vorticity(sx, sy, ss, sr, gx.ravel(), gy.ravel())
    total = w.sum() * h * h * (0.5 / np.pi)  # finalized field integrated over the plane
    assert abs(total - ss.sum()) <= 1e-9 * np.abs(ss).sum()


def test_linearity_and_empty():
    rng = np.random.default_rng(3)
    sx, sy, sr = rng.uniform(-1, 1, 50), rng.uniform(-1, 1, 50), rng.uniform(0.1, 0.3, 50)
    a, b = rng.normal(size=50), rng.normal(size=50)
    tx, ty = rng.uniform(-1, 1, 40), rng.uniform(-1, 1, 40)
    wa = O.points_vorticity(sx, sy, a, sr, tx, ty)
    wb = O.points_vorticity(sx, sy, b, sr, tx, ty)
    wab = O.points_vorticity(sx, sy, 2 * a - 3 * b, sr, tx, ty)
    assert rel_l2(wab, 2 * wa - 3 * wb) <= 1e-14
    assert np.all(O.points_vorticity([], [], [], [], tx, ty) == 0.0)
