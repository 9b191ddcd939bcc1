"""The GPU particle sums against the reference's own outputs for every core (the per-core builds
of the reference harness, tests/golden/core_vectors.npz), and the fp64 Gaussian's cutoff branch
at single-ulp distances from 16, 0.001f and 0.001 against the oracle.

Bars: plain fp64 <= 1e-12 relative L2 (per target where one pair decides it); split fp32 <= 1e-6
for velocity, the form's stated error target.
"""
import os

import numpy as np
import pytest

import oracle as O
import omega2d_amd as o2d
from helpers import rel_l2

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(__file__)
CV = np.load(os.path.join(HERE, "golden", "core_vectors.npz"))
CORES = ["rm", "gauss", "wl", "v2", "v3"]


def _inputs(tag):
    g = CV
    tr = g["cv_tr"] if tag == "0b" else None
    return (g["cv_sx"], g["cv_sy"], g["cv_ss"], g["cv_sr"], g["cv_tx"], g["cv_ty"], tr)


@pytest.mark.parametrize("res", ["vel", "vort", "shear"])
@pytest.mark.parametrize("tag", ["0p", "0b"])
@pytest.mark.parametrize("core", CORES)
def test_fp64_matches_reference_golden(core, tag, res):
    got = o2d.points_on_points(*_inputs(tag), core=core, results=res, precision="fp64")
    for k, a in zip("uvw", got):
        assert rel_l2(a, CV[f"cv_{core}_{tag}_{res}_{k}"]) <= 1e-12, (core, tag, res, k)


@pytest.mark.parametrize("tag", ["0p", "0b"])
@pytest.mark.parametrize("core", CORES)
def test_split_matches_reference_golden(core, tag):
    u, v = o2d.points_on_points(*_inputs(tag), core=core, precision="split")
    assert rel_l2(u, CV[f"cv_{core}_{tag}_vel_u"]) <= 1e-6
    assert rel_l2(v, CV[f"cv_{core}_{tag}_vel_v"]) <= 1e-6


@pytest.mark.parametrize("core", CORES)
@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_self_interaction_matches_reference_golden(core, precision):
    g = CV
    x, y, s, r = g["cv_sx"], g["cv_sy"], g["cv_ss"], g["cv_sr"]
    u, v = o2d.points_on_points(x, y, s, r, x, y, r, core=core, precision=precision)
    tol = 1e-12 if precision == "fp64" else 1e-6
    assert rel_l2(u, g[f"cv_{core}_self_u"]) <= tol and rel_l2(v, g[f"cv_{core}_self_v"]) <= tol


@pytest.mark.parametrize("tag", ["0p", "0b"])
def test_fp64_gaussian_cutoffs_match_reference_golden(tag):
    g = CV
    sx, sy, ss, sr = (np.array([v]) for v in g["gc_src"])
    tx, ty = g[f"gc_{tag}_tx"], g[f"gc_{tag}_ty"]
    tr = None if tag == "0p" else np.full(len(tx), g["gc_trad"][0])
    u, v = o2d.points_on_points(sx, sy, ss, sr, tx, ty, tr, core="gauss", precision="fp64")
    for a, ref in ((u, g[f"gc_{tag}_vel_u"]), (v, g[f"gc_{tag}_vel_v"])):
        # one pair per target: a target in the wrong branch is off by 1.1e-7 to 5e-4
        assert np.all(np.abs(a - ref) <= 1e-13 * np.abs(ref))


def _ulp_targets(sx, sy, w, cut, k=12):
    """Targets whose squared distance from (sx, sy) steps through d^2 = cut * w one ulp of the
    offset at a time, so that the reference's predicate (1/w)/(1/d^2) > cut (or < cut) flips
    inside the set."""
    d0 = np.sqrt(cut * w)
    ds = d0 + np.arange(-k, k + 1) * np.spacing(d0)
    ang = np.linspace(0.1, 6.0, len(ds))
    return sx + ds * np.cos(ang), sy + ds * np.sin(ang)


@pytest.mark.parametrize("particle_targets", [False, True])
def test_fp64_gaussian_branch_at_single_ulps(particle_targets):
    """At every cutoff (16, the reference's 0.001f, and 0.001 which it does not use) the GPU
    must take the branch the oracle takes, target by target."""
    sx, sy, ss, sr = 0.37, -0.21, 1.3, 0.031
    trad = 0.017 if particle_targets else 0.0
    w = sr * sr + trad * trad
    tx, ty = [], []
    for cut in (16.0, float(np.float32(0.001)), 0.001):
        a, b = _ulp_targets(sx, sy, w, cut)
        tx.append(a)
        ty.append(b)
    tx, ty = np.concatenate(tx), np.concatenate(ty)
    tr = np.full(len(tx), trad) if particle_targets else None
    src = (np.array([sx]), np.array([sy]), np.array([ss]), np.array([sr]))
    ru, rv = O.points_on_points("gauss", *src, tx, ty, tr)
    gu, gv = o2d.points_on_points(*src, tx, ty, tr, core="gauss", precision="fp64")
    for a, ref in ((gu, ru), (gv, rv)):
        assert np.all(np.abs(a - ref) <= 1e-14 * np.abs(ref))
    # and the set spans both branches at each cutoff
    dx, dy = tx - sx, ty - sy
    reld2 = (1.0 / w) / (1.0 / (dx * dx + dy * dy))
    for cut in (16.0, float(np.float32(0.001))):
        seg = np.abs(reld2 / cut - 1.0) < 1e-12
        assert np.any(reld2[seg] > cut) and np.any(reld2[seg] <= cut)
