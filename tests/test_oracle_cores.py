"""Pins every core of the fp64 oracle's particle sums (oracle/o2d_oracle.c, core_k and core_kg).

The reference compiles one core function, chosen by a hard-wired #define
(src/CoreFunc.h:35-39). oracle/Makefile builds the reference harness once per core with only that
define switched, and tests/golden/core_vectors.npz holds their outputs: five cores x {point,
particle targets} x {velocity, vel+vorticity, vel+shear}, the self interaction, float storage,
and a Gaussian set that straddles the core's cutoffs at 16, 0.001 and the float literal 0.001f
that the reference actually compares with.

What the mathematics fixes is checked too, independently of both the reference and the oracle's
formulas:
  * the vorticity and the strain rate of one particle equal the curl and the strain of its
    velocity field, differentiated by central differences in 60-digit decimal arithmetic of the
    closed-form velocity profiles (for every core, point and particle targets);
  * the Gaussian core's vorticity is that of the smooth Lamb-Oseen vortex in every branch (the
    reference's gradient factor carries e^-x whichever branch K took), and it equals the
    separately pinned InfluenceVort kernel (src/InfluenceVort.h:47-126);
  * a Gaussian particle seen by a Gaussian particle target is the Lamb-Oseen vortex with
    sigma^2 = sigma_s^2 + sigma_t^2, the convolution of the two Gaussians.
CPU only.
"""
import decimal
import os
from decimal import Decimal as D

import numpy as np
import pytest

import oracle as O
from helpers import particles, rel_l2

HERE = os.path.dirname(__file__)
CV = np.load(os.path.join(HERE, "golden", "core_vectors.npz"))
CORES = ["rm", "gauss", "wl", "v2", "v3"]
NEAR_CUT = float(np.float32(0.001))  # the reference's 0.001f, promoted to double


def _cv_inputs(tag):
    g = CV
    tr = g["cv_tr"] if tag == "0b" else None
    return (g["cv_sx"], g["cv_sy"], g["cv_ss"], g["cv_sr"], g["cv_tx"], g["cv_ty"], tr)


# ----------------------------------------------------------------------------- golden vectors

@pytest.mark.parametrize("res", ["vel", "vort", "shear"])
@pytest.mark.parametrize("tag", ["0p", "0b"])
@pytest.mark.parametrize("core", CORES)
def test_core_golden(core, tag, res):
    out = O.points_on_points(core, *_cv_inputs(tag), res)
    for k, a in zip("uvw", out):
        assert rel_l2(a, CV[f"cv_{core}_{tag}_{res}_{k}"]) <= 1e-14, (core, tag, res, k)


@pytest.mark.parametrize("core", CORES)
def test_core_golden_self(core):
    g = CV
    out = O.points_on_points(core, g["cv_sx"], g["cv_sy"], g["cv_ss"], g["cv_sr"], g["cv_sx"],
                             g["cv_sy"], g["cv_sr"], "vort")
    for k, a in zip("uvw", out):
        ref = g[f"cv_{core}_self_{k}"]
        if core == "gauss" and k == "w":
            # the reference's Gaussian gradient factor at the coincident pair is inf*0 = NaN
            # (src/CoreFunc.h:164-165 with 1/d^2 = inf), which poisons every target's
            # vorticity; the oracle takes that factor as 0 (it multiplies dx = dy = 0)
            assert np.all(np.isnan(ref)) and np.all(np.isfinite(a))
            continue
        assert rel_l2(a, ref) <= 1e-14, (core, k)


@pytest.mark.parametrize("tag", ["0p", "0b"])
@pytest.mark.parametrize("core", CORES)
def test_core_golden_float_storage(core, tag):
    # the application's float storage with double sums agrees with fp64 to float precision,
    # the accuracy bar the split-precision GPU form is held to
    f = [None if a is None else a.astype(np.float32).astype(np.float64) for a in _cv_inputs(tag)]
    u, v = O.points_on_points(core, *f)
    assert rel_l2(u, CV[f"cvf_{core}_{tag}_u"]) <= 1e-6
    assert rel_l2(v, CV[f"cvf_{core}_{tag}_v"]) <= 1e-6


@pytest.mark.parametrize("tag", ["0p", "0b"])
def test_gaussian_cutoffs_golden(tag):
    """Targets at x = d^2/w on both sides of 16, 0.001 and 0.001f. The branch taken must be the
    reference's for every target: K jumps by e^-16 (1.1e-7) at 16 and by about 5e-4 between
    0.001 and 0.001f."""
    g = CV
    sx, sy, ss, sr = (np.array([v]) for v in g["gc_src"])
    tx, ty = g[f"gc_{tag}_tx"], g[f"gc_{tag}_ty"]
    tr = None if tag == "0p" else np.full(len(tx), g["gc_trad"][0])
    for res in ("vel", "vort", "shear"):
        out = O.points_on_points("gauss", sx, sy, ss, sr, tx, ty, tr, res)
        for k, a in zip("uvw", out):
            ref = g[f"gc_{tag}_{res}_{k}"]
            if k in "uv":
                # one pair per target: every target on its own, to a few ulps
                assert np.all(np.abs(a - ref) <= 1e-14 * np.abs(ref)), (res, k)
            else:
                # the reference's vorticity and shear cancel in the outer branch (bbb d^2 + 2K);
                # the residue is rounding, far below the 5e-4 jump a wrong branch would make
                assert np.all(np.abs(a - ref) <= 1e-13 * np.max(np.abs(ref))), (res, k)
    # and the set does reach into the window between 0.001 and 0.001f
    xs = g[f"gc_{tag}_x"]
    assert np.sum((xs >= 0.001) & (xs < NEAR_CUT)) >= 25


# ----------------------------------------------------------------------------- live builds

@pytest.mark.skipif(not all(O.have_ref(c) for c in CORES), reason="per-core reference builds absent")
def test_reference_builds_carry_their_core():
    for c in CORES:
        assert O.ref(c).ref_core_id() == O.CORES[c]


@pytest.mark.skipif(not all(O.have_ref(c) for c in CORES), reason="per-core reference builds absent")
@pytest.mark.parametrize("core", CORES)
def test_oracle_matches_live_reference_build_each_core(core):
    rng = np.random.default_rng(31)
    sx, sy, ss, sr = particles(rng, 500, rad=(0.02, 0.3))
    tx, ty, _, tr = particles(rng, 250, rad=(0.02, 0.3))
    for res in ("vel", "vort", "shear"):
        for trr in (None, tr):
            a = O.points_on_points(core, sx, sy, ss, sr, tx, ty, trr, res)
            b = O.ref_points_on_points(sx, sy, ss, sr, tx, ty, trr, res, core=core)
            for x, y in zip(a, b):
                assert rel_l2(x, y) <= 1e-14, (core, res, trr is None)


# ----------------------------------------------------------------------------- closed forms

decimal.getcontext().prec = 60
_ONE, _TWO, _THREE = D(1), D(2), D(3)


def _k_closed(core, q, w):
    """K(q = d^2) of each core in Decimal, from its textbook velocity profile:
    RM u_theta ~ r/(r^2+w); Lamb-Oseen (1-e^{-r^2/w})/r; Winckelmans-Leonard
    r(r^2+2w)/(r^2+w)^2; Vatistas n=2 r/(r^4+w)^(1/2) and n=3 r/(r^6+w)^(1/3), with w the
    core's radius term."""
    if core == "rm":
        return _ONE / (q + w)
    if core == "gauss":
        return (_ONE - (-q / w).exp()) / q
    if core == "wl":
        return (q + _TWO * w) / (q + w) ** 2
    if core == "v2":
        return _ONE / (q * q + w).sqrt()
    return _ONE / (q * q * q + w) ** (_ONE / _THREE)


def _vel(core, w, x, y):
    q = x * x + y * y
    k = _k_closed(core, q, w)
    return -k * y, k * x


def _grad(core, w, x, y, h=D("1e-22")):
    """(du/dx, du/dy, dv/dx) of one unit particle at the origin, by central differences."""
    upx, vpx = _vel(core, w, x + h, y)
    umx, vmx = _vel(core, w, x - h, y)
    upy, _ = _vel(core, w, x, y + h)
    umy, _ = _vel(core, w, x, y - h)
    return (upx - umx) / (2 * h), (upy - umy) / (2 * h), (vpx - vmx) / (2 * h)


def _rterm(core, r):
    r = D(float(r))  # the exact binary value
    return r ** 4 if core == "v2" else (r ** 6 if core == "v3" else r * r)


@pytest.mark.parametrize("tag", ["0p", "0b"])
@pytest.mark.parametrize("core", CORES)
def test_vorticity_and_strain_are_curl_and_strain_of_velocity(core, tag):
    """One particle: the oracle's vel+vort and vel+shear outputs against the curl
    dv/dx - du/dy and the reference's strain measure sqrt(u_x^2 + (v_x + u_y)^2 / 4)
    (src/Kernels.h:139-231) of the closed-form velocity field, differentiated numerically.
    A dropped term, a wrong sign or factor in the gradient factor bbb fails here."""
    sr, trad = 0.07, (0.05 if tag == "0b" else 0.0)
    w = _rterm(core, sr) + (_rterm(core, trad) if trad else D(0))
    seff = (sr * sr + trad * trad) ** 0.5
    for rad, ang in ((0.3, 0.4), (1.0, 1.9), (2.2, 3.0), (3.7, -2.2)):
        x, y = rad * seff * np.cos(ang), rad * seff * np.sin(ang)
        tr = None if tag == "0p" else [trad]
        _, _, wo = O.points_on_points(core, [0.0], [0.0], [1.0], [sr], [x], [y], tr, "vort")
        _, _, eo = O.points_on_points(core, [0.0], [0.0], [1.0], [sr], [x], [y], tr, "shear")
        ux, uy, vx = _grad(core, w, D(float(x)), D(float(y)))
        curl = float(vx - uy)
        strain = float((ux * ux + (vx + uy) ** 2 / 4).sqrt())
        assert wo[0] == pytest.approx(curl, rel=1e-12), (core, tag, rad)
        xg = rad * rad  # d^2 / w for the Gaussian
        if core != "gauss" or NEAR_CUT < xg < 16.0:
            # the Gaussian's strain is the smooth profile's only between its cutoffs
            assert eo[0] == pytest.approx(strain, rel=1e-12), (core, tag, rad)


@pytest.mark.parametrize("tag", ["0p", "0b"])
def test_gaussian_vorticity_is_lamb_oseen_in_every_branch(tag):
    """bbb = -2 K/d^2 + 2 e^-x/(w d^2) whichever branch K took (src/CoreFunc.h:156-180), so the
    vorticity bbb d^2 + 2K is 2 e^-x / w, the smooth Lamb-Oseen vortex's, inside the core
    (K = 1/w), between the cutoffs and beyond them (K = 1/d^2)."""
    sr, trad = 0.05, (0.03 if tag == "0b" else 0.0)
    w = sr * sr + trad * trad
    for x in (1e-5, 5e-4, 0.01, 0.5, 3.0, 9.0, 15.0, 17.0, 20.0):
        d = np.sqrt(x * w)
        tr = None if tag == "0p" else [trad]
        _, _, wo = O.points_on_points("gauss", [0.0], [0.0], [1.0], [sr], [d], [0.0], tr, "vort")
        exact = 2.0 * np.exp(-x) / w
        # beyond x ~ 10 the reference's form cancels 2/d^2 against itself: absolute bound
        assert abs(wo[0] - exact) <= 1e-12 * exact + 4e-16 * 2.0 / (d * d), x


def test_gaussian_vorticity_equals_influence_vort():
    """The Gaussian core's vel+vort vorticity against InfluenceVort (pinned by its own golden
    vectors), on many particles at point targets."""
    rng = np.random.default_rng(41)
    sx, sy, ss, sr = particles(rng, 400, rad=(0.1, 0.3))
    tx, ty, _, _ = particles(rng, 300)
    _, _, w = O.points_on_points("gauss", sx, sy, ss, sr, tx, ty, None, "vort")
    wv = O.points_vorticity(sx, sy, ss, sr, tx, ty)
    assert rel_l2(w, wv) <= 1e-12


def test_gaussian_particle_target_is_convolution():
    """A Gaussian particle seen by a Gaussian particle target: the Lamb-Oseen profile of
    sigma^2 = sigma_s^2 + sigma_t^2, K = (1 - e^{-d^2/sigma^2}) / d^2, between the cutoffs."""
    ss_, st = 0.04, 0.03
    s2 = ss_ * ss_ + st * st
    for x in (0.01, 0.2, 1.0, 4.0, 12.0):
        d = np.sqrt(x * s2)
        u, v = O.points_on_points("gauss", [0.0], [0.0], [1.0], [ss_], [0.0], [d], [st])
        assert -u[0] == pytest.approx(-np.expm1(-x) / d, rel=1e-14), x
        assert abs(v[0]) <= 1e-300


def test_closed_form_radius_combinations():
    """Particle targets combine the radius terms by core: r^2 (RM, Gaussian, WL), r^4 (V2),
    r^6 (V3)."""
    s, t, d = 0.1, 0.2, 0.3
    cases = {
        "rm": d / (d * d + s * s + t * t),
        "wl": d * (d * d + 2 * (s * s + t * t)) / (d * d + s * s + t * t) ** 2,
        "v2": d / np.sqrt(d ** 4 + s ** 4 + t ** 4),
        "v3": d / np.cbrt(d ** 6 + s ** 6 + t ** 6),
        "gauss": -np.expm1(-d * d / (s * s + t * t)) / d,
    }
    for core, speed in cases.items():
        _, v = O.points_on_points(core, [0.0], [0.0], [1.0], [s], [d], [0.0], [t])
        assert v[0] == pytest.approx(speed, rel=1e-14), core
