"""Pins the fp64 oracle (oracle/o2d_oracle.c) before it is trusted as the GPU checker.

Each oracle function is checked against the golden vectors written from the reference's own
templates (tests/golden/reference_vectors.npz) and against what the mathematics fixes: closed
forms, invariants, quadrature, brute force and a LAPACK solve. CPU only.
"""
import os

import numpy as np
import pytest

import oracle as O
from helpers import circle_panels, particles, rel_l2

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))
CORES = ["rm", "gauss", "wl", "v2", "v3"]


# ----------------------------------------------------------------------------- golden vectors

@pytest.mark.parametrize("res", ["vel", "vort", "shear"])
@pytest.mark.parametrize("tag", ["0p", "0b"])
def test_points_on_points_golden_v2(res, tag):
    g = GOLD
    tr = g["pp_tr"] if tag == "0b" else None
    out = O.points_on_points("v2", g["pp_sx"], g["pp_sy"], g["pp_ss"], g["pp_sr"], g["pp_tx"],
                             g["pp_ty"], tr, res)
    for k, a in zip("uvw", out):
        assert rel_l2(a, g[f"pp_{tag}_{res}_{k}"]) <= 1e-14


def test_points_on_points_golden_self():
    g = GOLD
    out = O.points_on_points("v2", g["pp_sx"], g["pp_sy"], g["pp_ss"], g["pp_sr"], g["pp_sx"],
                             g["pp_sy"], g["pp_sr"], "vort")
    for k, a in zip("uvw", out):
        assert rel_l2(a, g[f"pp_self_{k}"]) <= 1e-14


def test_points_on_points_golden_float_storage():
    # the application stores float and sums in double: agreement to float precision
    g = GOLD
    f = [g[k].astype(np.float32).astype(np.float64)
         for k in ("pp_sx", "pp_sy", "pp_ss", "pp_sr", "pp_tx", "pp_ty", "pp_tr")]
    u, v = O.points_on_points("v2", *f)
    assert rel_l2(u, g["ppf_0b_u"]) <= 1e-6 and rel_l2(v, g["ppf_0b_v"]) <= 1e-6


def test_panels_on_points_golden():
    g = GOLD
    u, v = O.panels_on_points(g["pan_nx"], g["pan_ny"], g["pan_idx"], g["pan_vs"], g["pan_ss"],
                              g["pan_tx"], g["pan_ty"])
    assert rel_l2(u, g["pan_u"]) <= 1e-14 and rel_l2(v, g["pan_v"]) <= 1e-14


def test_points_on_panels_golden():
    g = GOLD
    u, v = O.points_on_panels(g["ptp_sx"], g["ptp_sy"], g["ptp_ss"], g["pan_nx"], g["pan_ny"],
                              g["pan_idx"])
    assert rel_l2(u, g["ptp_u"]) <= 1e-14 and rel_l2(v, g["ptp_v"]) <= 1e-14


@pytest.mark.parametrize("nunk", [1, 2])
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_bem_assemble_golden(nunk, kind):
    g = GOLD
    nx, ny, idx = g["bem_nx"], g["bem_ny"], g["bem_idx"]
    pre = f"bem_{nunk}_{kind}_"
    Aref = g[pre + "A"]
    augmented = Aref.shape[0] == len(idx) // 2 * nunk + 1
    assert augmented == (kind != 0)
    tx, ty, _, _, area = O.panel_geometry(nx, ny, idx)
    assert rel_l2(tx, g[pre + "tang_x"]) <= 1e-15 and rel_l2(area, g[pre + "area"]) <= 1e-15
    vol, cx, cy = O.body_volume(nx, ny, idx)
    if augmented:
        assert abs(vol - g[pre + "vol"][0]) <= 1e-15
        # the augmented column from the restated unit-rotation strengths and panel velocities
        vs, ss = O.unit_rot_strengths(nx, ny, idx, cx, cy)
        _, _, nnx, nny, _ = O.panel_geometry(nx, ny, idx)
        # collocation points just off each panel, stored as float by represent_as_particles
        # (src/Surfaces.h:830-878): wgt*x1 + (1-wgt)*x0 + offset*normal with wgt = 1/2
        a, b = idx[0::2], idx[1::2]
        mx = (0.5 * nx[b].astype(np.float64) + 0.5 * nx[a] + 1e-4 * nnx).astype(np.float32)
        my = (0.5 * ny[b].astype(np.float64) + 0.5 * ny[a] + 1e-4 * nny).astype(np.float32)
        uu, uv = O.panels_on_points(nx, ny, idx, vs, ss, mx, my)
        uu, uv = O.finalize(uu, uv)
        assert rel_l2(uu, g[pre + "urot_u"]) <= 1e-12 and rel_l2(uv, g[pre + "urot_v"]) <= 1e-12
    A = O.bem_assemble(nx, ny, idx, nunk, augmented, g[pre + "urot_u"], g[pre + "urot_v"],
                       g[pre + "vol"][0] if augmented else 0.0)
    assert rel_l2(A, Aref) <= 1e-14
    assert np.max(np.abs(A - Aref)) <= 1e-15


@pytest.mark.parametrize("nunk", [1, 2])
@pytest.mark.parametrize("kinds", ["22", "20", "02", "00"])
@pytest.mark.parametrize("tag", ["ab", "ba"])
def test_bem_block_between_bodies_golden(nunk, kinds, tag):
    """Block of one body on another (src/Coefficients.h with &src != &targ), augmented or not."""
    g = GOLD
    a = (g["bp_ax"].astype(np.float64), g["bp_ay"].astype(np.float64), g["bp_ai"])
    b = (g["bp_bx"].astype(np.float64), g["bp_by"].astype(np.float64), g["bp_bi"])
    src, targ = (a, b) if tag == "ab" else (b, a)
    key = f"bp_{tag}_{nunk}_{kinds}"
    ref = g[key + "_A"]
    got = O.bem_assemble_pair(*src, *targ, nunk, kinds[0] == "2", kinds[1] == "2",
                              g[key + "_uu"], g[key + "_uv"])
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) <= 1e-15 * np.max(np.abs(ref))


def test_bem_internal_flow_not_augmented():
    # reversed panel order bounds internal flow: negative volume, no augmentation for ground
    g = GOLD
    vol, _, _ = O.body_volume(g["bem_nx"], g["bem_ny"], g["bem_int_idx"])
    assert vol < 0 and abs(vol - g["bem_int_vol"][0]) <= 1e-15
    A = O.bem_assemble(g["bem_nx"], g["bem_ny"], g["bem_int_idx"], 2, False)
    assert rel_l2(A, g["bem_int_A"]) <= 1e-14


@pytest.mark.parametrize("nunk", [1, 2])
def test_rhs_golden(nunk):
    g = GOLD
    tx, ty, _, _, _ = O.panel_geometry(g["bem_nx"], g["bem_ny"], g["bem_idx"])
    rhs = O.vels_to_rhs(g["rhs_pu"], g["rhs_pv"], tx, ty, nunk, g["rhs_nbc"].astype(np.float64))
    assert rel_l2(rhs, g[f"rhs_{nunk}"]) <= 1e-15


def test_moves_and_finalize_golden():
    g = GOLD
    us = [g["mv_u1"], g["mv_u2"], g["mv_u3"]]
    vs = [g["mv_v1"], g["mv_v2"], g["mv_v3"]]
    # the reference build may fuse multiply-adds; the oracle never does, so compare to an ulp
    x2, y2 = O.rk_combine(g["mv_x"], g["mv_y"], us[:2], vs[:2], [0.25, 0.75], 0.01)
    assert rel_l2(x2, g["mv2_x"]) <= 1e-16 and rel_l2(y2, g["mv2_y"]) <= 1e-16
    x3, y3 = O.rk_combine(g["mv_x"], g["mv_y"], us, vs, [2 / 9, 3 / 9, 4 / 9], 0.01)
    assert rel_l2(x3, g["mv3_x"]) <= 1e-16 and rel_l2(y3, g["mv3_y"]) <= 1e-16
    fu, fv = O.finalize(us[0], vs[0], (1.0, -0.5))
    assert rel_l2(fu, g["fin_u"]) <= 1e-16 and rel_l2(fv, g["fin_v"]) <= 1e-16


# ----------------------------------------------------------------------------- closed forms

def _single_vortex_speed(core, r, sig, gamma=1.0):
    """Unscaled tangential speed of one particle at distance r (point target)."""
    u, v = O.points_on_points(core, [0.0], [0.0], [gamma], [sig], [r], [0.0])
    assert abs(u[0]) < 1e-300 or abs(u[0]) <= 1e-15 * abs(v[0])
    return v[0]


@pytest.mark.parametrize("r", [1e-3, 0.01, 0.05, 0.2, 1.0])
def test_closed_form_profiles(r):
    s = 0.1
    x = r * r / (s * s)
    # u_theta * 2 pi = Gamma r K(r^2): RM, Gaussian (three branches), WL, V2, V3
    assert _single_vortex_speed("rm", r, s) == pytest.approx(r / (r * r + s * s), rel=1e-15)
    if x > 16:
        gk = 1.0 / r
    elif x < float(np.float32(0.001)):  # the reference compares with the float 0.001f
        gk = r / (s * s)
    else:
        gk = -np.expm1(-x) / r
    assert _single_vortex_speed("gauss", r, s) == pytest.approx(gk, rel=1e-12)
    assert _single_vortex_speed("wl", r, s) == pytest.approx(
        r * (r * r + 2 * s * s) / (r * r + s * s) ** 2, rel=1e-14)
    assert _single_vortex_speed("v2", r, s) == pytest.approx(r / np.sqrt(r**4 + s**4), rel=1e-15)
    assert _single_vortex_speed("v3", r, s) == pytest.approx(r / np.cbrt(r**6 + s**6), rel=1e-15)


def test_gaussian_cutoff_branches():
    s = 0.1
    # just outside the compact cutoff the reference returns exactly 1/d^2
    r = 4.0001 * s
    assert _single_vortex_speed("gauss", r, s) == 1.0 / r
    # deep inside: the constant-core branch 1/sigma^2
    r = 0.01 * s
    assert _single_vortex_speed("gauss", r, s) == pytest.approx(r / (s * s), rel=1e-15)


def test_particle_target_core_adds_radii():
    # 0b kernels regularise with both radii: RM K = 1/(d^2 + s^2 + t^2)
    u, v = O.points_on_points("rm", [0.0], [0.0], [1.0], [0.1], [0.3], [0.0], [0.2])
    assert v[0] == pytest.approx(0.3 / (0.09 + 0.01 + 0.04), rel=1e-15)
    u, v = O.points_on_points("v2", [0.0], [0.0], [1.0], [0.1], [0.3], [0.0], [0.2])
    assert v[0] == pytest.approx(0.3 / np.sqrt(0.09**2 + 0.1**4 + 0.2**4), rel=1e-15)


# ----------------------------------------------------------------------------- invariants

@pytest.mark.parametrize("core", CORES)
def test_invariants(core):
    rng = np.random.default_rng(1)
    sx, sy, ss, sr = particles(rng, 200, rad=(0.05, 0.2))
    tx, ty, _, tr = particles(rng, 100, rad=(0.05, 0.2))
    u, v = O.points_on_points(core, sx, sy, ss, sr, tx, ty, tr)
    # linearity in strength
    u2, v2 = O.points_on_points(core, sx, sy, 3.0 * ss, sr, tx, ty, tr)
    assert rel_l2(u2, 3 * u) <= 1e-14 and rel_l2(v2, 3 * v) <= 1e-14
    # translation invariance
    ut, vt = O.points_on_points(core, sx + 0.75, sy - 1.25, ss, sr, tx + 0.75, ty - 1.25, tr)
    assert rel_l2(ut, u) <= 1e-12 and rel_l2(vt, v) <= 1e-12
    # rotation equivariance
    c, s = np.cos(0.7), np.sin(0.7)
    ur, vr = O.points_on_points(core, c * sx - s * sy, s * sx + c * sy, ss, sr,
                                c * tx - s * ty, s * tx + c * ty, tr)
    assert rel_l2(ur, c * u - s * v) <= 1e-12 and rel_l2(vr, s * u + c * v) <= 1e-12
    # symmetric self-interaction conserves linear impulse: sum_i Gamma_i u_i = 0
    us, vs = O.points_on_points(core, sx, sy, ss, sr, sx, sy, sr)
    scale = np.sum(np.abs(ss * us))
    assert abs(np.sum(ss * us)) <= 1e-12 * scale and abs(np.sum(ss * vs)) <= 1e-12 * scale


def test_brute_force_long_double():
    rng = np.random.default_rng(2)
    sx, sy, ss, sr = particles(rng, 40, rad=(0.05, 0.2))
    tx, ty, _, tr = particles(rng, 30, rad=(0.05, 0.2))
    L = np.longdouble
    for core in CORES:
        u, v = O.points_on_points(core, sx, sy, ss, sr, tx, ty, tr)
        bu = np.zeros(30, dtype=L)
        bv = np.zeros(30, dtype=L)
        for i in range(30):
            dx = L(tx[i]) - sx.astype(L)
            dy = L(ty[i]) - sy.astype(L)
            d2 = dx * dx + dy * dy
            s2 = sr.astype(L) ** 2 + L(tr[i]) ** 2
            if core == "rm":
                k = 1 / (d2 + s2)
            elif core == "gauss":
                x = d2 / s2
                k = np.where(x > 16, 1 / d2,
                             np.where(x < L(np.float32(0.001)), 1 / s2, -np.expm1(-x) / d2))
            elif core == "wl":
                k = (d2 + 2 * s2) / (d2 + s2) ** 2
            elif core == "v2":
                k = 1 / np.sqrt(d2 * d2 + sr.astype(L) ** 4 + L(tr[i]) ** 4)
            else:
                k = 1 / np.cbrt(d2**3 + sr.astype(L) ** 6 + L(tr[i]) ** 6)
            bu[i] = np.sum(-ss.astype(L) * k * dy)
            bv[i] = np.sum(ss.astype(L) * k * dx)
        assert rel_l2(u, bu.astype(np.float64)) <= 1e-13, core
        assert rel_l2(v, bv.astype(np.float64)) <= 1e-13, core


# ----------------------------------------------------------------------------- panels

def _gl_panel(x0, y0, x1, y1, vs, ss, tx, ty, n=64):
    """Gauss-Legendre quadrature of point vortices (vs) and point sources (ss) along a panel."""
    t, w = np.polynomial.legendre.leggauss(n)
    t = 0.5 * (t + 1.0)
    L = np.hypot(x1 - x0, y1 - y0)
    w = 0.5 * w * L
    px, py = x0 + t * (x1 - x0), y0 + t * (y1 - y0)
    dx, dy = tx - px, ty - py
    r2 = dx * dx + dy * dy
    u = np.sum(w * (-vs * dy + ss * dx) / r2)
    v = np.sum(w * (vs * dx + ss * dy) / r2)
    return u, v


def test_panel_kernel_vs_quadrature():
    rng = np.random.default_rng(3)
    for _ in range(20):
        x0, y0, x1, y1 = rng.uniform(-0.1, 0.1, 4)
        ang = rng.uniform(0, 2 * np.pi)
        dist = rng.uniform(0.5, 2.0)
        tx, ty = dist * np.cos(ang), dist * np.sin(ang)
        vs, ss = rng.normal(size=2)
        u, v = O.panels_on_points([x0, x1], [y0, y1], [0, 1], [vs], [ss], [tx], [ty])
        qu, qv = _gl_panel(x0, y0, x1, y1, vs, ss, tx, ty)
        assert u[0] == pytest.approx(qu, rel=1e-11, abs=1e-14)
        assert v[0] == pytest.approx(qv, rel=1e-11, abs=1e-14)
        u1, v1 = O.panels_on_points([x0, x1], [y0, y1], [0, 1], [vs], None, [tx], [ty])
        qu1, qv1 = _gl_panel(x0, y0, x1, y1, vs, 0.0, tx, ty)
        assert u1[0] == pytest.approx(qu1, rel=1e-11, abs=1e-14)
        assert v1[0] == pytest.approx(qv1, rel=1e-11, abs=1e-14)


def test_uniform_vortex_circle():
    # a closed ring of constant sheet strength: no velocity inside, Gamma/(2 pi r) outside
    npan = 2000
    nx, ny, idx = circle_panels(npan)
    nx, ny = nx.astype(np.float64), ny.astype(np.float64)
    gam = 1.0
    vs = np.full(npan, gam)
    _, _, _, _, area = O.panel_geometry(nx, ny, idx)
    circ = np.sum(vs * area)
    tx = np.array([0.0, 0.1, -0.2, 1.0, 0.0, -2.0])
    ty = np.array([0.0, 0.05, 0.1, 0.0, 1.5, 0.3])
    u, v = O.panels_on_points(nx, ny, idx, vs, None, tx, ty)
    u, v = O.finalize(u, v)
    r = np.hypot(tx, ty)
    inside = r < 0.4
    assert np.max(np.hypot(u[inside], v[inside])) < 1e-6
    speed = np.hypot(u[~inside], v[~inside])
    assert np.allclose(speed, abs(circ) / (2 * np.pi * r[~inside]), rtol=1e-5)


def test_points_on_panels_vs_quadrature_average():
    # panel-averaged velocity of singular point vortices, by quadrature along each panel
    rng = np.random.default_rng(4)
    nx, ny, idx = circle_panels(16)
    nx, ny = nx.astype(np.float64), ny.astype(np.float64)
    r = rng.uniform(0.8, 2.0, 50)
    th = rng.uniform(0, 2 * np.pi, 50)
    sx, sy, ss = r * np.cos(th), r * np.sin(th), rng.normal(size=50)
    pu, pv = O.points_on_panels(sx, sy, ss, nx, ny, idx)
    t, w = np.polynomial.legendre.leggauss(64)
    t, w = 0.5 * (t + 1), 0.5 * w
    for i in range(16):
        a, b = idx[2 * i], idx[2 * i + 1]
        qx = nx[a] + t * (nx[b] - nx[a])
        qy = ny[a] + t * (ny[b] - ny[a])
        dx = qx[:, None] - sx[None, :]
        dy = qy[:, None] - sy[None, :]
        r2 = dx * dx + dy * dy
        eu = np.sum(w[:, None] * (-ss * dy / r2))
        ev = np.sum(w[:, None] * (ss * dx / r2))
        assert pu[i] == pytest.approx(eu, rel=1e-10, abs=1e-13)
        assert pv[i] == pytest.approx(ev, rel=1e-10, abs=1e-13)


# ----------------------------------------------------------------------------- BEM

@pytest.mark.parametrize("nunk", [1, 2])
def test_bem_structure(nunk):
    nx, ny, idx = circle_panels(40)
    nx, ny = nx.astype(np.float64), ny.astype(np.float64)
    vol, _, _ = O.body_volume(nx, ny, idx)
    A = O.bem_assemble(nx, ny, idx, nunk, True, np.ones(40), np.zeros(40), vol)
    n = 40 * nunk
    assert np.allclose(np.diag(A)[:n], 0.5, rtol=0, atol=1e-15)
    _, _, _, _, area = O.panel_geometry(nx, ny, idx)
    if nunk == 1:
        assert np.array_equal(A[n, :n], area)
    else:
        assert np.array_equal(A[n, 0:n:2], area) and np.all(A[n, 1:n:2] == 0)
    assert A[n, n] == 2 * vol
    # polygon area of the inscribed 40-gon
    assert vol == pytest.approx(0.5 * 40 * 0.25 * np.sin(2 * np.pi / 40), rel=1e-6)


def test_bem_circle_potential_flow():
    # uniform flow past a circle: the slip vortex sheet is 2 U |sin(theta)|
    npan = 256
    nx, ny, idx = circle_panels(npan)
    nx, ny = nx.astype(np.float64), ny.astype(np.float64)
    tx, ty, _, _, area = O.panel_geometry(nx, ny, idx)
    vol, cx, cy = O.body_volume(nx, ny, idx)
    vs, ss = O.unit_rot_strengths(nx, ny, idx, cx, cy)
    A = O.bem_assemble(nx, ny, idx, 1, True, np.zeros(npan), np.zeros(npan), vol)
    pu, pv = O.finalize(np.zeros(npan), np.zeros(npan), (1.0, 0.0))
    rhs = np.concatenate([O.vels_to_rhs(pu, pv, tx, ty, 1), [0.0]])
    gam = np.linalg.solve(A, rhs)
    mx = 0.5 * (nx[idx[0::2]] + nx[idx[1::2]])
    my = 0.5 * (ny[idx[0::2]] + ny[idx[1::2]])
    th = np.arctan2(my, mx)
    assert np.max(np.abs(np.abs(gam[:npan]) - 2 * np.abs(np.sin(th)))) < 0.02
    # circulation constraint row: total sheet circulation equals the prescribed 0
    assert abs(np.dot(gam[:npan], area)) < 1e-10


# ----------------------------------------------------------------------------- adjoint

@pytest.mark.parametrize("core", CORES)
def test_adjoint_vs_central_differences(core):
    rng = np.random.default_rng(6)
    ns, nt = 12, 9
    sx, sy, ss, sr = particles(rng, ns, rad=(0.15, 0.4))
    tx, ty = rng.uniform(-1, 1, nt), rng.uniform(-1, 1, nt)
    gu, gv = rng.normal(size=nt), rng.normal(size=nt)

    def J(sx_, sy_, ss_, sr_):
        u, v = O.points_on_points(core, sx_, sy_, ss_, sr_, tx, ty)
        return np.dot(gu, u) + np.dot(gv, v)

    gx, gy, gs, gr = O.points_on_points_adjoint(core, sx, sy, ss, sr, tx, ty, gu, gv)
    h = 1e-6
    for j in range(ns):
        for arr, g in ((sx, gx), (sy, gy), (ss, gs), (sr, gr)):
            p, m = arr.copy(), arr.copy()
            p[j] += h
            m[j] -= h
            args_p = [p if a is arr else a for a in (sx, sy, ss, sr)]
            args_m = [m if a is arr else a for a in (sx, sy, ss, sr)]
            fd = (J(*args_p) - J(*args_m)) / (2 * h)
            assert g[j] == pytest.approx(fd, rel=2e-6, abs=1e-7), (core, j)


def test_adjoint_strength_gradient_is_transposed_sum():
    # dJ/dGamma_j = sum_k K_kj (gv dx - gu dy): the velocity of unit-strength source j ...
    rng = np.random.default_rng(8)
    sx, sy, ss, sr = particles(rng, 30)
    tx, ty = rng.uniform(-1, 1, 20), rng.uniform(-1, 1, 20)
    gu, gv = rng.normal(size=20), rng.normal(size=20)
    _, _, gs, _ = O.points_on_points_adjoint("v2", sx, sy, ss, sr, tx, ty, gu, gv)
    for j in range(30):
        u, v = O.points_on_points("v2", sx[j:j + 1], sy[j:j + 1], [1.0], sr[j:j + 1], tx, ty)
        assert gs[j] == pytest.approx(np.dot(gu, u) + np.dot(gv, v), rel=1e-13, abs=1e-15)


# ----------------------------------------------------------------------------- Morton keys

def test_morton_keys_brute_force():
    rng = np.random.default_rng(10)
    x, y = rng.uniform(-3, 5, 500), rng.uniform(-2, 2, 500)
    x0, y0 = x.min(), y.min()
    inv = 1.0 / max(x.max() - x0, y.max() - y0)
    keys = O.morton_keys(x, y, x0, y0, inv)
    for i in range(500):
        qx = min(65535, max(0, int(np.floor((x[i] - x0) * inv * 65536.0))))
        qy = min(65535, max(0, int(np.floor((y[i] - y0) * inv * 65536.0))))
        k = 0
        for b in range(16):
            k |= ((qx >> b) & 1) << (2 * b)
            k |= ((qy >> b) & 1) << (2 * b + 1)
        assert keys[i] == k


# ----------------------------------------------------------------------------- reference build

@pytest.mark.skipif(not O.have_ref(), reason="reference build oracle/_ref not present")
def test_oracle_matches_live_reference_build():
    rng = np.random.default_rng(11)
    sx, sy, ss, sr = particles(rng, 700)
    tx, ty, _, tr = particles(rng, 333)
    for res in ("vel", "vort", "shear"):
        a = O.points_on_points("v2", sx, sy, ss, sr, tx, ty, tr, res)
        b = O.ref_points_on_points(sx, sy, ss, sr, tx, ty, tr, res)
        for x, y in zip(a, b):
            assert rel_l2(x, y) <= 1e-14
