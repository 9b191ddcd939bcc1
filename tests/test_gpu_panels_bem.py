"""GPU parity of the panel kernels, BEM assembly/solve, RK update and adjoint vs the oracle."""
import numpy as np
import pytest

import oracle as O
import omega2d_amd as o2d
from helpers import circle_panels, particles, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("with_source", [False, True])
@pytest.mark.parametrize("npan", [5, 64, 300])
def test_panels_on_points(with_source, npan):
    rng = np.random.default_rng(npan)
    nx, ny, idx = circle_panels(npan)
    nx, ny = nx.astype(np.float64), ny.astype(np.float64)
    vs = rng.normal(size=npan)
    ss = rng.normal(size=npan) if with_source else None
    tx, ty, _, _ = particles(rng, 3001, lo=-2.0, hi=2.0)
    got = o2d.panels_on_points(nx, ny, idx, vs, ss, tx, ty)
    ref = O.panels_on_points(nx, ny, idx, vs, ss, tx, ty)
    for g, r in zip(got, ref):
        assert rel_l2(g, r) <= 1e-12


@pytest.mark.parametrize("npan", [7, 128, 1000])
@pytest.mark.parametrize("ns", [1, 300, 20000])
def test_points_on_panels(npan, ns):
    rng = np.random.default_rng(npan + ns)
    nx, ny, idx = circle_panels(npan)
    nx, ny = nx.astype(np.float64), ny.astype(np.float64)
    r = rng.uniform(0.55, 2.5, ns)
    th = rng.uniform(0, 2 * np.pi, ns)
    sx, sy, ss = r * np.cos(th), r * np.sin(th), rng.normal(size=ns)
    got = o2d.points_on_panels(sx, sy, ss, nx, ny, idx)
    ref = O.points_on_panels(sx, sy, ss, nx, ny, idx)
    for g, rr in zip(got, ref):
        assert rel_l2(g, rr) <= 1e-12


def test_clustered_panel_sums_over_two_bodies():
    """From 128 panels the panel sums expand far clusters of 32 panels; here clusters straddle
    the two bodies (200 + 129 panels), targets crowd both surfaces and one lies on a cluster's
    end node, and a NaN particle must reach every panel sum as in the per-panel kernels."""
    ax, ay, ai = circle_panels(200)
    bx, by, bi = circle_panels(129, diam=0.6, cx=1.6, cy=0.3)
    nx = np.concatenate([ax, bx]).astype(np.float64)
    ny = np.concatenate([ay, by]).astype(np.float64)
    idx = np.concatenate([ai, bi + len(ax)]).astype(np.uint32)
    npan = len(idx) // 2
    rng = np.random.default_rng(11)
    vs, ss = rng.normal(size=npan), rng.normal(size=npan)
    n = 6000
    t = rng.uniform(0, 2 * np.pi, n)
    r = np.where(rng.uniform(size=n) < 0.5, 0.5 * rng.uniform(1.0, 1.3, n), rng.uniform(0.2, 3.0, n))
    tx, ty = r * np.cos(t), r * np.sin(t)
    tx[:n // 3] = 1.6 + 0.3 * np.cos(t[:n // 3]) * rng.uniform(1.0, 1.5, n // 3)
    ty[:n // 3] = 0.3 + 0.3 * np.sin(t[:n // 3]) * rng.uniform(1.0, 1.5, n // 3)
    tx[5], ty[5] = nx[idx[2 * 40]], ny[idx[2 * 40]]  # on an end node
    got = o2d.panels_on_points(nx, ny, idx, vs, ss, tx, ty)
    ref = O.panels_on_points(nx, ny, idx, vs, ss, tx, ty)
    ok = np.isfinite(ref[0]) & np.isfinite(ref[1])
    for g, rr in zip(got, ref):
        assert np.array_equal(np.isfinite(g), np.isfinite(rr))
        assert rel_l2(g[ok], rr[ok]) <= 1e-12
    qs = rng.normal(size=n)
    keep = np.hypot(tx - 0.0, ty) > 0.52  # particles off the first body
    keep &= np.hypot(tx - 1.6, ty - 0.3) > 0.32
    sx, sy, sq = tx[keep], ty[keep], qs[keep]
    got = o2d.points_on_panels(sx, sy, sq, nx, ny, idx)
    ref = O.points_on_panels(sx, sy, sq, nx, ny, idx)
    for g, rr in zip(got, ref):
        assert rel_l2(g, rr) <= 1e-12
    sx = sx.copy()
    sx[len(sx) // 2] = np.nan
    got = o2d.points_on_panels(sx, sy, sq, nx, ny, idx)
    assert np.all(np.isnan(got[0])) and np.all(np.isnan(got[1]))


@pytest.mark.parametrize("nunk", [1, 2])
@pytest.mark.parametrize("augmented", [False, True])
@pytest.mark.parametrize("npan", [16, 257])
def test_bem_assemble(nunk, augmented, npan):
    nx, ny, idx = circle_panels(npan)
    nx, ny = nx.astype(np.float64), ny.astype(np.float64)
    rng = np.random.default_rng(npan)
    uu, uv = rng.normal(size=npan), rng.normal(size=npan)
    vol, _, _ = O.body_volume(nx, ny, idx)
    got = o2d.bem_assemble(nx, ny, idx, nunk, augmented, uu, uv, vol)
    ref = O.bem_assemble(nx, ny, idx, nunk, augmented, uu, uv, vol)
    assert got.shape == ref.shape
    assert rel_l2(got, ref) <= 1e-12
    assert np.max(np.abs(got - ref)) <= 1e-13


@pytest.mark.parametrize("nunk", [1, 2])
@pytest.mark.parametrize("src_aug,targ_aug", [(True, True), (True, False), (False, True),
                                              (False, False)])
def test_bem_block_between_bodies(nunk, src_aug, targ_aug):
    ax, ay, ai = circle_panels(300)
    bx, by, bi = circle_panels(129, diam=0.6, cx=1.6, cy=0.3)
    ax, ay, bx, by = (v.astype(np.float64) for v in (ax, ay, bx, by))
    rng = np.random.default_rng(nunk)
    uu, uv = rng.normal(size=129), rng.normal(size=129)
    got = o2d.bem_assemble_pair(ax, ay, ai, bx, by, bi, nunk, src_aug, targ_aug, uu, uv)
    ref = O.bem_assemble_pair(ax, ay, ai, bx, by, bi, nunk, src_aug, targ_aug, uu, uv)
    assert got.shape == ref.shape == (129 * nunk + targ_aug, 300 * nunk + src_aug)
    assert rel_l2(got, ref) <= 1e-12
    assert np.max(np.abs(got - ref)) <= 1e-13


@pytest.mark.parametrize("nunk", [1, 2])
def test_bem_factor_solve_matches_lapack(nunk):
    npan = 200
    nx, ny, idx = circle_panels(npan)
    nx, ny = nx.astype(np.float64), ny.astype(np.float64)
    tx, ty, tnx, tny, area = O.panel_geometry(nx, ny, idx)
    vol, cx, cy = O.body_volume(nx, ny, idx)
    vs, ss = O.unit_rot_strengths(nx, ny, idx, cx, cy)
    A = O.bem_assemble(nx, ny, idx, nunk, True, np.zeros(npan), np.zeros(npan), vol)
    rng = np.random.default_rng(1)
    b = rng.normal(size=A.shape[0])
    o2d.bem_factor(A)
    x = o2d.bem_solve(b)
    xr = np.linalg.solve(A, b)
    assert rel_l2(x, xr) <= 1e-12
    # solve again with the same factors (the reference re-uses its solver until A changes): the
    # second solve builds the explicit inverse, later ones are matrix-vector products
    for k in range(5):
        bk = rng.normal(size=A.shape[0])
        assert rel_l2(o2d.bem_solve(bk), np.linalg.solve(A, bk)) <= 1e-12
    # a new factorisation drops the inverse of the old matrix
    A2 = A + np.diag(np.full(A.shape[0], 0.25))
    o2d.bem_factor(A2)
    for k in range(6):
        bk = rng.normal(size=A.shape[0])
        assert rel_l2(o2d.bem_solve(bk), np.linalg.solve(A2, bk)) <= 1e-12


def test_bem_repeated_device_solves_in_place():
    """Device entry with x aliasing b, on the inverse path, against LAPACK."""
    import torch
    import omega2d_amd.device as dv
    npan = 300
    nx, ny, idx = circle_panels(npan)
    nx, ny = nx.astype(np.float64), ny.astype(np.float64)
    vol, _, _ = O.body_volume(nx, ny, idx)
    A = O.bem_assemble(nx, ny, idx, 2, True, np.zeros(npan), np.zeros(npan), vol)
    dA = torch.from_numpy(np.asfortranarray(A).ravel(order="F").copy()).cuda()
    dv.bem_factor(dA.view(A.shape[0], A.shape[0]))
    rng = np.random.default_rng(3)
    for k in range(6):  # triangular solves, then the inverse (built on the second)
        b = rng.normal(size=A.shape[0])
        db = torch.from_numpy(b).cuda()
        dv.bem_solve(db, db)
        assert rel_l2(db.cpu().numpy(), np.linalg.solve(A, b)) <= 1e-12


@pytest.mark.parametrize("n", [1, 7, 8, 9, 63, 64, 65, 130, 1000, 4100, 6145])
def test_lu_of_general_matrices_matches_lapack(n):
    """Random matrices need row interchanges at almost every column: the blocked factorisation
    (orders up to 6144) and the library fallback above it solve like LAPACK."""
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n))
    o2d.bem_factor(A)
    for k in range(5):  # the fresh factors, then the inverse built from them (second solve)
        b = rng.normal(size=n)
        x = o2d.bem_solve(b)
        assert rel_l2(x, np.linalg.solve(A, b)) <= 1e-9, (n, k)
        assert np.linalg.norm(A @ x - b) <= 1e-12 * np.linalg.norm(A) * np.linalg.norm(x)


@pytest.mark.parametrize("n", [65, 700, 2049, 4097])
def test_lu_factors_and_pivots_equal_lapack(n):
    """The kept factors are LAPACK getrf's: the same interchanges at every column and the same
    L and U up to rounding (a different pivot anywhere would change whole rows of both)."""
    import scipy.linalg as sl
    rng = np.random.default_rng(n + 11)
    A = rng.normal(size=(n, n))
    o2d.bem_factor(A)
    lu, piv = o2d.bem_factors(n)
    rlu, rpiv = sl.lu_factor(A)
    assert np.array_equal(piv, rpiv + 1)
    assert np.max(np.abs(lu - rlu)) <= 1e-9 * np.max(np.abs(rlu))


@pytest.mark.parametrize("K", [0, 11, 70])
@pytest.mark.parametrize("rows", [(3, 10, 70), (5, 517, 1029), (130, 700, 1300), (600, 1100),
                                  (1590,)])
def test_lu_pivot_ties_take_the_first_row(K, rows):
    """Exact ties for the largest magnitude in pivot column K: idamax takes the first row. The
    panel kernel's thread t holds rows t, t + 512, ...; the tied rows sit in one thread, in
    one wave and in different waves. Columns k < K pivot on row 1599 - 2k (entry 10, zero in
    the other pivot rows), and column K is zero in those rows, so no update reaches it before
    its own pivot search and the tie stays exact."""
    import scipy.linalg as sl
    n = 1600
    rows = [r + 80 if r < K else r for r in rows]
    rng = np.random.default_rng(K + len(rows))
    A = rng.uniform(-1.0, 1.0, size=(n, n))
    prow = [n - 1 - 2 * k for k in range(K)]
    for k, r in enumerate(prow):
        A[prow, k] = 0.0
        A[r, k] = 10.0
    A[prow, K] = 0.0
    for i, r in enumerate(rows):
        A[r, K] = 5.0 if i % 2 == 0 else -5.0
    o2d.bem_factor(A)
    lu, piv = o2d.bem_factors(n)
    rlu, rpiv = sl.lu_factor(A)
    assert piv[K] == min(rows) + 1
    assert np.array_equal(piv, rpiv + 1)
    assert np.max(np.abs(lu - rlu)) <= 1e-9 * np.max(np.abs(rlu))


def test_bem_factors_need_factors_of_that_size():
    o2d.bem_factor(np.eye(3) * 2.0)
    with pytest.raises(o2d.O2DError):
        o2d.bem_factors(4)


@pytest.mark.parametrize("n", [70, 300])
def test_lu_reports_a_singular_matrix(n):
    rng = np.random.default_rng(1)
    A = rng.normal(size=(n, n))
    A[:, 5] = 0.0
    with pytest.raises(o2d.O2DError) as e:
        o2d.bem_factor(A)
    assert e.value.code == -7
    assert "singular" in str(e.value)


def test_bem_solve_without_factor_of_that_size_fails():
    o2d.bem_factor(np.eye(3) * 2.0)
    with pytest.raises(o2d.O2DError):
        o2d.bem_solve(np.ones(5))


@pytest.mark.parametrize("nstage,w", [(1, [1.0]), (2, [0.25, 0.75]), (3, [2 / 9, 3 / 9, 4 / 9])])
def test_rk_move_bit_exact(nstage, w):
    rng = np.random.default_rng(nstage)
    n = 10007
    x, y = rng.normal(size=n), rng.normal(size=n)
    us = [rng.normal(size=n) for _ in range(nstage)]
    vs = [rng.normal(size=n) for _ in range(nstage)]
    gx, gy = o2d.rk_move(x.copy(), y.copy(), us, vs, w, 0.01)
    rx, ry = O.rk_combine(x, y, us, vs, w, 0.01)
    assert np.array_equal(gx, rx) and np.array_equal(gy, ry)


@pytest.mark.parametrize("core", ["rm", "gauss", "wl", "v2", "v3"])
def test_adjoint_matches_oracle(core):
    rng = np.random.default_rng(4)
    sx, sy, ss, sr = particles(rng, 3000, rad=(0.05, 0.3))
    th = 2 * np.pi * np.arange(500) / 500
    tx, ty = 0.8 * np.cos(th), 0.8 * np.sin(th)
    gu, gv = rng.normal(size=500), rng.normal(size=500)
    got = o2d.points_on_points_adjoint(sx, sy, ss, sr, tx, ty, gu, gv, core=core)
    ref = O.points_on_points_adjoint(core, sx, sy, ss, sr, tx, ty, gu, gv)
    for g, r in zip(got, ref):
        assert rel_l2(g, r) <= 1e-12
