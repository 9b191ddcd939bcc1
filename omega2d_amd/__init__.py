"""omega2d_amd - MI355X (gfx950) direct Biot-Savart velocity path of Omega2D.

The compute lives in the HIP library omega2d_amd/lib/libomega2d_hip.so behind the C ABI in
include/omega2d_hip.h. This module is a thin numpy binding over its synchronous host-array
entry points (the drop-in form: arrays in, velocities accumulated out); ``omega2d_amd.device``
binds the device-pointer entry points for torch tensors already in HBM, and
``omega2d_amd.autograd`` wraps the adjoint as a torch.autograd.Function.

Every function raises O2DError when the library reports a failure; there is no CPU fallback.
"""
import ctypes

import numpy as np

from ._cabi import (CORES, PRECISIONS, RESULTS, Env, O2DError, TreeOpts, call, lib,  # noqa: F401
                    make_env,
                    CORE_RM, CORE_GAUSSIAN, CORE_WL, CORE_V2, CORE_V3, VELONLY, VELANDVORT,
                    VELANDSHEAR, PLAIN_FP64, SPLIT_FP32)

__all__ = ["Env", "O2DError", "make_env", "points_on_points", "panels_on_points",
           "points_on_panels", "bem_assemble", "bem_assemble_pair", "bem_factor", "bem_solve", "rk_move",
           "points_on_points_adjoint", "points_on_points_tree", "points_vorticity", "nearest_distances", "reflect",
           "clear_inner",
           "flops_per_pair", "release"]


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _arr(a, dtype):
    a = np.asarray(a)
    if a.dtype != dtype:
        raise TypeError(f"expected {np.dtype(dtype).name} array, got {a.dtype}")
    return np.ascontiguousarray(a)


def _out(a, n, dtype):
    if a is None:
        return np.zeros(n, dtype)
    if a.dtype != dtype or not a.flags.c_contiguous or a.shape != (n,):
        raise TypeError("output arrays must be contiguous, of the storage dtype and length n")
    return a


def flops_per_pair(core, point_targets, panel=0):
    c = CORES[core] if isinstance(core, str) else int(core)
    return int(lib().o2d_hip_flops_per_pair(c, int(bool(point_targets)), int(panel)))


def release():
    call("o2d_hip_release")


def points_on_points(sx, sy, ss, sr, tx, ty, tr=None, *, core="v2", precision="fp64",
                     results="vel", device=0, tu=None, tv=None, aux=None):
    """points_affect_points (src/Influence.h:56-416) on the GPU.

    Float64 inputs use the _d entry point, float32 inputs the _f one (the reference app's
    STORE=float). Unscaled velocities are accumulated into tu, tv (new zero arrays if None);
    results="vort"/"shear" also accumulates into aux. tr=None means point targets.
    Returns (tu, tv) or (tu, tv, aux).
    """
    dt = np.asarray(sx).dtype
    if dt not in (np.float64, np.float32):
        raise TypeError("inputs must be float64 or float32")
    sx, sy, ss, sr, tx, ty = (_arr(a, dt) for a in (sx, sy, ss, sr, tx, ty))
    tr = None if tr is None else _arr(tr, dt)
    nt = len(tx)
    tu, tv = _out(tu, nt, dt), _out(tv, nt, dt)
    env = make_env(core, precision, results, device)
    want_aux = env.results != VELONLY
    aux = _out(aux, nt, dt) if want_aux else None
    tw = aux if env.results == VELANDVORT else None
    te = aux if env.results == VELANDSHEAR else None
    fn = "o2d_hip_points_on_points_d" if dt == np.float64 else "o2d_hip_points_on_points_f"
    call(fn, ctypes.byref(env), len(sx), _p(sx), _p(sy), _p(ss), _p(sr), nt, _p(tx), _p(ty),
         _p(tr), _p(tu), _p(tv), _p(tw), _p(te))
    return (tu, tv, aux) if want_aux else (tu, tv)


def panels_on_points(nx, ny, idx, vs, ss, tx, ty, *, device=0, tu=None, tv=None):
    """panels_affect_points (src/Influence.h:422-614): ss=None means vortex-only panels."""
    nx, ny, vs, tx, ty = (_arr(a, np.float64) for a in (nx, ny, vs, tx, ty))
    ss = None if ss is None else _arr(ss, np.float64)
    idx = _arr(idx, np.uint32)
    nt = len(tx)
    tu, tv = _out(tu, nt, np.float64), _out(tv, nt, np.float64)
    env = make_env("v2", "fp64", "vel", device)
    call("o2d_hip_panels_on_points_d", ctypes.byref(env), len(nx), _p(nx), _p(ny), len(idx) // 2,
         _p(idx), _p(vs), _p(ss), nt, _p(tx), _p(ty), _p(tu), _p(tv))
    return tu, tv


def points_on_panels(sx, sy, ss, nx, ny, idx, *, device=0, pu=None, pv=None):
    """points_affect_panels (src/Influence.h:633-771): panel-averaged velocity, accumulated."""
    sx, sy, ss, nx, ny = (_arr(a, np.float64) for a in (sx, sy, ss, nx, ny))
    idx = _arr(idx, np.uint32)
    npan = len(idx) // 2
    pu, pv = _out(pu, npan, np.float64), _out(pv, npan, np.float64)
    env = make_env("v2", "fp64", "vel", device)
    call("o2d_hip_points_on_panels_d", ctypes.byref(env), len(sx), _p(sx), _p(sy), _p(ss), len(nx),
         _p(nx), _p(ny), npan, _p(idx), _p(pu), _p(pv))
    return pu, pv


def bem_assemble(nx, ny, idx, nunk, augmented=False, urot_u=None, urot_v=None, vol=0.0, *,
                 device=0):
    """panels_on_panels_coeff for a body on itself (src/Coefficients.h:155-573).
    Returns the n x n matrix as A[row, col] (stored column-major, like the reference)."""
    nx, ny = _arr(nx, np.float64), _arr(ny, np.float64)
    idx = _arr(idx, np.uint32)
    npan = len(idx) // 2
    n = npan * nunk + (1 if augmented else 0)
    A = np.zeros(n * n)
    uu = None if urot_u is None else _arr(urot_u, np.float64)
    uv = None if urot_v is None else _arr(urot_v, np.float64)
    env = make_env("v2", "fp64", "vel", device)
    call("o2d_hip_bem_assemble_d", ctypes.byref(env), len(nx), _p(nx), _p(ny), npan, _p(idx),
         int(nunk), int(bool(augmented)), _p(uu), _p(uv), float(vol), _p(A), n)
    return A.reshape(n, n).T


def bem_assemble_pair(snx, sny, sidx, tnx, tny, tidx, nunk, src_aug=False, targ_aug=False,
                      urot_u=None, urot_v=None, *, device=0):
    """Block of source surface s acting on a different target surface t
    (src/Coefficients.h:155-573 with &src != &targ); returns A[row, col]."""
    snx, sny, tnx, tny = (_arr(a, np.float64) for a in (snx, sny, tnx, tny))
    sidx, tidx = _arr(sidx, np.uint32), _arr(tidx, np.uint32)
    snp, tnp = len(sidx) // 2, len(tidx) // 2
    nr, nc = tnp * nunk + int(bool(targ_aug)), snp * nunk + int(bool(src_aug))
    uu = None if urot_u is None else _arr(urot_u, np.float64)
    uv = None if urot_v is None else _arr(urot_v, np.float64)
    A = np.zeros(nr * nc)
    env = make_env("v2", "fp64", "vel", device)
    call("o2d_hip_bem_assemble_pair_d", ctypes.byref(env), len(snx), _p(snx), _p(sny), snp,
         _p(sidx), len(tnx), _p(tnx), _p(tny), tnp, _p(tidx), int(nunk), int(bool(src_aug)),
         int(bool(targ_aug)), _p(uu), _p(uv), _p(A), nr)
    return A.reshape(nc, nr).T.copy()


def bem_factor(A, *, device=0):
    """LU-factor A (given as A[row, col]) on the device and keep the factors."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    colmajor = np.ascontiguousarray(A.T)
    env = make_env("v2", "fp64", "vel", device)
    call("o2d_hip_bem_factor_d", ctypes.byref(env), n, _p(colmajor), n)


def bem_factors(n, *, device=0):
    """The kept factors as LAPACK getrf returns them: (lu[row, col], 1-based ipiv)."""
    lu = np.empty(n * n, dtype=np.float64)
    ipiv = np.empty(n, dtype=np.int32)
    env = make_env("v2", "fp64", "vel", device)
    call("o2d_hip_bem_factors_d", ctypes.byref(env), n, _p(lu), _p(ipiv))
    return lu.reshape(n, n).T.copy(), ipiv


def bem_solve(b, *, device=0):
    b = _arr(b, np.float64).copy()
    env = make_env("v2", "fp64", "vel", device)
    call("o2d_hip_bem_solve_d", ctypes.byref(env), len(b), _p(b), _p(b))
    return b


def rk_move(x, y, us, vs, w, dt, *, device=0):
    """ElementBase::move (src/ElementBase.h:276-359): x += dt * sum_s w_s u_s, in place."""
    x, y = _arr(x, np.float64), _arr(y, np.float64)
    us = [_arr(a, np.float64) for a in us]
    vs = [_arr(a, np.float64) for a in vs]
    w = _arr(np.asarray(w, dtype=np.float64), np.float64)
    up = (ctypes.c_void_p * len(us))(*[a.ctypes.data for a in us])
    vp = (ctypes.c_void_p * len(vs))(*[a.ctypes.data for a in vs])
    env = make_env("v2", "fp64", "vel", device)
    call("o2d_hip_rk_move_d", ctypes.byref(env), len(x), float(dt), len(us), _p(w), up, vp, _p(x),
         _p(y))
    return x, y


def points_on_points_adjoint(sx, sy, ss, sr, tx, ty, gu, gv, *, core="gauss", device=0):
    """dJ/d(sx, sy, strength, radius) of the unscaled point-target sum given dJ/dtu, dJ/dtv."""
    sx, sy, ss, sr, tx, ty, gu, gv = (_arr(a, np.float64) for a in (sx, sy, ss, sr, tx, ty, gu, gv))
    ns = len(sx)
    out = [np.zeros(ns) for _ in range(4)]
    env = make_env(core, "fp64", "vel", device)
    call("o2d_hip_points_on_points_adjoint_d", ctypes.byref(env), ns, _p(sx), _p(sy), _p(ss),
         _p(sr), len(tx), _p(tx), _p(ty), _p(gu), _p(gv), *[_p(o) for o in out])
    return tuple(out)


def points_on_points_tree(sx, sy, ss, sr, tx, ty, tr=None, *, core="gauss", theta=0.6,
                          order=28, core_tol=1e-12, device=0, tu=None, tv=None):
    """Treecode form of points_on_points (summation_t barneshut, src/ExecEnv.h:26-31): same
    unscaled, accumulated velocity sum, fp64, with far cells replaced by their moments."""
    sx, sy, ss, sr, tx, ty = (_arr(a, np.float64) for a in (sx, sy, ss, sr, tx, ty))
    tr = None if tr is None else _arr(tr, np.float64)
    nt = len(tx)
    tu, tv = _out(tu, nt, np.float64), _out(tv, nt, np.float64)
    env = make_env(core, "fp64", "vel", device)
    opt = TreeOpts(float(theta), int(order), 0, float(core_tol))
    call("o2d_hip_points_on_points_tree_d", ctypes.byref(env), ctypes.byref(opt), len(sx), _p(sx),
         _p(sy), _p(ss), _p(sr), nt, _p(tx), _p(ty), _p(tr), _p(tu), _p(tv))
    return tu, tv


def points_vorticity(sx, sy, ss, sr, tx, ty, *, device=0, tw=None):
    """points_affect_points_vorticity (src/InfluenceVort.h:47-126): accumulates the unscaled
    Gaussian vorticity sum into tw (fp64)."""
    sx, sy, ss, sr, tx, ty = (_arr(a, np.float64) for a in (sx, sy, ss, sr, tx, ty))
    tw = _out(tw, len(tx), np.float64)
    env = make_env("gauss", "fp64", "vel", device)
    call("o2d_hip_points_vorticity_d", ctypes.byref(env), len(sx), _p(sx), _p(sy), _p(ss),
         _p(sr), len(tx), _p(tx), _p(ty), _p(tw))
    return tw


def nearest_distances(nx, ny, idx, tx, ty, *, device=0):
    """get_nearest_distances (src/Reflect.h:792-826): distance from each target to its nearest
    panel or node."""
    nx, ny, tx, ty = (_arr(a, np.float64) for a in (nx, ny, tx, ty))
    idx = _arr(idx, np.uint32)
    dist = np.zeros(len(tx))
    env = make_env("v2", "fp64", "vel", device)
    call("o2d_hip_nearest_distances_d", ctypes.byref(env), len(nx), _p(nx), _p(ny),
         len(idx) // 2, _p(idx), len(tx), _p(tx), _p(ty), _p(dist))
    return dist


def reflect(nx, ny, idx, pnx, pny, tx, ty, *, device=0):
    """reflect_panp2 (src/Reflect.h:167-311) on copies of tx, ty: returns (x, y, reflected)."""
    nx, ny, pnx, pny = (_arr(a, np.float64) for a in (nx, ny, pnx, pny))
    tx, ty = _arr(tx, np.float64).copy(), _arr(ty, np.float64).copy()
    idx = _arr(idx, np.uint32)
    n = ctypes.c_int64()
    env = make_env("v2", "fp64", "vel", device)
    call("o2d_hip_reflect_d", ctypes.byref(env), len(nx), _p(nx), _p(ny), len(idx) // 2, _p(idx),
         _p(pnx), _p(pny), len(tx), _p(tx), _p(ty), ctypes.byref(n))
    return tx, ty, n.value


def clear_inner(method, nx, ny, idx, pnx, pny, cutoff_mult, ips, tx, ty, ts=None, tr=None, *,
                device=0):
    """clear_inner_panp2 (src/Reflect.h:487-727) on copies of the targets: returns dict(x, y,
    str, circ_removed, num_cropped). tr=None makes the targets field points."""
    nx, ny, pnx, pny = (_arr(a, np.float64) for a in (nx, ny, pnx, pny))
    tx, ty = _arr(tx, np.float64).copy(), _arr(ty, np.float64).copy()
    ts = None if ts is None else _arr(ts, np.float64).copy()
    tr = None if tr is None else _arr(tr, np.float64)
    idx = _arr(idx, np.uint32)
    circ, n = ctypes.c_double(), ctypes.c_int64()
    env = make_env("v2", "fp64", "vel", device)
    call("o2d_hip_clear_inner_d", ctypes.byref(env), int(method), float(cutoff_mult), float(ips),
         len(nx), _p(nx), _p(ny), len(idx) // 2, _p(idx), _p(pnx), _p(pny), len(tx), _p(tx),
         _p(ty), _p(ts), _p(tr), ctypes.byref(circ), ctypes.byref(n))
    return dict(x=tx, y=ty, str=ts, circ_removed=circ.value, num_cropped=n.value)

