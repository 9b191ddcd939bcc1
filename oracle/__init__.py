"""ctypes bindings to the CPU checkers. TEST INFRASTRUCTURE ONLY.

Two libraries, both built by ``oracle/Makefile``:

* ``build/libo2d_oracle.so`` -- the fp64 C restatement (``o2d_oracle.c``), the parity oracle.
* ``_ref/libo2d_ref.so``     -- the reference's own templates compiled from its sources
  (``ref_harness.cpp``) with its compiled-in core (Vatistas n=2); used to write golden
  fixtures, to validate the restatement, and as the "reference" CPU baseline. Absent where the
  reference tree was never available.
* ``_ref/libo2d_ref_<core>.so`` -- the same harness built with the reference's core define
  switched to ``rm``, ``gauss``, ``wl`` or ``v3`` (see ``oracle/Makefile``).
* ``_ref/libo2d_ref_dropin.so`` -- the reference's own ``points_affect_points`` built with
  ``EXTERNAL_VEL_SOLVE`` and linked to the HIP library's hook (``ref_dropin.cpp``).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import this
module, always as the checker or the timed CPU baseline, never as the thing measured.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "build", "libo2d_oracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libo2d_ref.so")
DROPIN_SO = os.path.join(_HERE, "_ref", "libo2d_ref_dropin.so")

_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_vp = ctypes.c_void_p

CORES = {"rm": 1, "gauss": 2, "wl": 3, "v2": 4, "v3": 5}
RESULTS = {"vel": 1, "vort": 4, "shear": 5}


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run make -C oracle)")
        L = ctypes.CDLL(ORACLE_SO)
        L.oc_points_on_points.argtypes = [_int, _int, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp,
                                          _vp, _vp, _vp, _vp, _vp]
        L.oc_panels_on_points.argtypes = [_i64, _vp, _vp, _i64, _vp, _vp, _vp, _i64, _vp, _vp,
                                          _vp, _vp]
        L.oc_points_on_panels.argtypes = [_i64, _vp, _vp, _vp, _i64, _vp, _vp, _i64, _vp, _vp,
                                          _vp]
        L.oc_panel_geometry.argtypes = [_i64, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp]
        L.oc_body_volume.argtypes = [_i64, _vp, _vp, _i64, _vp, _vp, _vp, _vp]
        L.oc_unit_rot_strengths.argtypes = [_i64, _vp, _vp, _i64, _vp, _dbl, _dbl, _vp, _vp]
        L.oc_bem_assemble.argtypes = [_i64, _vp, _vp, _i64, _vp, _int, _int, _vp, _vp, _dbl,
                                      _vp, _i64]
        L.oc_bem_assemble_pair.argtypes = [_i64, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _i64, _vp,
                                           _int, _int, _int, _vp, _vp, _vp, _i64]
        L.oc_vels_to_rhs.argtypes = [_i64, _int, _vp, _vp, _vp, _vp, _vp, _vp]
        L.oc_finalize.argtypes = [_i64, _vp, _vp, _dbl, _dbl]
        L.oc_rk_combine.argtypes = [_i64, _int, _vp, _vp, _vp, _dbl, _vp, _vp]
        L.oc_points_on_points_adjoint.argtypes = [_int, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp,
                                                  _vp, _vp, _vp, _vp, _vp, _vp]
        L.oc_morton_keys.argtypes = [_i64, _vp, _vp, _dbl, _dbl, _dbl, _vp]
        L.oc_points_vorticity.argtypes = [_i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp]
        L.oc_nearest_distances.argtypes = [_i64, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _vp]
        L.oc_reflect.argtypes = [_i64, _vp, _vp, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp]
        L.oc_cut_table.argtypes = [_dbl, _int, _vp, _vp, _vp]
        L.oc_clear_inner.argtypes = [_int, _i64, _vp, _vp, _i64, _vp, _vp, _vp, _dbl, _dbl, _i64,
                                     _vp, _vp, _vp, _vp, _int, _vp, _vp, _vp]
        _lib = L
    return _lib


def ref_so(core=None):
    """Path of the reference build carrying ``core`` (None or "v2": the headers as they are)."""
    if core in (None, "v2"):
        return REF_SO
    if core not in CORES:
        raise ValueError(f"unknown core {core!r}")
    return os.path.join(_HERE, "_ref", f"libo2d_ref_{core}.so")


def have_ref(core=None):
    return os.path.exists(ref_so(core))


_refs = {}


def ref(core=None):
    """ctypes handle of the reference build carrying ``core``; its compiled-in core is checked."""
    key = "v2" if core is None else core
    if key not in _refs:
        path = ref_so(core)
        if not os.path.exists(path):
            raise RuntimeError(f"reference build missing: {path}")
        R = ctypes.CDLL(path)
        pp = [_i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _int, _vp, _vp, _vp, _vp]
        R.ref_points_on_points_d.argtypes = pp
        R.ref_points_on_points_f.argtypes = pp
        pa = [_i64, _vp, _vp, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp]
        R.ref_panels_on_points_d.argtypes = pa
        R.ref_panels_on_points_f.argtypes = pa
        pb = [_i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _i64, _vp, _vp, _vp]
        R.ref_points_on_panels_d.argtypes = pb
        R.ref_points_on_panels_f.argtypes = pb
        pc = [_i64, _vp, _vp, _i64, _vp, _int, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        R.ref_bem_coeffs_d.argtypes = pc
        R.ref_bem_coeffs_f.argtypes = pc
        R.ref_rhs_panels_d.argtypes = [_i64, _vp, _vp, _i64, _vp, _int, _vp, _vp, _vp, _vp]
        R.ref_points_move2_d.argtypes = [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _dbl, _dbl, _dbl]
        R.ref_points_move3_d.argtypes = [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _dbl,
                                         _dbl, _dbl, _dbl]
        R.ref_finalize_d.argtypes = [_i64, _vp, _vp, _dbl, _dbl]
        R.ref_points_vorticity_d.argtypes = [_i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp]
        R.ref_points_vorticity_f.argtypes = [_i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp]
        R.ref_nearest_distances_d.argtypes = [_i64, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _vp]
        R.ref_panel_normals_d.argtypes = [_i64, _vp, _vp, _i64, _vp, _vp, _vp]
        R.ref_reflect_d.argtypes = [_i64, _vp, _vp, _i64, _vp, _i64, _vp, _vp]
        R.ref_clear_inner_d.argtypes = [_int, _i64, _vp, _vp, _i64, _vp, _dbl, _dbl, _i64, _vp,
                                        _vp, _vp, _vp, _vp]
        R.ref_cut_table_d.argtypes = [_int, _vp, _vp, _vp]
        R.ref_bem_block_d.argtypes = [_i64, _vp, _vp, _i64, _vp, _int, _i64, _vp, _vp, _i64, _vp,
                                      _int, _int, _vp, _vp, _vp, _vp, _vp]
        R.ref_write_points_vtk_f.argtypes = [ctypes.c_char_p, _i64, _vp, _vp, _vp, _vp, _vp, _vp,
                                             _vp, _int, _i64, _i64, _dbl, ctypes.c_char_p, _i64]
        R.ref_write_panels_vtk_f.argtypes = [ctypes.c_char_p, _i64, _vp, _vp, _i64, _vp, _vp, _vp,
                                             _i64, _i64, _dbl, ctypes.c_char_p, _i64]
        R.ref_body_motion_d.argtypes = [_i64, _vp, _vp, _i64, _vp, ctypes.c_char_p,
                                        ctypes.c_char_p, ctypes.c_char_p, _dbl, _dbl, _vp, _vp,
                                        _vp, _vp, _vp, _vp, _vp]
        R.ref_core_id.argtypes = []
        got = R.ref_core_id()
        if got != CORES[key]:
            raise RuntimeError(f"{path} carries core id {got}, expected {CORES[key]} ({key})")
        _refs[key] = R
    return _refs[key]


_dropin = None


def have_dropin():
    return os.path.exists(DROPIN_SO)


def dropin():
    global _dropin
    if _dropin is None:
        if not have_dropin():
            raise RuntimeError(f"drop-in build missing: {DROPIN_SO}")
        D = ctypes.CDLL(DROPIN_SO)
        D.dropin_find_vels_f.argtypes = [_int, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _int,
                                         _dbl, _dbl, _vp, _vp]
        _dropin = D
    return _dropin


class _Quiet:
    """Silence fd 1 while the reference (which logs every call to stdout) runs."""

    def __enter__(self):
        import sys
        sys.stdout.flush()
        self._saved = os.dup(1)
        self._null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(self._null, 1)
        return self

    def __exit__(self, *exc):
        ctypes.CDLL(None).fflush(None)  # C stdio buffers still hold the reference's log lines
        os.dup2(self._saved, 1)
        os.close(self._saved)
        os.close(self._null)
        return False


quiet = _Quiet


# ----------------------------------------------------------------------------- restatement

def points_on_points(core, sx, sy, ss, sr, tx, ty, tr=None, restype="vel"):
    """fp64 unscaled sums (u, v[, w|e]) at targets; tr=None means point (inert) targets."""
    sx, sy, ss, sr, tx, ty = map(_f64, (sx, sy, ss, sr, tx, ty))
    tr = None if tr is None else _f64(tr)
    nt = len(tx)
    tu, tv = np.zeros(nt), np.zeros(nt)
    tw = np.zeros(nt) if restype == "vort" else None
    te = np.zeros(nt) if restype == "shear" else None
    rc = lib().oc_points_on_points(CORES[core], RESULTS[restype], len(sx), _p(sx), _p(sy), _p(ss),
                                   _p(sr), nt, _p(tx), _p(ty), _p(tr), _p(tu), _p(tv), _p(tw),
                                   _p(te))
    assert rc == 0, rc
    if restype == "vort":
        return tu, tv, tw
    if restype == "shear":
        return tu, tv, te
    return tu, tv


def panels_on_points(nx, ny, idx, vs, ss, tx, ty):
    nx, ny, vs, tx, ty = map(_f64, (nx, ny, vs, tx, ty))
    ss = None if ss is None else _f64(ss)
    idx = _u32(idx)
    nt = len(tx)
    tu, tv = np.zeros(nt), np.zeros(nt)
    rc = lib().oc_panels_on_points(len(nx), _p(nx), _p(ny), len(idx) // 2, _p(idx), _p(vs), _p(ss),
                                   nt, _p(tx), _p(ty), _p(tu), _p(tv))
    assert rc == 0, rc
    return tu, tv


def points_on_panels(sx, sy, ss, nx, ny, idx):
    sx, sy, ss, nx, ny = map(_f64, (sx, sy, ss, nx, ny))
    idx = _u32(idx)
    np_ = len(idx) // 2
    pu, pv = np.zeros(np_), np.zeros(np_)
    rc = lib().oc_points_on_panels(len(sx), _p(sx), _p(sy), _p(ss), len(nx), _p(nx), _p(ny), np_,
                                   _p(idx), _p(pu), _p(pv))
    assert rc == 0, rc
    return pu, pv


def panel_geometry(nx, ny, idx):
    nx, ny = _f64(nx), _f64(ny)
    idx = _u32(idx)
    np_ = len(idx) // 2
    out = [np.zeros(np_) for _ in range(5)]
    rc = lib().oc_panel_geometry(len(nx), _p(nx), _p(ny), np_, _p(idx), *[_p(o) for o in out])
    assert rc == 0, rc
    return tuple(out)  # tang_x, tang_y, norm_x, norm_y, area


def body_volume(nx, ny, idx):
    nx, ny = _f64(nx), _f64(ny)
    idx = _u32(idx)
    vol, cx, cy = _dbl(), _dbl(), _dbl()
    rc = lib().oc_body_volume(len(nx), _p(nx), _p(ny), len(idx) // 2, _p(idx), ctypes.byref(vol),
                              ctypes.byref(cx), ctypes.byref(cy))
    assert rc == 0, rc
    return vol.value, cx.value, cy.value


def unit_rot_strengths(nx, ny, idx, cx, cy):
    nx, ny = _f64(nx), _f64(ny)
    idx = _u32(idx)
    np_ = len(idx) // 2
    vs, ss = np.zeros(np_), np.zeros(np_)
    rc = lib().oc_unit_rot_strengths(len(nx), _p(nx), _p(ny), np_, _p(idx), cx, cy, _p(vs), _p(ss))
    assert rc == 0, rc
    return vs, ss


def bem_assemble(nx, ny, idx, nunk, augmented, urot_u=None, urot_v=None, vol=0.0):
    nx, ny = _f64(nx), _f64(ny)
    idx = _u32(idx)
    np_ = len(idx) // 2
    n = np_ * nunk + (1 if augmented else 0)
    A = np.zeros(n * n)
    uu = None if urot_u is None else _f64(urot_u)
    uv = None if urot_v is None else _f64(urot_v)
    rc = lib().oc_bem_assemble(len(nx), _p(nx), _p(ny), np_, _p(idx), nunk, int(augmented),
                               _p(uu), _p(uv), vol, _p(A), n)
    assert rc == 0, rc
    return A.reshape(n, n).T  # column-major storage -> A[row, col]


def bem_assemble_pair(snx, sny, sidx, tnx, tny, tidx, nunk, src_aug, targ_aug, urot_u=None,
                      urot_v=None):
    """Block of source surface s on a different target surface t: rows = t, columns = s."""
    snx, sny, tnx, tny = map(_f64, (snx, sny, tnx, tny))
    sidx, tidx = _u32(sidx), _u32(tidx)
    snp, tnp = len(sidx) // 2, len(tidx) // 2
    nr = tnp * nunk + (1 if targ_aug else 0)
    ncol = snp * nunk + (1 if src_aug else 0)
    uu = _f64(urot_u if urot_u is not None else np.zeros(tnp))
    uv = _f64(urot_v if urot_v is not None else np.zeros(tnp))
    A = np.zeros(nr * ncol)
    rc = lib().oc_bem_assemble_pair(len(snx), _p(snx), _p(sny), snp, _p(sidx), len(tnx), _p(tnx),
                                    _p(tny), tnp, _p(tidx), int(nunk), int(bool(src_aug)),
                                    int(bool(targ_aug)), _p(uu), _p(uv), _p(A), nr)
    assert rc == 0, rc
    return A.reshape(ncol, nr).T.copy()


def vels_to_rhs(pu, pv, tang_x, tang_y, nunk, norm_bc=None):
    pu, pv, tang_x, tang_y = map(_f64, (pu, pv, tang_x, tang_y))
    nb = None if norm_bc is None else _f64(norm_bc)
    rhs = np.zeros(len(pu) * nunk)
    rc = lib().oc_vels_to_rhs(len(pu), nunk, _p(pu), _p(pv), _p(tang_x), _p(tang_y), _p(nb), _p(rhs))
    assert rc == 0, rc
    return rhs


def finalize(u, v, fs=(0.0, 0.0)):
    u, v = _f64(u).copy(), _f64(v).copy()
    lib().oc_finalize(len(u), _p(u), _p(v), float(fs[0]), float(fs[1]))
    return u, v


def rk_combine(x, y, us, vs, w, dt):
    x, y = _f64(x).copy(), _f64(y).copy()
    us = [_f64(a) for a in us]
    vs = [_f64(a) for a in vs]
    w = _f64(w)
    up = (_vp * len(us))(*[a.ctypes.data for a in us])
    vp = (_vp * len(vs))(*[a.ctypes.data for a in vs])
    rc = lib().oc_rk_combine(len(x), len(us), _p(w), up, vp, float(dt), _p(x), _p(y))
    assert rc == 0, rc
    return x, y


def points_on_points_adjoint(core, sx, sy, ss, sr, tx, ty, gu, gv):
    sx, sy, ss, sr, tx, ty, gu, gv = map(_f64, (sx, sy, ss, sr, tx, ty, gu, gv))
    ns = len(sx)
    out = [np.zeros(ns) for _ in range(4)]
    rc = lib().oc_points_on_points_adjoint(CORES[core], ns, _p(sx), _p(sy), _p(ss), _p(sr), len(tx),
                                           _p(tx), _p(ty), _p(gu), _p(gv), *[_p(o) for o in out])
    assert rc == 0, rc
    return tuple(out)  # gx, gy, gs, gr


def morton_keys(x, y, x0, y0, inv_extent):
    x, y = _f64(x), _f64(y)
    keys = np.zeros(len(x), dtype=np.uint32)
    lib().oc_morton_keys(len(x), _p(x), _p(y), float(x0), float(y0), float(inv_extent), _p(keys))
    return keys


def points_vorticity(sx, sy, ss, sr, tx, ty):
    """Unscaled Gaussian vorticity sum at the targets (src/InfluenceVort.h:47-126)."""
    sx, sy, ss, sr, tx, ty = map(_f64, (sx, sy, ss, sr, tx, ty))
    tw = np.zeros(len(tx))
    rc = lib().oc_points_vorticity(len(sx), _p(sx), _p(sy), _p(ss), _p(sr), len(tx), _p(tx),
                                   _p(ty), _p(tw))
    assert rc == 0, rc
    return tw


def nearest_distances(nx, ny, idx, tx, ty):
    nx, ny, tx, ty = map(_f64, (nx, ny, tx, ty))
    idx = _u32(idx)
    d = np.zeros(len(tx))
    rc = lib().oc_nearest_distances(len(nx), _p(nx), _p(ny), len(idx) // 2, _p(idx), len(tx),
                                    _p(tx), _p(ty), _p(d))
    assert rc == 0, rc
    return d


def reflect(nx, ny, idx, pnx, pny, tx, ty):
    """Returns (new x, new y, number reflected)."""
    nx, ny, pnx, pny = map(_f64, (nx, ny, pnx, pny))
    tx, ty = _f64(tx).copy(), _f64(ty).copy()
    idx = _u32(idx)
    n = _i64()
    rc = lib().oc_reflect(len(nx), _p(nx), _p(ny), len(idx) // 2, _p(idx), _p(pnx), _p(pny),
                          len(tx), _p(tx), _p(ty), ctypes.byref(n))
    assert rc == 0, rc
    return tx, ty, n.value


def cut_table(dx=0.1):
    pos, sm, ds = (np.zeros(256) for _ in range(3))
    rows = lib().oc_cut_table(float(dx), 256, _p(pos), _p(sm), _p(ds))
    assert rows > 0, rows
    return pos[:rows], sm[:rows], ds[:rows]


def clear_inner(method, nx, ny, idx, pnx, pny, cutoff_mult, ips, tx, ty, ts=None, tr=None,
                max_passes=64):
    """Returns dict(x, y, str, circ_removed, num_cropped, unsettled)."""
    nx, ny, pnx, pny = map(_f64, (nx, ny, pnx, pny))
    tx, ty = _f64(tx).copy(), _f64(ty).copy()
    ts = None if ts is None else _f64(ts).copy()
    tr = None if tr is None else _f64(tr)
    idx = _u32(idx)
    circ, ncrop, nuns = _dbl(), _i64(), _i64()
    rc = lib().oc_clear_inner(int(method), len(nx), _p(nx), _p(ny), len(idx) // 2, _p(idx),
                              _p(pnx), _p(pny), float(cutoff_mult), float(ips), len(tx), _p(tx),
                              _p(ty), _p(ts), _p(tr), int(max_passes), ctypes.byref(circ),
                              ctypes.byref(ncrop), ctypes.byref(nuns))
    assert rc == 0, rc
    return dict(x=tx, y=ty, str=ts, circ_removed=circ.value, num_cropped=ncrop.value,
                unsettled=nuns.value)


# ----------------------------------------------------------------------------- reference build

def ref_points_on_points(sx, sy, ss, sr, tx, ty, tr=None, restype="vel", dtype=np.float64,
                         core=None):
    """Reference points_affect_points at S=dtype (A=double), with the reference build carrying
    ``core`` (None: the compiled default, Vatistas n=2)."""
    c = lambda a: np.ascontiguousarray(a, dtype=dtype)
    sx, sy, ss, sr, tx, ty = map(c, (sx, sy, ss, sr, tx, ty))
    tr = None if tr is None else c(tr)
    nt = len(tx)
    tu, tv = np.zeros(nt, dtype), np.zeros(nt, dtype)
    tw = np.zeros(nt, dtype) if restype == "vort" else None
    te = np.zeros(nt, dtype) if restype == "shear" else None
    R = ref(core)
    fn = R.ref_points_on_points_d if dtype == np.float64 else R.ref_points_on_points_f
    with quiet():
        rc = fn(len(sx), _p(sx), _p(sy), _p(ss), _p(sr), nt, _p(tx), _p(ty), _p(tr),
                RESULTS[restype], _p(tu), _p(tv), _p(tw), _p(te))
    assert rc == 0, rc
    if restype == "vort":
        return tu, tv, tw
    if restype == "shear":
        return tu, tv, te
    return tu, tv


def ref_panels_on_points(nx, ny, idx, vs, ss, tx, ty, dtype=np.float64):
    c = lambda a: np.ascontiguousarray(a, dtype=dtype)
    nx, ny = np.ascontiguousarray(nx, np.float32), np.ascontiguousarray(ny, np.float32)
    vs, tx, ty = map(c, (vs, tx, ty))
    ss = None if ss is None else c(ss)
    idx = _u32(idx)
    nt = len(tx)
    tu, tv = np.zeros(nt, dtype), np.zeros(nt, dtype)
    fn = ref().ref_panels_on_points_d if dtype == np.float64 else ref().ref_panels_on_points_f
    with quiet():
        rc = fn(len(nx), _p(nx), _p(ny), len(idx) // 2, _p(idx), _p(vs), _p(ss), nt, _p(tx), _p(ty),
                _p(tu), _p(tv))
    assert rc == 0, rc
    return tu, tv


def ref_points_on_panels(sx, sy, ss, sr, nx, ny, idx, dtype=np.float64):
    c = lambda a: np.ascontiguousarray(a, dtype=dtype)
    sx, sy, ss, sr = map(c, (sx, sy, ss, sr))
    nx, ny = np.ascontiguousarray(nx, np.float32), np.ascontiguousarray(ny, np.float32)
    idx = _u32(idx)
    np_ = len(idx) // 2
    pu, pv = np.zeros(np_, dtype), np.zeros(np_, dtype)
    fn = ref().ref_points_on_panels_d if dtype == np.float64 else ref().ref_points_on_panels_f
    with quiet():
        rc = fn(len(sx), _p(sx), _p(sy), _p(ss), _p(sr), len(nx), _p(nx), _p(ny), np_, _p(idx),
                _p(pu), _p(pv))
    assert rc == 0, rc
    return pu, pv


def ref_bem_coeffs(nx, ny, idx, nunk, body_kind, dtype=np.float64):
    """Reference BEM self-block (augmented when the body calls for it), plus the inputs the
    restatement needs: tangents, lengths, volume, unit-rotation panel velocities."""
    nx, ny = np.ascontiguousarray(nx, np.float32), np.ascontiguousarray(ny, np.float32)
    idx = _u32(idx)
    np_ = len(idx) // 2
    nmax = np_ * nunk + 1
    A = np.zeros(nmax * nmax, dtype)
    nr, nc = _i64(), _i64()
    tx, ty, ar, uu, uv = (np.zeros(np_, dtype) for _ in range(5))
    vol = np.zeros(1, dtype)
    fn = ref().ref_bem_coeffs_d if dtype == np.float64 else ref().ref_bem_coeffs_f
    with quiet():
        rc = fn(len(nx), _p(nx), _p(ny), np_, _p(idx), nunk, body_kind, _p(A), ctypes.byref(nr),
                ctypes.byref(nc), _p(tx), _p(ty), _p(ar), _p(vol), _p(uu), _p(uv))
    assert rc == 0, rc
    n = nr.value
    return dict(A=A[: n * n].reshape(n, n).T.copy(), tang_x=tx, tang_y=ty, area=ar,
                vol=float(vol[0]), urot_u=uu, urot_v=uv)


def ref_rhs_panels(nx, ny, idx, nunk, pu, pv, norm_bc):
    nx, ny = np.ascontiguousarray(nx, np.float32), np.ascontiguousarray(ny, np.float32)
    idx = _u32(idx)
    pu, pv = _f64(pu), _f64(pv)
    nb = np.ascontiguousarray(norm_bc, np.float32)
    np_ = len(idx) // 2
    rhs = np.zeros(np_ * nunk)
    with quiet():
        rc = ref().ref_rhs_panels_d(len(nx), _p(nx), _p(ny), np_, _p(idx), nunk, _p(pu), _p(pv),
                                    _p(nb), _p(rhs))
    assert rc == 0, rc
    return rhs


def ref_move(x, y, us, vs, w, dt):
    x, y = _f64(x).copy(), _f64(y).copy()
    us = [_f64(a) for a in us]
    vs = [_f64(a) for a in vs]
    with quiet():
        if len(us) == 2:
            rc = ref().ref_points_move2_d(len(x), _p(x), _p(y), _p(us[0]), _p(vs[0]), _p(us[1]),
                                          _p(vs[1]), float(dt), float(w[0]), float(w[1]))
        else:
            rc = ref().ref_points_move3_d(len(x), _p(x), _p(y), _p(us[0]), _p(vs[0]), _p(us[1]),
                                          _p(vs[1]), _p(us[2]), _p(vs[2]), float(dt), float(w[0]),
                                          float(w[1]), float(w[2]))
    assert rc == 0, rc
    return x, y


def ref_finalize(u, v, fs):
    u, v = _f64(u).copy(), _f64(v).copy()
    with quiet():
        ref().ref_finalize_d(len(u), _p(u), _p(v), float(fs[0]), float(fs[1]))
    return u, v


def _f32nodes(nx, ny):
    return np.ascontiguousarray(nx, np.float32), np.ascontiguousarray(ny, np.float32)


def ref_nearest_distances(nx, ny, idx, tx, ty):
    nx, ny = _f32nodes(nx, ny)
    idx = _u32(idx)
    tx, ty = _f64(tx), _f64(ty)
    d = np.zeros(len(tx))
    with quiet():
        rc = ref().ref_nearest_distances_d(len(nx), _p(nx), _p(ny), len(idx) // 2, _p(idx),
                                           len(tx), _p(tx), _p(ty), _p(d))
    assert rc == 0, rc
    return d


def ref_panel_normals(nx, ny, idx):
    nx, ny = _f32nodes(nx, ny)
    idx = _u32(idx)
    pnx, pny = np.zeros(len(idx) // 2), np.zeros(len(idx) // 2)
    with quiet():
        rc = ref().ref_panel_normals_d(len(nx), _p(nx), _p(ny), len(idx) // 2, _p(idx), _p(pnx),
                                       _p(pny))
    assert rc == 0, rc
    return pnx, pny


def ref_reflect(nx, ny, idx, tx, ty):
    nx, ny = _f32nodes(nx, ny)
    idx = _u32(idx)
    tx, ty = _f64(tx).copy(), _f64(ty).copy()
    with quiet():
        rc = ref().ref_reflect_d(len(nx), _p(nx), _p(ny), len(idx) // 2, _p(idx), len(tx), _p(tx),
                                 _p(ty))
    assert rc == 0, rc
    return tx, ty


def ref_clear_inner(method, nx, ny, idx, cutoff_mult, ips, tx, ty, ts=None, tr=None):
    """Reference clear_inner_panp2<double>. Method 0 needs particles (ts and tr)."""
    assert method == 1 or (ts is not None and tr is not None)
    nx, ny = _f32nodes(nx, ny)
    idx = _u32(idx)
    tx, ty = _f64(tx).copy(), _f64(ty).copy()
    ts = None if ts is None else _f64(ts).copy()
    tr = None if tr is None else _f64(tr)
    circ = np.zeros(1)
    with quiet():
        rc = ref().ref_clear_inner_d(int(method), len(nx), _p(nx), _p(ny), len(idx) // 2,
                                     _p(idx), float(cutoff_mult), float(ips), len(tx), _p(tx),
                                     _p(ty), _p(ts), _p(tr), _p(circ))
    assert rc == 0, rc
    return dict(x=tx, y=ty, str=ts, circ_removed=float(circ[0]))


def ref_cut_table():
    pos, sm, ds = (np.zeros(256) for _ in range(3))
    rows = ref().ref_cut_table_d(256, _p(pos), _p(sm), _p(ds))
    assert rows > 0, rows
    return pos[:rows], sm[:rows], ds[:rows]


def ref_points_vorticity(sx, sy, ss, sr, tx, ty, dtype=np.float64):
    c = lambda a: np.ascontiguousarray(a, dtype=dtype)  # noqa: E731
    sx, sy, ss, sr, tx, ty = map(c, (sx, sy, ss, sr, tx, ty))
    tw = np.zeros(len(tx), dtype)
    fn = ref().ref_points_vorticity_d if dtype == np.float64 else ref().ref_points_vorticity_f
    with quiet():
        rc = fn(len(sx), _p(sx), _p(sy), _p(ss), _p(sr), len(tx), _p(tx), _p(ty), _p(tw))
    assert rc == 0, rc
    return tw


def ref_write_points_vtk(directory, x, y, s, r, u, v, w=None, inert=False, index=0, frameno=0,
                         time=0.0):
    """Reference Points<float>::write_vtk inside `directory`; returns the file name."""
    c = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)  # noqa: E731
    x, y, s, r, u, v, w = map(c, (x, y, s, r, u, v, w))
    name = ctypes.create_string_buffer(256)
    with quiet():
        rc = ref().ref_write_points_vtk_f(str(directory).encode(), len(x), _p(x), _p(y), _p(s),
                                          _p(r), _p(u), _p(v), _p(w), int(bool(inert)), index,
                                          frameno, float(time), name, 256)
    assert rc == 0, rc
    return name.value.decode()


def ref_write_panels_vtk(directory, nx, ny, idx, vs, ss, index=0, frameno=0, time=0.0):
    """Reference Surfaces<float>::write_vtk inside `directory`; returns the file name."""
    nx, ny = _f32nodes(nx, ny)
    idx = _u32(idx)
    vs, ss = np.ascontiguousarray(vs, np.float32), np.ascontiguousarray(ss, np.float32)
    name = ctypes.create_string_buffer(256)
    with quiet():
        rc = ref().ref_write_panels_vtk_f(str(directory).encode(), len(nx), _p(nx), _p(ny),
                                          len(idx) // 2, _p(idx), _p(vs), _p(ss), index, frameno,
                                          float(time), name, 256)
    assert rc == 0, rc
    return name.value.decode()


def ref_bem_block(snx, sny, sidx, skind, tnx, tny, tidx, tkind, nunk):
    """Reference block of one surface on another (solve_bem's steps): dict(A, urot_u, urot_v)."""
    snx, sny = _f32nodes(snx, sny)
    tnx, tny = _f32nodes(tnx, tny)
    sidx, tidx = _u32(sidx), _u32(tidx)
    snp, tnp = len(sidx) // 2, len(tidx) // 2
    A = np.zeros((snp * nunk + 1) * (tnp * nunk + 1))
    nr, nc = _i64(), _i64()
    uu, uv = np.zeros(tnp), np.zeros(tnp)
    with quiet():
        rc = ref().ref_bem_block_d(len(snx), _p(snx), _p(sny), snp, _p(sidx), int(skind), len(tnx),
                                   _p(tnx), _p(tny), tnp, _p(tidx), int(tkind), int(nunk), _p(A),
                                   ctypes.byref(nr), ctypes.byref(nc), _p(uu), _p(uv))
    assert rc == 0, rc
    r, c = nr.value, nc.value
    return dict(A=A[: r * c].reshape(c, r).T.copy(), urot_u=uu, urot_v=uv)


def ref_body_motion(nx, ny, idx, xexpr, yexpr, rexpr, t, factor=-1.0):
    """Reference body moving by tinyexpr expressions of t: its body-bound Surfaces transformed to
    time t and the body's motion added (times factor) to zeroed panel velocities. Returns
    dict(x, y, tx, ty, pu, pv, body=[x, y, u, v, theta, omega])."""
    nx, ny = _f32nodes(nx, ny)
    idx = _u32(idx)
    nn, np_ = len(nx), len(idx) // 2
    out = {k: np.zeros(nn if k in ("x", "y") else np_) for k in ("x", "y", "tx", "ty", "pu", "pv")}
    body = np.zeros(6)
    with quiet():
        rc = ref().ref_body_motion_d(nn, _p(nx), _p(ny), np_, _p(idx), xexpr.encode(),
                                     yexpr.encode(), rexpr.encode(), float(t), float(factor),
                                     *[_p(out[k]) for k in ("x", "y", "tx", "ty", "pu", "pv")],
                                     _p(body))
    assert rc == 0, rc
    out["body"] = body
    return out


def dropin_find_vels(sets, tx, ty, tr=None, finalize=False, fs=(0.0, 0.0)):
    """The reference's find_vels particle loop (src/Convection.h:185-213) with an external
    environment: every (x, y, strength, radius) set in ``sets`` acts on the targets through
    src/Influence.h:74-98, i.e. through the HIP library's external_vel_solver_f_/_tr_f_.
    float32 storage, as the application stores. Returns (u, v) as float32."""
    f = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    sets = [tuple(map(f, st)) for st in sets]
    k = len(sets)
    ns = np.array([len(st[0]) for st in sets], dtype=np.int64)
    ptrs = [(ctypes.c_void_p * max(k, 1))(*[st[c].ctypes.data for st in sets]) for c in range(4)]
    tx, ty = f(tx), f(ty)
    tr = None if tr is None else f(tr)
    nt = len(tx)
    tu, tv = np.zeros(nt, np.float32), np.zeros(nt, np.float32)
    with quiet():
        rc = dropin().dropin_find_vels_f(k, _p(ns), *[ctypes.cast(p, _vp) for p in ptrs], nt,
                                         _p(tx), _p(ty), _p(tr), int(bool(finalize)),
                                         float(fs[0]), float(fs[1]), _p(tu), _p(tv))
    assert rc == 0, rc
    return tu, tv
