"""Device-pointer binding of the C ABI for torch tensors already resident in HBM.

Every call enqueues on torch's current HIP stream (the stream handle is passed through
o2d_env.stream) and returns without synchronising, exactly like a torch op. Tensors must be
contiguous float64 on the same device (uint32 panel indices travel as int32 tensors).
"""
import ctypes

import torch

from ._cabi import VELANDSHEAR, VELANDVORT, VELONLY, TreeOpts, call, make_env

__all__ = ["points_on_points", "panels_on_points", "points_on_panels", "bem_assemble",
           "bem_factor", "bem_solve", "rk_move", "points_on_points_adjoint", "morton_order",
           "points_on_points_tree", "points_vorticity", "nearest_distances", "reflect", "clear_inner", "env_for"]


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _check(*ts, dtype=torch.float64):
    dev = None
    for t in ts:
        if t is None:
            continue
        if not t.is_cuda or t.dtype != dtype or not t.is_contiguous():
            raise TypeError(f"expected a contiguous {dtype} device tensor, got {t.dtype} "
                            f"on {t.device} (contiguous={t.is_contiguous()})")
        if dev is None:
            dev = t.device
        elif t.device != dev:
            raise ValueError("tensors live on different devices")
    return dev


def env_for(device, core="v2", precision="fp64", results="vel"):
    idx = device.index if device.index is not None else torch.cuda.current_device()
    stream = torch.cuda.current_stream(device).cuda_stream
    return make_env(core, precision, results, idx, stream)


def points_on_points(sx, sy, ss, sr, tx, ty, tr=None, *, core="v2", precision="fp64",
                     results="vel", tu=None, tv=None, aux=None):
    dev = _check(sx, sy, ss, sr, tx, ty, tr, tu, tv, aux)
    nt = tx.numel()
    tu = torch.zeros(nt, dtype=torch.float64, device=dev) if tu is None else tu
    tv = torch.zeros(nt, dtype=torch.float64, device=dev) if tv is None else tv
    env = env_for(dev, core, precision, results)
    if env.results != VELONLY and aux is None:
        aux = torch.zeros(nt, dtype=torch.float64, device=dev)
    tw = aux if env.results == VELANDVORT else None
    te = aux if env.results == VELANDSHEAR else None
    call("o2d_hip_points_on_points_dev_d", ctypes.byref(env), sx.numel(), _ptr(sx), _ptr(sy),
         _ptr(ss), _ptr(sr), nt, _ptr(tx), _ptr(ty), _ptr(tr), _ptr(tu), _ptr(tv), _ptr(tw),
         _ptr(te))
    return (tu, tv, aux) if env.results != VELONLY else (tu, tv)


def panels_on_points(nx, ny, idx, vs, ss, tx, ty, *, tu=None, tv=None):
    dev = _check(nx, ny, vs, ss, tx, ty, tu, tv)
    _check(idx, dtype=torch.int32)
    nt = tx.numel()
    tu = torch.zeros(nt, dtype=torch.float64, device=dev) if tu is None else tu
    tv = torch.zeros(nt, dtype=torch.float64, device=dev) if tv is None else tv
    env = env_for(dev)
    call("o2d_hip_panels_on_points_dev_d", ctypes.byref(env), nx.numel(), _ptr(nx), _ptr(ny),
         idx.numel() // 2, _ptr(idx), _ptr(vs), _ptr(ss), nt, _ptr(tx), _ptr(ty), _ptr(tu),
         _ptr(tv))
    return tu, tv


def points_on_panels(sx, sy, ss, nx, ny, idx, *, pu=None, pv=None):
    dev = _check(sx, sy, ss, nx, ny, pu, pv)
    _check(idx, dtype=torch.int32)
    npan = idx.numel() // 2
    pu = torch.zeros(npan, dtype=torch.float64, device=dev) if pu is None else pu
    pv = torch.zeros(npan, dtype=torch.float64, device=dev) if pv is None else pv
    env = env_for(dev)
    call("o2d_hip_points_on_panels_dev_d", ctypes.byref(env), sx.numel(), _ptr(sx), _ptr(sy),
         _ptr(ss), nx.numel(), _ptr(nx), _ptr(ny), npan, _ptr(idx), _ptr(pu), _ptr(pv))
    return pu, pv


def bem_assemble(nx, ny, idx, nunk, augmented=False, urot_u=None, urot_v=None, vol=0.0, *,
                 out=None):
    """Returns the column-major matrix storage as a (ncols, nrows) tensor (A^T in row-major)."""
    dev = _check(nx, ny, urot_u, urot_v, out)
    _check(idx, dtype=torch.int32)
    npan = idx.numel() // 2
    n = npan * nunk + (1 if augmented else 0)
    A = torch.empty((n, n), dtype=torch.float64, device=dev) if out is None else out
    env = env_for(dev)
    call("o2d_hip_bem_assemble_dev_d", ctypes.byref(env), nx.numel(), _ptr(nx), _ptr(ny), npan,
         _ptr(idx), int(nunk), int(bool(augmented)), _ptr(urot_u), _ptr(urot_v), float(vol),
         _ptr(A), n)
    return A


def bem_factor(A_colmajor):
    dev = _check(A_colmajor)
    n = A_colmajor.shape[0]
    env = env_for(dev)
    call("o2d_hip_bem_factor_dev_d", ctypes.byref(env), n, _ptr(A_colmajor), n)


def bem_solve(b, x=None):
    dev = _check(b, x)
    x = b.clone() if x is None else x
    env = env_for(dev)
    call("o2d_hip_bem_solve_dev_d", ctypes.byref(env), b.numel(), _ptr(b), _ptr(x))
    return x


def rk_move(x, y, us, vs, w, dt):
    dev = _check(x, y, *us, *vs)
    env = env_for(dev)
    wv = (ctypes.c_double * len(w))(*[float(a) for a in w])
    up = (ctypes.c_void_p * len(us))(*[u.data_ptr() for u in us])
    vp = (ctypes.c_void_p * len(vs))(*[v.data_ptr() for v in vs])
    call("o2d_hip_rk_move_dev_d", ctypes.byref(env), x.numel(), float(dt), len(us), wv, up, vp,
         _ptr(x), _ptr(y))
    return x, y


def points_on_points_adjoint(sx, sy, ss, sr, tx, ty, gu, gv, *, core="gauss", out=None):
    dev = _check(sx, sy, ss, sr, tx, ty, gu, gv)
    ns = sx.numel()
    if out is None:
        out = tuple(torch.zeros(ns, dtype=torch.float64, device=dev) for _ in range(4))
    _check(*out)
    env = env_for(dev, core)
    call("o2d_hip_points_on_points_adjoint_dev_d", ctypes.byref(env), ns, _ptr(sx), _ptr(sy),
         _ptr(ss), _ptr(sr), tx.numel(), _ptr(tx), _ptr(ty), _ptr(gu), _ptr(gv),
         *[_ptr(o) for o in out])
    return out


def morton_order(x, y, x0, y0, inv_extent):
    dev = _check(x, y)
    n = x.numel()
    keys = torch.empty(n, dtype=torch.int32, device=dev)
    order = torch.empty(n, dtype=torch.int32, device=dev)
    env = env_for(dev)
    call("o2d_hip_morton_order_dev_d", ctypes.byref(env), n, _ptr(x), _ptr(y), float(x0),
         float(y0), float(inv_extent), _ptr(keys), _ptr(order))
    return keys, order


def points_on_points_tree(sx, sy, ss, sr, tx, ty, tr=None, *, core="gauss", theta=0.6, order=28,
                          core_tol=1e-12, tu=None, tv=None):
    """Treecode velocity sum on device tensors (accumulates into tu, tv; zeros when absent)."""
    dev = _check(sx, sy, ss, sr, tx, ty, tr, tu, tv)
    nt = tx.numel()
    tu = torch.zeros(nt, dtype=torch.float64, device=dev) if tu is None else tu
    tv = torch.zeros(nt, dtype=torch.float64, device=dev) if tv is None else tv
    env = env_for(dev, core=core)
    opt = TreeOpts(float(theta), int(order), 0, float(core_tol))
    call("o2d_hip_points_on_points_tree_dev_d", ctypes.byref(env), ctypes.byref(opt), sx.numel(),
         _ptr(sx), _ptr(sy), _ptr(ss), _ptr(sr), nt, _ptr(tx), _ptr(ty), _ptr(tr), _ptr(tu),
         _ptr(tv))
    return tu, tv


def points_vorticity(sx, sy, ss, sr, tx, ty, tw=None):
    """Accumulates the unscaled Gaussian vorticity sum into tw (zeros when not given)."""
    dev = _check(sx, sy, ss, sr, tx, ty, tw)
    tw = torch.zeros(tx.numel(), dtype=torch.float64, device=dev) if tw is None else tw
    env = env_for(dev, core="gauss")
    call("o2d_hip_points_vorticity_dev_d", ctypes.byref(env), sx.numel(), _ptr(sx), _ptr(sy),
         _ptr(ss), _ptr(sr), tx.numel(), _ptr(tx), _ptr(ty), _ptr(tw))
    return tw


def nearest_distances(nx, ny, idx, tx, ty, dist=None):
    dev = _check(nx, ny, tx, ty)
    _check(idx, dtype=torch.int32)
    nt = tx.numel()
    dist = torch.empty(nt, dtype=torch.float64, device=dev) if dist is None else dist
    env = env_for(dev)
    call("o2d_hip_nearest_distances_dev_d", ctypes.byref(env), nx.numel(), _ptr(nx), _ptr(ny),
         idx.numel() // 2, _ptr(idx), nt, _ptr(tx), _ptr(ty), _ptr(dist))
    return dist


def reflect(nx, ny, idx, pnx, pny, tx, ty):
    """In place on tx, ty; returns a device int64[3] (reflected, 0, bad panels)."""
    dev = _check(nx, ny, pnx, pny, tx, ty)
    _check(idx, dtype=torch.int32)
    counts = torch.zeros(3, dtype=torch.int64, device=dev)
    env = env_for(dev)
    call("o2d_hip_reflect_dev_d", ctypes.byref(env), nx.numel(), _ptr(nx), _ptr(ny),
         idx.numel() // 2, _ptr(idx), _ptr(pnx), _ptr(pny), tx.numel(), _ptr(tx), _ptr(ty),
         _ptr(counts))
    return counts


def clear_inner(method, nx, ny, idx, pnx, pny, cutoff_mult, ips, tx, ty, ts=None, tr=None):
    """In place on tx, ty (and ts for method 0); returns (circ_removed, counts) as device
    tensors, counts = (moves, unsettled targets, bad panels)."""
    dev = _check(nx, ny, pnx, pny, tx, ty, ts, tr)
    _check(idx, dtype=torch.int32)
    counts = torch.zeros(3, dtype=torch.int64, device=dev)
    circ = torch.zeros(1, dtype=torch.float64, device=dev)
    env = env_for(dev)
    call("o2d_hip_clear_inner_dev_d", ctypes.byref(env), int(method), float(cutoff_mult),
         float(ips), nx.numel(), _ptr(nx), _ptr(ny), idx.numel() // 2, _ptr(idx), _ptr(pnx),
         _ptr(pny), tx.numel(), _ptr(tx), _ptr(ty), _ptr(ts), _ptr(tr), _ptr(circ),
         _ptr(counts))
    return circ, counts

