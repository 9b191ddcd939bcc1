"""Per-configuration measurements beside the headline bench line (one JSON object per config).

    python bench_configs.py [--configs c1,c2host,c2tree,c3,c3step,c4,c5,...] [--out FILE]

C1  10k disc, RM core, one Ralston RK2 step: split and plain velocity evaluations, fp64 update;
    the Gaussian vorticity of the disc on itself
C2tree  the 1M blob workload with the treecode (summation_t barneshut): time, equivalent pair
    rate and error against the fp64 direct sum, at two accuracy settings
C3  circle of 2048 panels + 200k wake, V2 core: BEM assembly (vortex-only + augmented 2049^2,
    and vortex+source + augmented 4097^2) with HBM GB/s, LU factor + solve, RHS (points on
    panels), panels on particles, particles on particles (split and plain), and the
    particle/panel proximity scans run after every move (clearing pass, nearest distances),
    and the Gaussian vorticity of the wake on itself
C3step  whole RK2 steps of C3 through the C++ host mirror (tests/cpp/host_driver.cpp, built
    by python -m omega2d_amd.build), host arrays over PCIe
C4  250k particles -> 4096 probes, Gaussian core: forward and adjoint (fp64), ratio
C5  4M particles in a shear layer, RM core, one Ralston RK3 step in the split form with the
    fp64 position update; error of the split velocity against the fp64 form on 16,384 targets

Short drivers for profiler runs (each prints one JSON line; run them under rocprofv3):
lu        three factorisations of a random 2049^2 matrix, then three of a 4097^2 one
assemble  the C3 assembly: vortex-only + augmented 2049^2, then vortex+source + augmented
          4097^2, three times each
treewalk  the C2 treecode evaluation (theta 0.5, order 20) twice on the particle targets
near      the C3 proximity scans: three clearing passes (method 1) and three nearest-distance
          scans of the 200k wake against the 2,048 panels
panels    the C3 panel sums: three points-on-panels (RHS) and three panels-on-points launches
vort      three Gaussian vorticity sums of the C3 wake on itself
fp64c2    two fp64 evaluations of the C2 particles on themselves (Gaussian core)

Timing: torch.cuda.synchronize() around each measured piece; kernel times from the library's
HIP-event timers where a single kernel dominates. Inputs are resident in HBM.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402
import torch  # noqa: E402

from omega2d_amd import device as dv  # noqa: E402
from omega2d_amd import flops_per_pair, workloads as wl  # noqa: E402
from omega2d_amd._cabi import (KT_ADJOINT, KT_BEM_ASSEMBLE, KT_NBODY_PLAIN,  # noqa: E402
                               KT_NBODY_PLAIN_PT, KT_NBODY_SPLIT, KT_NEAR, KT_PANEL_RHS,
                               KT_PANELS, KT_TREE_BUILD, KT_TREE_EVAL, KT_VORT, timing_enable,
                               timing_read)

PEAK_FP32, PEAK_FP64, PEAK_HBM = 157.3e12, 78.6e12, 8.0e12


def T(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


def timed(fn, reps=1):
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def kernel_ms(cls, fn, reps=1):
    fn()
    torch.cuda.synchronize()
    timing_read(cls)
    timing_enable(True)
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    timing_enable(False)
    ms, n = timing_read(cls)
    return ms / max(n, 1), n / reps


def rel(a, b):
    a, b = (x.detach().cpu().numpy() if torch.is_tensor(x) else x for x in (a, b))
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def vel(x, y, s, r, tx, ty, tr, core, prec, out=None):
    u, v = dv.points_on_points(x, y, s, r, tx, ty, tr, core=core, precision=prec)
    c = 0.5 / math.pi
    return u * c, v * c


def c1():
    w = wl.c1_disc()
    x, y, s, r = (T(w[k]) for k in ("x", "y", "s", "r"))
    n = x.numel()
    res = {"config": "C1", "name": w["name"], "particles": n, "core": "rm"}
    for prec in ("split", "fp64"):
        xs, ys = x.clone(), y.clone()

        def step():
            xs.copy_(x)
            ys.copy_(y)
            u1, v1 = vel(xs, ys, s, r, xs, ys, r, "rm", prec)
            x1, y1 = xs.clone(), ys.clone()
            dv.rk_move(x1, y1, [u1], [v1], [1.0], 2.0 / 3.0 * w["dt"])
            u2, v2 = vel(x1, y1, s, r, x1, y1, r, "rm", prec)
            dv.rk_move(xs, ys, [u1, u2], [v1, v2], [0.25, 0.75], w["dt"])
        t = timed(step, reps=5)
        res[prec] = {"ms_per_step": round(t * 1e3, 4),
                     "Gint_s": round(2.0 * n * n / t / 1e9, 2)}
        res[prec + "_x"] = xs
    res["split_pos_err_vs_fp64"] = rel(res.pop("split_x"), res["fp64_x"])
    res.pop("fp64_x")
    # Gaussian vorticity of the disc on itself (src/InfluenceVort.h), 13 flops per pair in the
    # reference's count
    ms, _ = kernel_ms(KT_VORT, lambda: dv.points_vorticity(x, y, s, r, x, y), reps=5)
    res["vorticity"] = {"kernel_ms": round(ms, 4),
                        "Gint_s": round(float(n) * n / (ms * 1e-3) / 1e9, 2),
                        "frac_fp64": round(13.0 * n * n / (ms * 1e-3) / PEAK_FP64, 4)}
    return res


def c3():
    w = wl.c3_circle_wake()
    npan = len(w["idx"]) // 2
    nx, ny = T(w["nx"]), T(w["ny"])
    idx = torch.from_numpy(w["idx"].astype(np.int32)).cuda()
    x, y, s, r = (T(w[k]) for k in ("x", "y", "s", "r"))
    nw = x.numel()
    res = {"config": "C3", "name": w["name"], "panels": npan, "wake": nw, "core": "v2"}
    rng = np.random.default_rng(0)
    uu, uv = T(rng.normal(size=npan)), T(rng.normal(size=npan))
    for nunk in (1, 2):
        n = npan * nunk + 1
        A = torch.empty((n, n), dtype=torch.float64, device="cuda")
        ms, _ = kernel_ms(KT_BEM_ASSEMBLE, lambda: dv.bem_assemble(nx, ny, idx, nunk, True, uu, uv,
                                                                  0.78, out=A), reps=3)
        entries = float(npan) * npan * nunk * nunk
        flops = float(npan) * npan * (92.0 if nunk == 1 else 122.0)
        t_fac = timed(lambda: dv.bem_factor(A))
        b = T(rng.normal(size=n))
        for _ in range(2):  # the first round loads the solver's kernels for this size
            dv.bem_factor(A)
            for _ in range(5):
                dv.bem_solve(b)
        dv.bem_factor(A)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dv.bem_solve(b)  # triangular solves with the fresh factors
        torch.cuda.synchronize()
        t_first = time.perf_counter() - t0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dv.bem_solve(b)  # second solve: builds the explicit inverse, then one matrix-vector product
        torch.cuda.synchronize()
        t_inv = time.perf_counter() - t0
        t_sol = timed(lambda: dv.bem_solve(b), reps=10)
        res[f"assemble_{n}"] = {"kernel_ms": round(ms, 4),
                                "GB_s": round(8.0 * entries / (ms * 1e-3) / 1e9, 1),
                                "frac_hbm": round(8.0 * entries / (ms * 1e-3) / PEAK_HBM, 4),
                                "TFLOP_s": round(flops / (ms * 1e-3) / 1e12, 2),
                                "frac_fp64": round(flops / (ms * 1e-3) / PEAK_FP64, 4),
                                "lu_factor_ms": round(t_fac * 1e3, 3),
                                "lu_first_solve_ms": round(t_first * 1e3, 4),
                                "inverse_build_ms": round(t_inv * 1e3, 3),
                                "lu_solve_ms": round(t_sol * 1e3, 4)}
    vs, ss = T(rng.normal(size=npan)), T(rng.normal(size=npan))
    ms, _ = kernel_ms(KT_PANEL_RHS, lambda: dv.points_on_panels(x, y, s, nx, ny, idx), reps=3)
    # the panel sums take far clusters of panels by an expansion and the proximity scans skip
    # tiles, so their rates count the (target, panel) pairs of the full sum: equivalent rates,
    # with no fraction of a peak
    res["points_on_panels"] = {"kernel_ms": round(ms, 4),
                               "equivalent_Gint_s": round(float(nw) * npan / (ms * 1e-3) / 1e9, 2)}
    ms, _ = kernel_ms(KT_PANELS, lambda: dv.panels_on_points(nx, ny, idx, vs, ss, x, y), reps=3)
    res["panels_on_points"] = {"kernel_ms": round(ms, 4),
                               "equivalent_Gint_s": round(float(nw) * npan / (ms * 1e-3) / 1e9, 2)}
    for prec, cls in (("split", KT_NBODY_SPLIT), ("fp64", KT_NBODY_PLAIN)):
        ms, _ = kernel_ms(cls, lambda: dv.points_on_points(x, y, s, r, x, y, r, core="v2",
                                                           precision=prec), reps=2)
        fp = flops_per_pair("v2", False)
        res[f"particles_{prec}"] = {"kernel_ms": round(ms, 3),
                                    "Gint_s": round(float(nw) * nw / (ms * 1e-3) / 1e9, 1),
                                    "frac_peak": round(fp * float(nw) * nw / (ms * 1e-3) /
                                                       (PEAK_FP32 if prec == "split" else PEAK_FP64), 4)}
    # proximity scans (src/Reflect.h): one clearing pass (method 1, the reference's cutoff
    # 0.5/sqrt(2 pi) at spacing 0.005) and the nearest distances, every wake particle against
    # every panel
    a, b = w["idx"][0::2], w["idx"][1::2]
    tx_, ty_ = w["nx"][b] - w["nx"][a], w["ny"][b] - w["ny"][a]
    ln = np.sqrt(tx_ * tx_ + ty_ * ty_)
    pnx, pny = T(-ty_ * (1.0 / ln)), T(tx_ * (1.0 / ln))
    cm, ips = 0.5 / math.sqrt(2.0 * math.pi), 0.005
    for name, fn in (("clear_inner", lambda: dv.clear_inner(1, nx, ny, idx, pnx, pny, cm, ips,
                                                            x.clone(), y.clone())),
                     ("nearest_distances", lambda: dv.nearest_distances(nx, ny, idx, x, y))):
        ms, _ = kernel_ms(KT_NEAR, fn, reps=3)
        res[name] = {"kernel_ms": round(ms, 4),
                     "equivalent_Gint_s": round(float(nw) * npan / (ms * 1e-3) / 1e9, 2)}
    # Gaussian vorticity of the wake on itself (src/InfluenceVort.h): the kernel alone, and the
    # whole call with the Morton orders and the source tiles' boxes
    ms, _ = kernel_ms(KT_VORT, lambda: dv.points_vorticity(x, y, s, r, x, y), reps=3)
    t = timed(lambda: dv.points_vorticity(x, y, s, r, x, y), reps=3)
    res["vorticity"] = {"kernel_ms": round(ms, 4), "call_ms": round(t * 1e3, 4),
                        "equivalent_Gint_s": round(float(nw) * nw / (ms * 1e-3) / 1e9, 2)}
    us, vs_ = vel(x, y, s, r, x, y, r, "v2", "split")
    u6, v6 = vel(x, y, s, r, x, y, r, "v2", "fp64")
    res["split_err_vs_fp64"] = math.sqrt((rel(us, u6) ** 2 * float((u6 ** 2).sum())
                                          + rel(vs_, v6) ** 2 * float((v6 ** 2).sum()))
                                         / float((u6 ** 2).sum() + (v6 ** 2).sum()))
    return res


def c4():
    w = wl.c4_adjoint()
    x, y, s, r, tx, ty = (T(w[k]) for k in ("x", "y", "s", "r", "tx", "ty"))
    n, m = x.numel(), tx.numel()
    u = torch.zeros(m, dtype=torch.float64, device="cuda")
    v = torch.zeros_like(u)
    g = tuple(torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(4))

    def fwd():
        u.zero_()
        v.zero_()
        dv.points_on_points(x, y, s, r, tx, ty, None, core="gauss", tu=u, tv=v)

    def adj():
        for a in g:
            a.zero_()
        dv.points_on_points_adjoint(x, y, s, r, tx, ty, 2 * u, 2 * v, core="gauss", out=g)
    tf = timed(fwd, reps=10)
    ta = timed(adj, reps=10)
    kf, _ = kernel_ms(KT_NBODY_PLAIN_PT, fwd, reps=5)  # probes are point targets
    ka, _ = kernel_ms(KT_ADJOINT, adj, reps=5)
    pairs = float(n) * m
    return {"config": "C4", "name": w["name"], "particles": n, "probes": m,
            "fwd_ms": round(tf * 1e3, 4), "adjoint_ms": round(ta * 1e3, 4),
            "adjoint_over_fwd": round(ta / tf, 3), "fwd_kernel_ms": round(kf, 4),
            "adjoint_kernel_ms": round(ka, 4), "fwd_Gint_s": round(pairs / tf / 1e9, 1),
            "adjoint_Gint_s": round(pairs / ta / 1e9, 1)}


def c5():
    w = wl.c5_shear_layer()
    x, y, s, r = (T(w[k]) for k in ("x", "y", "s", "r"))
    n = x.numel()
    dt = w["dt"]
    xs, ys = x.clone(), y.clone()
    stage = {}

    def step():
        # Ralston RK3 (src/Convection.h:472-598): stages at t, t+dt/2, t+3dt/4 (from x0 with u1)
        xs.copy_(x)
        ys.copy_(y)
        u0, v0 = vel(xs, ys, s, r, xs, ys, r, "rm", "split")
        x1, y1 = xs.clone(), ys.clone()
        dv.rk_move(x1, y1, [u0], [v0], [1.0], 0.5 * dt)
        u1, v1 = vel(x1, y1, s, r, x1, y1, r, "rm", "split")
        x2, y2 = xs.clone(), ys.clone()
        dv.rk_move(x2, y2, [u1], [v1], [1.0], 0.75 * dt)
        u2, v2 = vel(x2, y2, s, r, x2, y2, r, "rm", "split")
        dv.rk_move(xs, ys, [u0, u1, u2], [v0, v1, v2], [2 / 9, 3 / 9, 4 / 9], dt)
        stage["u0"], stage["v0"] = u0, v0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    ms, launches = kernel_ms(KT_NBODY_SPLIT, lambda: vel(x, y, s, r, x, y, r, "rm", "split"))
    # error of the first-stage split velocity against the fp64 form on a target subsample
    rng = np.random.default_rng(4)
    sub = torch.from_numpy(np.sort(rng.choice(n, 16384, replace=False))).cuda()
    u6, v6 = vel(x, y, s, r, x[sub].contiguous(), y[sub].contiguous(), r[sub].contiguous(),
                 "rm", "fp64")
    us, vs_ = stage["u0"][sub], stage["v0"][sub]
    err = math.sqrt(float(((us - u6) ** 2).sum() + ((vs_ - v6) ** 2).sum())
                    / float((u6 ** 2).sum() + (v6 ** 2).sum()))
    fp = flops_per_pair("rm", False)
    pairs = float(n) * n
    return {"config": "C5", "name": w["name"], "particles": n, "core": "rm",
            "ms_per_step": round(t * 1e3, 1), "Gint_s_step": round(3 * pairs / t / 1e9, 1),
            "eval_kernel_ms": round(ms, 2), "Gint_s_kernel": round(pairs / (ms * 1e-3) / 1e9, 1),
            "frac_fp32_peak": round(fp * pairs / (ms * 1e-3) / PEAK_FP32, 4),
            "split_err_vs_fp64_16k": err}


def c3step(nsteps=10):
    """Whole Ralston RK2 advection steps of C3 through the C++ host mirror (omega2d_amd/host,
    src/Convection.h:390-466): each of the two stages solves the BEM (RHS from the wake, the
    2049^2 vortex-only system), sums panels and particles on the 200,000 wake particles, moves
    them and clears the inner layer (method 1). The driver passes host arrays, so every call
    copies its inputs over PCIe, as the reference's call would. Wall time per step after three
    untimed steps, split and fp64 pair sums."""
    import subprocess
    import tempfile
    from omega2d_amd import build as B
    w = wl.c3_circle_wake()
    npan, nw = len(w["idx"]) // 2, len(w["x"])
    res = {"config": "C3step", "name": w["name"], "panels": npan, "wake": nw, "core": "v2",
           "unknowns_per_panel": 1, "steps_timed": nsteps}
    with tempfile.TemporaryDirectory() as tmp:
        for prec, pid in (("split", 2), ("fp64", 1)):
            inp, out = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
            with open(inp, "wb") as f:
                f.write(np.array([npan, len(w["nx"]), nw, 1, 4, 2, nsteps, pid],
                                 np.int64).tobytes())
                f.write(np.array([w["fs"][0], w["fs"][1], 0.01, 0.5]).tobytes())
                for a in (w["nx"].astype(np.float32), w["ny"].astype(np.float32),
                          w["idx"].astype(np.uint32), w["x"], w["y"], w["s"], w["r"]):
                    f.write(np.ascontiguousarray(a).tobytes())
            r = subprocess.run([B.HOST_TEST_BIN, "bemtime", inp, out], capture_output=True,
                               text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError(r.stderr[-2000:])
            d = np.fromfile(out, dtype=np.float64)
            assert d.size == 1 + 2 * nw and np.all(np.isfinite(d))
            res[prec] = {"ms_per_step": round(float(d[0]), 3)}
    return res


def c2host():
    """The C2 evaluation (1M particles on themselves plus 65,536 tracers, split form) through the
    host entry point, as the reference calls it: host arrays in, copies over PCIe both ways,
    results in host arrays on return. Reported beside the same evaluation on device-resident
    arrays."""
    import omega2d_amd as o2d
    w = wl.c2_blobs()
    n, nt = len(w["x"]), len(w["tx"])
    pairs = float(n) * n + float(nt) * n

    def host_step():
        o2d.points_on_points(w["x"], w["y"], w["s"], w["r"], w["x"], w["y"], w["r"], core="gauss",
                             precision="split")
        o2d.points_on_points(w["x"], w["y"], w["s"], w["r"], w["tx"], w["ty"], core="gauss",
                             precision="split")
    x, y, s, r = (T(w[k]) for k in ("x", "y", "s", "r"))
    tx, ty = T(w["tx"]), T(w["ty"])

    def dev_step():
        dv.points_on_points(x, y, s, r, x, y, r, core="gauss", precision="split")
        dv.points_on_points(x, y, s, r, tx, ty, None, core="gauss", precision="split")
    host_step()
    th = timed(host_step, reps=3)
    dev_step()
    td = timed(dev_step, reps=3)
    return {"config": "C2host", "name": w["name"], "particles": n, "tracers": nt,
            "host_entry_ms": round(th * 1e3, 2), "host_entry_Gint_s": round(pairs / th / 1e9, 1),
            "device_entry_ms": round(td * 1e3, 2),
            "device_entry_Gint_s": round(pairs / td / 1e9, 1)}


def c2tree():
    """The 1M blob workload (1,048,576 particles on themselves + 65,536 tracers, Gaussian) with
    the treecode: build and walk kernel times, the equivalent direct-sum pair rate, and the error
    against the fp64 direct sum on 4,096 particle targets."""
    w = wl.c2_blobs()
    x, y, s, r = (T(w[k]) for k in ("x", "y", "s", "r"))
    tx, ty = T(w["tx"]), T(w["ty"])
    n, nt = x.numel(), tx.numel()
    pairs = float(n) * n + float(nt) * n
    m = 4096
    idx = torch.arange(0, n, n // m, device="cuda")[:m]
    ru, rv = dv.points_on_points(x, y, s, r, x[idx].contiguous(), y[idx].contiguous(),
                                 r[idx].contiguous(), core="gauss", precision="fp64")
    res = {"config": "C2tree", "name": w["name"], "particles": n, "tracers": nt,
           "equivalent_pairs": pairs}
    for theta, order in ((0.6, 28), (0.5, 20), (0.4, 30)):
        def step():
            u, v = dv.points_on_points_tree(x, y, s, r, x, y, r, theta=theta, order=order)
            dv.points_on_points_tree(x, y, s, r, tx, ty, theta=theta, order=order)
            return u, v
        u, v = step()
        torch.cuda.synchronize()
        t = timed(step, reps=3)
        timing_read(KT_TREE_BUILD)
        timing_read(KT_TREE_EVAL)
        timing_enable(True)
        step()
        torch.cuda.synchronize()
        timing_enable(False)
        bms, _ = timing_read(KT_TREE_BUILD)
        ems, _ = timing_read(KT_TREE_EVAL)
        err = float(torch.sqrt(((u[idx] - ru) ** 2 + (v[idx] - rv) ** 2).sum() /
                               (ru ** 2 + rv ** 2).sum()))
        res[f"theta{theta}_order{order}"] = {
            "ms_per_step": round(t * 1e3, 3), "build_ms": round(bms, 3), "walk_ms": round(ems, 3),
            "equivalent_Gint_s": round(pairs / t / 1e9, 1), "err_vs_fp64_direct": err}
    return res


def _c3_panels():
    w = wl.c3_circle_wake()
    nx, ny = T(w["nx"]), T(w["ny"])
    idx = torch.from_numpy(w["idx"].astype(np.int32)).cuda()
    return w, nx, ny, idx


def lu():
    """Profiler driver: three LU factorisations of a random 2049^2 matrix, then of a 4097^2
    one (the wall time of each third factorisation is reported)."""
    rng = np.random.default_rng(0)
    res = {"config": "lu"}
    for n in (2049, 4097):
        A0 = T(rng.normal(size=(n, n)))
        for _ in range(3):
            A = A0.clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            dv.bem_factor(A)
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
        res[f"factor_{n}_ms"] = round(t * 1e3, 3)
    return res


def assemble():
    """Profiler driver: the C3 assembly, vortex-only + augmented 2049^2 and vortex+source +
    augmented 4097^2, three launches each."""
    w, nx, ny, idx = _c3_panels()
    npan = len(w["idx"]) // 2
    rng = np.random.default_rng(0)
    uu, uv = T(rng.normal(size=npan)), T(rng.normal(size=npan))
    res = {"config": "assemble"}
    for nunk in (1, 2):
        n = npan * nunk + 1
        A = torch.empty((n, n), dtype=torch.float64, device="cuda")
        ms, _ = kernel_ms(KT_BEM_ASSEMBLE, lambda: dv.bem_assemble(nx, ny, idx, nunk, True, uu,
                                                                  uv, 0.78, out=A), reps=3)
        res[f"assemble_{n}_kernel_ms"] = round(ms, 4)
        res[f"assemble_{n}_bytes_written"] = 8 * n * n
    return res


def treewalk():
    """Profiler driver: the C2 treecode evaluation (theta 0.5, order 20) on the 1M particle
    targets, twice."""
    w = wl.c2_blobs()
    x, y, s, r = (T(w[k]) for k in ("x", "y", "s", "r"))
    for _ in range(2):
        dv.points_on_points_tree(x, y, s, r, x, y, r, theta=0.5, order=20)
    torch.cuda.synchronize()
    return {"config": "treewalk", "particles": x.numel()}


def near():
    """Profiler driver: three clearing passes (method 1, the reference's layer at spacing
    0.005) and three nearest-distance scans of the C3 wake against its 2,048 panels."""
    w, nx, ny, idx = _c3_panels()
    x, y = T(w["x"]), T(w["y"])
    a, b = w["idx"][0::2], w["idx"][1::2]
    tx_, ty_ = w["nx"][b] - w["nx"][a], w["ny"][b] - w["ny"][a]
    ln = np.sqrt(tx_ * tx_ + ty_ * ty_)
    pnx, pny = T(-ty_ * (1.0 / ln)), T(tx_ * (1.0 / ln))
    cm, ips = 0.5 / math.sqrt(2.0 * math.pi), 0.005
    res = {"config": "near"}
    for name, fn in (("clear_inner", lambda: dv.clear_inner(1, nx, ny, idx, pnx, pny, cm, ips,
                                                            x.clone(), y.clone())),
                     ("nearest_distances", lambda: dv.nearest_distances(nx, ny, idx, x, y))):
        ms, _ = kernel_ms(KT_NEAR, fn, reps=3)
        res[name + "_kernel_ms"] = round(ms, 4)
    return res


def panels():
    """Profiler driver: three points-on-panels (the BEM right-hand side from the 200k wake) and
    three panels-on-points (the 2,048 panels on the wake) launches."""
    w, nx, ny, idx = _c3_panels()
    npan = len(w["idx"]) // 2
    x, y, s = T(w["x"]), T(w["y"]), T(w["s"])
    rng = np.random.default_rng(0)
    vs, ss = T(rng.normal(size=npan)), T(rng.normal(size=npan))
    res = {"config": "panels"}
    ms, _ = kernel_ms(KT_PANEL_RHS, lambda: dv.points_on_panels(x, y, s, nx, ny, idx), reps=3)
    res["points_on_panels_kernel_ms"] = round(ms, 4)
    ms, _ = kernel_ms(KT_PANELS, lambda: dv.panels_on_points(nx, ny, idx, vs, ss, x, y), reps=3)
    res["panels_on_points_kernel_ms"] = round(ms, 4)
    return res


def vort():
    """Profiler driver: three Gaussian vorticity sums of the C3 wake on itself."""
    w = wl.c3_circle_wake()
    x, y, s, r = (T(w[k]) for k in ("x", "y", "s", "r"))
    ms, _ = kernel_ms(KT_VORT, lambda: dv.points_vorticity(x, y, s, r, x, y), reps=3)
    return {"config": "vort", "particles": x.numel(), "kernel_ms": round(ms, 4)}


def fp64c2():
    """Profiler driver: two fp64 evaluations of the C2 particles on themselves (Gaussian)."""
    w = wl.c2_blobs()
    x, y, s, r = (T(w[k]) for k in ("x", "y", "s", "r"))
    ms, n = kernel_ms(KT_NBODY_PLAIN, lambda: dv.points_on_points(x, y, s, r, x, y, r,
                                                                  core="gauss",
                                                                  precision="fp64"))
    return {"config": "fp64c2", "particles": x.numel(), "kernel_ms": round(ms, 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--configs", default="c1,c3,c4,c5")
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    fns = {"c1": c1, "c2host": c2host, "c2tree": c2tree, "c3": c3, "c3step": c3step,
           "c4": c4, "c5": c5, "lu": lu, "assemble": assemble, "treewalk": treewalk,
           "near": near, "panels": panels, "vort": vort, "fp64c2": fp64c2}
    out = []
    for name in args.configs.split(","):
        t0 = time.perf_counter()
        res = fns[name]()
        res["wall_s"] = round(time.perf_counter() - t0, 2)
        print(json.dumps(res), flush=True)
        out.append(res)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(out, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
