"""Write tests/golden/reference_vectors.npz from the reference's own templates.

Runs in a container that has the reference tree: it calls oracle/_ref/libo2d_ref.so (built by
``make -C oracle ref`` from src/Influence.h, src/Coefficients.h, src/RHS.h, src/Points.h,
src/Surfaces.h, src/ElementBase.h, src/Reflect.h, src/InfluenceVort.h where they lie) and, under
tests/golden/vtk/, the .vtu files its own writers produce (src/VtkXmlWriter.h) on small seeded inputs and stores inputs and
outputs. The fixture is data only: inputs and the reference's outputs.

    python tests/golden/make_golden.py

Core function: the reference compiles in Vatistas n=2 (src/CoreFunc.h:38), so every
particle-on-point vector of reference_vectors.npz uses that core. tests/golden/core_vectors.npz
holds the particle-on-point vectors of every core: oracle/Makefile also builds the harness once
per other core (rm, gauss, wl, v3) with only the reference's core define switched, and this
script calls each build (oracle.ref(core) checks the core each build reports).

    python tests/golden/make_golden.py            # both files
    python tests/golden/make_golden.py cores      # core_vectors.npz only
"""
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle as O  # noqa: E402
from helpers import circle_panels  # noqa: E402

OUT = os.path.join(HERE, "reference_vectors.npz")
OUT_CORES = os.path.join(HERE, "core_vectors.npz")
CORE_NAMES = ("rm", "gauss", "wl", "v2", "v3")


def gauss_cutoff_targets(sx, sy, w, angle):
    """Targets around one source at squared distances x*w, with x on both sides of the Gaussian
    core's cutoffs (src/CoreFunc.h:119-131): 16, 0.001 and the float literal 0.001f the
    reference compares with, and inside the window between the last two. Offsets start at
    1e-13 relative, clear of the last-ulp differences that fused multiply-adds make."""
    cuts = (16.0, 0.001, float(np.float32(0.001)))
    rels = (1e-13, 1e-11, 1e-9, 1e-7, 1e-5, 1e-3)
    xs = [c * (1 + sg * r) for c in cuts for r in rels for sg in (-1, 1)]
    lo, hi = 0.001, float(np.float32(0.001))
    xs += list(lo + (hi - lo) * np.linspace(0.02, 0.98, 25))
    xs = np.array(xs)
    d = np.sqrt(xs * w)
    return sx + d * np.cos(angle), sy + d * np.sin(angle), xs


def core_vectors():
    """Every core x {point, particle targets} x {vel, vel+vort, vel+shear}, S=double, from the
    reference build carrying that core; velocity at S=float too; the Gaussian cutoff sets."""
    for c in CORE_NAMES:
        if not O.have_ref(c):
            raise SystemExit(f"{O.ref_so(c)} missing: run make -C oracle ref")
    rng = np.random.default_rng(20261203)
    d = {}
    ns, nt = 300, 200
    # radii large against the spacing, so that many pairs sit inside the cores
    sx, sy = rng.uniform(-1, 1, ns), rng.uniform(-1, 1, ns)
    ss, sr = rng.normal(0, 1, ns), rng.uniform(0.02, 0.25, ns)
    tx, ty = rng.uniform(-1.1, 1.1, nt), rng.uniform(-1.1, 1.1, nt)
    tr = rng.uniform(0.02, 0.25, nt)
    d.update(cv_sx=sx, cv_sy=sy, cv_ss=ss, cv_sr=sr, cv_tx=tx, cv_ty=ty, cv_tr=tr)
    for c in CORE_NAMES:
        for res in ("vel", "vort", "shear"):
            for tag, trr in (("0p", None), ("0b", tr)):
                out = O.ref_points_on_points(sx, sy, ss, sr, tx, ty, trr, res, core=c)
                for k, a in zip("uvw", out):
                    d[f"cv_{c}_{tag}_{res}_{k}"] = a
        # the self interaction with vorticity (targets are the sources, particle targets)
        out = O.ref_points_on_points(sx, sy, ss, sr, sx, sy, sr, "vort", core=c)
        for k, a in zip("uvw", out):
            d[f"cv_{c}_self_{k}"] = a
        # float storage (the application's STORE=float), velocity
        f = [a.astype(np.float32) for a in (sx, sy, ss, sr, tx, ty, tr)]
        for tag, trr in (("0p", None), ("0b", f[6])):
            fu, fv = O.ref_points_on_points(*f[:6], trr, "vel", dtype=np.float32, core=c)
            d[f"cvf_{c}_{tag}_u"], d[f"cvf_{c}_{tag}_v"] = fu, fv

    # the Gaussian at its cutoffs: one source, point targets (w = sr^2) and particle targets of
    # one radius (w = sr^2 + tr^2)
    s0 = np.array([0.1]), np.array([-0.2]), np.array([0.7]), np.array([0.05])
    trad = 0.03
    for tag, w, trr in (("0p", 0.05 ** 2, None), ("0b", 0.05 ** 2 + trad ** 2, trad)):
        gx, gy, xs = gauss_cutoff_targets(0.1, -0.2, w, 0.3)
        tt = None if trr is None else np.full(len(gx), trr)
        d[f"gc_{tag}_tx"], d[f"gc_{tag}_ty"], d[f"gc_{tag}_x"] = gx, gy, xs
        for res in ("vel", "vort", "shear"):
            out = O.ref_points_on_points(*s0, gx, gy, tt, res, core="gauss")
            for k, a in zip("uvw", out):
                d[f"gc_{tag}_{res}_{k}"] = a
    d["gc_src"] = np.concatenate(s0)
    d["gc_trad"] = np.array([trad])
    np.savez_compressed(OUT_CORES, **d)
    print(f"wrote {OUT_CORES} ({os.path.getsize(OUT_CORES)} bytes, {len(d)} arrays)")


def main():
    if not O.have_ref():
        raise SystemExit("oracle/_ref/libo2d_ref.so missing: run make -C oracle ref")
    rng = np.random.default_rng(20240601)
    d = {}

    # particles on points, S=double A=double and S=float A=double
    ns, nt = 512, 300
    sx, sy = rng.uniform(-1, 1, ns), rng.uniform(-1, 1, ns)
    ss, sr = rng.normal(0, 1, ns), rng.uniform(0.02, 0.1, ns)
    tx, ty = rng.uniform(-1.2, 1.2, nt), rng.uniform(-1.2, 1.2, nt)
    tr = rng.uniform(0.02, 0.1, nt)
    d.update(pp_sx=sx, pp_sy=sy, pp_ss=ss, pp_sr=sr, pp_tx=tx, pp_ty=ty, pp_tr=tr)
    for res in ("vel", "vort", "shear"):
        for tag, trr in (("0p", None), ("0b", tr)):
            out = O.ref_points_on_points(sx, sy, ss, sr, tx, ty, trr, res)
            for k, a in zip("uvw", out):
                d[f"pp_{tag}_{res}_{k}"] = a
    # self interaction (targets are the sources), vorticity included
    out = O.ref_points_on_points(sx, sy, ss, sr, sx, sy, sr, "vort")
    d["pp_self_u"], d["pp_self_v"], d["pp_self_w"] = out
    # float storage (the application's STORE=float)
    f = [a.astype(np.float32) for a in (sx, sy, ss, sr, tx, ty, tr)]
    fu, fv = O.ref_points_on_points(*f, "vel", dtype=np.float32)
    d["ppf_0b_u"], d["ppf_0b_v"] = fu, fv

    # panels on points (vortex + source panels), points on panels
    npan = 64
    nx, ny, idx = circle_panels(npan)
    d.update(pan_nx=nx, pan_ny=ny, pan_idx=idx)
    vs, sst = rng.normal(0, 1, npan), rng.normal(0, 1, npan)
    px, py = rng.uniform(-1.5, 1.5, 400), rng.uniform(-1.5, 1.5, 400)
    d.update(pan_vs=vs, pan_ss=sst, pan_tx=px, pan_ty=py)
    d["pan_u"], d["pan_v"] = O.ref_panels_on_points(nx, ny, idx, vs, sst, px, py)
    qx, qy = rng.uniform(-1.5, 1.5, 500), rng.uniform(-1.5, 1.5, 500)
    qs, qr = rng.normal(0, 1, 500), np.full(500, 0.01)
    d.update(ptp_sx=qx, ptp_sy=qy, ptp_ss=qs, ptp_sr=qr)
    d["ptp_u"], d["ptp_v"] = O.ref_points_on_panels(qx, qy, qs, qr, nx, ny, idx)

    # BEM self blocks: vortex-only and vortex+source, no body / ground body / moving body
    bnp = 32
    bx, by, bidx = circle_panels(bnp)
    d.update(bem_nx=bx, bem_ny=by, bem_idx=bidx)
    for nunk in (1, 2):
        for kind in (0, 1, 2):
            R = O.ref_bem_coeffs(bx, by, bidx, nunk, kind)
            for k in ("A", "tang_x", "tang_y", "area", "urot_u", "urot_v"):
                d[f"bem_{nunk}_{kind}_{k}"] = R[k]
            d[f"bem_{nunk}_{kind}_vol"] = np.array([R["vol"]])
    # internal flow (reversed panel order): negative volume, the ground body is not augmented
    ix, iy, iidx = circle_panels(bnp, external=False)
    R = O.ref_bem_coeffs(ix, iy, iidx, 2, 1)
    d["bem_int_idx"], d["bem_int_A"], d["bem_int_vol"] = iidx, R["A"], np.array([R["vol"]])

    # RHS from panel velocities
    pu, pv = rng.normal(0, 1, bnp), rng.normal(0, 1, bnp)
    nbc = rng.normal(0, 0.1, bnp).astype(np.float32)
    d.update(rhs_pu=pu, rhs_pv=pv, rhs_nbc=nbc)
    for nunk in (1, 2):
        d[f"rhs_{nunk}"] = O.ref_rhs_panels(bx, by, bidx, nunk, pu, pv, nbc)

    # finalize and RK moves
    n = 257
    x, y = rng.normal(0, 1, n), rng.normal(0, 1, n)
    us = [rng.normal(0, 1, n) for _ in range(3)]
    vv = [rng.normal(0, 1, n) for _ in range(3)]
    d.update(mv_x=x, mv_y=y, mv_u1=us[0], mv_u2=us[1], mv_u3=us[2], mv_v1=vv[0], mv_v2=vv[1],
             mv_v3=vv[2])
    d["mv2_x"], d["mv2_y"] = O.ref_move(x, y, us[:2], vv[:2], [0.25, 0.75], 0.01)
    d["mv3_x"], d["mv3_y"] = O.ref_move(x, y, us, vv, [2 / 9, 3 / 9, 4 / 9], 0.01)
    d["fin_u"], d["fin_v"] = O.ref_finalize(us[0], vv[0], (1.0, -0.5))

    # particle/panel proximity (src/Reflect.h), its own seed so the arrays above do not move
    r2 = np.random.default_rng(20241202)
    cnp = 48
    cx_, cy_, cidx = circle_panels(cnp)
    d.update(rf_nx=cx_, rf_ny=cy_, rf_idx=cidx)
    d["rf_pnx"], d["rf_pny"] = O.ref_panel_normals(cx_, cy_, cidx)
    nrf = 700
    rho, ang = r2.uniform(0.0, 0.8, nrf), r2.uniform(0.0, 2 * np.pi, nrf)
    fx, fy = rho * np.cos(ang), rho * np.sin(ang)
    d.update(rf_tx=fx, rf_ty=fy)
    d["rf_dist"] = O.ref_nearest_distances(cx_, cy_, cidx, fx, fy)
    d["rf_refl_x"], d["rf_refl_y"] = O.ref_reflect(cx_, cy_, cidx, fx, fy)
    cm, ips = 0.5 / np.sqrt(2.0 * np.pi), 0.05
    d["rf_cut"] = np.array([cm, ips])
    ps, pr = r2.normal(0, 1e-3, nrf), r2.uniform(0.015, 0.025, nrf)
    d.update(rf_ts=ps, rf_tr=pr)
    R = O.ref_clear_inner(1, cx_, cy_, cidx, cm, ips, fx, fy, ps, pr)
    d["rf_c1_x"], d["rf_c1_y"] = R["x"], R["y"]
    R = O.ref_clear_inner(1, cx_, cy_, cidx, cm, ips, fx, fy)
    d["rf_c1f_x"], d["rf_c1f_y"] = R["x"], R["y"]
    # method 0 on a shell that starts inside the cut table (a target more than 3.05 radii
    # below the layer is moved to exactly 1 radius above it, where rounding decides whether
    # it is trimmed once more), and on a deep set for the band check
    for tag, lo, hi in (("sh", 0.44, 0.62), ("dp", 0.0, 0.40)):
        rr = r2.uniform(lo, hi, 400)
        aa = r2.uniform(0.0, 2 * np.pi, 400)
        mx, my = rr * np.cos(aa), rr * np.sin(aa)
        ms, mr = r2.normal(0, 1e-3, 400), np.full(400, 0.025)
        R = O.ref_clear_inner(0, cx_, cy_, cidx, cm, ips, mx, my, ms, mr)
        d.update({f"rf_c0{tag}_tx": mx, f"rf_c0{tag}_ty": my, f"rf_c0{tag}_ts": ms,
                  f"rf_c0{tag}_tr": mr, f"rf_c0{tag}_x": R["x"], f"rf_c0{tag}_y": R["y"],
                  f"rf_c0{tag}_s": R["str"], f"rf_c0{tag}_circ": np.array([R["circ_removed"]])})
    d["rf_ct_pos"], d["rf_ct_sm"], d["rf_ct_ds"] = O.ref_cut_table()
    # an open wall with a spur: a node shared by three panels
    tnx = np.array([-1.0, -0.5, 0.0, 0.5, 1.0, 0.0, 0.0], np.float32)
    tny = np.array([0.0, 0.0, 0.0, 0.0, 0.0, 0.5, 1.0], np.float32)
    tidx = np.array([0, 1, 1, 2, 2, 3, 3, 4, 2, 5, 5, 6], np.uint32)
    d.update(rt_nx=tnx, rt_ny=tny, rt_idx=tidx)
    d["rt_pnx"], d["rt_pny"] = O.ref_panel_normals(tnx, tny, tidx)
    gx, gy = r2.uniform(-1.2, 1.2, 500), r2.uniform(-0.5, 1.2, 500)
    d.update(rt_tx=gx, rt_ty=gy)
    d["rt_dist"] = O.ref_nearest_distances(tnx, tny, tidx, gx, gy)
    d["rt_refl_x"], d["rt_refl_y"] = O.ref_reflect(tnx, tny, tidx, gx, gy)
    R = O.ref_clear_inner(1, tnx, tny, tidx, cm, ips, gx, gy)
    d["rt_c1_x"], d["rt_c1_y"] = R["x"], R["y"]

    # Gaussian vorticity of particles (src/InfluenceVort.h), its own seed
    r3 = np.random.default_rng(20241203)
    vn, vt = 400, 350
    vsx, vsy = r3.uniform(-1, 1, vn), r3.uniform(-1, 1, vn)
    vss, vsr = r3.normal(0, 1, vn), r3.uniform(0.05, 0.2, vn)
    vtx, vty = r3.uniform(-1.2, 1.2, vt), r3.uniform(-1.2, 1.2, vt)
    d.update(vo_sx=vsx, vo_sy=vsy, vo_ss=vss, vo_sr=vsr, vo_tx=vtx, vo_ty=vty)
    d["vo_w"] = O.ref_points_vorticity(vsx, vsy, vss, vsr, vtx, vty)
    d["vo_wf"] = O.ref_points_vorticity(vsx, vsy, vss, vsr, vtx, vty, dtype=np.float32)

    # BEM blocks between two different bodies (src/Coefficients.h with &src != &targ, built as
    # src/BEMHelper.h:184-216 builds them); body kind 2 = rotating (augmented), 0 = no body
    ax_, ay_, ai_ = circle_panels(24)
    bx_, by_, bi_ = circle_panels(16, diam=0.6, cx=1.6, cy=0.3)
    d.update(bp_ax=ax_, bp_ay=ay_, bp_ai=ai_, bp_bx=bx_, bp_by=by_, bp_bi=bi_)
    for nunk in (1, 2):
        for sk, tk in ((2, 2), (2, 0), (0, 2), (0, 0)):
            for tag, geo in (("ab", (ax_, ay_, ai_, bx_, by_, bi_)), ("ba", (bx_, by_, bi_, ax_, ay_, ai_))):
                R = O.ref_bem_block(geo[0], geo[1], geo[2], sk, geo[3], geo[4], geo[5], tk, nunk)
                key = f"bp_{tag}_{nunk}_{sk}{tk}"
                d[key + "_A"], d[key + "_uu"], d[key + "_uv"] = R["A"], R["urot_u"], R["urot_v"]

    # prescribed body motion (src/Body.cpp, src/ElementBase.h:250-273, src/Surfaces.h:492-523,
    # 728-752): a body-bound circle moved by tinyexpr expressions of t, transformed and its
    # motion added (factor -1) to zeroed panel velocities, at three times
    mx_, my_, mi_ = circle_panels(20, diam=0.5, cx=0.4, cy=-0.2)
    d.update(bm_nx=mx_, bm_ny=my_, bm_ni=mi_, bm_times=np.array([0.0, 0.37, 1.9]))
    motions = {1: ("0.3*sin(2*t)", "0.1*t*t", "0.5*t"), 2: ("1.5+0.1*sin(t)", "", "0.3*t"),
               3: ("", "", "7*t")}
    for kind, ex in motions.items():
        for ti, t in enumerate(d["bm_times"]):
            R = O.ref_body_motion(mx_, my_, mi_, *ex, t, -1.0)
            for k in ("x", "y", "tx", "ty", "pu", "pv", "body"):
                d[f"bm_{kind}_{ti}_{k}"] = R[k]

    # VTU files written by the reference's own writers (src/Points.h:556-740,
    # src/Surfaces.h:1238-1379, src/VtkXmlWriter.h), kept byte for byte under tests/golden/vtk/
    r4 = np.random.default_rng(20241204)
    kn = 37
    kx, ky = r4.normal(0, 1, kn), r4.normal(0, 1, kn)
    ks, kr = r4.normal(0, 1e-2, kn), r4.uniform(0.01, 0.02, kn)
    ku, kv, kw = r4.normal(0, 1, kn), r4.normal(0, 1, kn), r4.normal(0, 10, kn)
    d.update(vk_x=kx, vk_y=ky, vk_s=ks, vk_r=kr, vk_u=ku, vk_v=kv, vk_w=kw)
    pnx, pny, pidx = circle_panels(12)
    pvs, pss = r4.normal(0, 1, 12), r4.normal(0, 1, 12)
    d.update(vk_pvs=pvs, vk_pss=pss)
    vdir = os.path.join(HERE, "vtk")
    os.makedirs(vdir, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        names = [O.ref_write_points_vtk(tmp, kx, ky, ks, kr, ku, kv, kw, index=0, frameno=12,
                                        time=0.125),
                 O.ref_write_points_vtk(tmp, kx, ky, None, None, ku, kv, None, inert=True,
                                        index=1, frameno=12, time=0.125),
                 O.ref_write_panels_vtk(tmp, pnx, pny, pidx, pvs, pss, index=0, frameno=3,
                                        time=2.5)]
        for nm in names:
            shutil.copyfile(os.path.join(tmp, nm), os.path.join(vdir, nm))

    np.savez_compressed(OUT, **d)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes, {len(d)} arrays)")


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "cores":
        core_vectors()
    else:
        main()
        core_vectors()
