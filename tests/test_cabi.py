"""The C ABI library loads, exports every symbol include/omega2d_hip.h declares, and rejects bad
arguments with the documented status codes before touching a device (CPU only: no entry point
here reaches a kernel)."""
import ctypes
import os
import re

import numpy as np
import pytest

from omega2d_amd import _cabi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "omega2d_hip.h")


def declared(prefix=r"o2d_hip_[a-z0-9_]+"):
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(" + prefix + r")\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _cabi.lib()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_cabi.SIGNATURES), "ctypes table and header disagree"
    ext = declared(r"external_vel_solver_[a-z_]+")
    assert ext == ["external_vel_solver_d_", "external_vel_solver_f_", "external_vel_solver_tr_f_"]
    for n in ext:
        assert hasattr(L, n), n
    assert set(ext) == set(_cabi.EXTERNAL_SIGNATURES)


_EXT_SCRIPT = """
import ctypes, sys
import numpy as np
from omega2d_amd import _cabi
L = _cabi.lib()
a = np.ones(4, dtype=np.float32)
p = lambda x: x.ctypes.data_as(ctypes.c_void_p)
ns, nt = ctypes.c_int(int(sys.argv[1])), ctypes.c_int(int(sys.argv[2]))
print(L.external_vel_solver_f_(ctypes.byref(ns), p(a), p(a), p(a), p(a), ctypes.byref(nt), p(a),
                               p(a), p(a), p(a)))
"""


@pytest.mark.parametrize("env,ns,nt,expect", [
    ({}, 0, 3, 0.0),                                  # nothing to sum: no device needed
    ({}, 3, 0, 0.0),
    ({"O2D_EXT_CORE": "v7"}, 3, 3, -1.0),             # rejected before any device call
    ({"O2D_EXT_PRECISION": "fp16"}, 3, 3, -1.0),
    ({"O2D_EXT_DEVICE": "x"}, 3, 3, -1.0),
    ({}, -1, 3, -1.0),
])
def test_external_hook_configuration_and_counts(env, ns, nt, expect):
    import subprocess
    import sys
    e = dict(os.environ, PYTHONPATH=ROOT, **env)
    r = subprocess.run([sys.executable, "-c", _EXT_SCRIPT, str(ns), str(nt)], env=e,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert float(r.stdout.split()[-1]) == expect
    if expect < 0:
        assert "external_vel_solver" in r.stderr


def test_abi_version_and_flop_convention():
    L = _cabi.lib()
    assert L.o2d_hip_abi_version() == 1
    # reference convention (src/Kernels.h:81,95; src/CoreFunc.h flops_* counters)
    f = L.o2d_hip_flops_per_pair
    assert [f(c, 0, 0) for c in (1, 2, 3, 4, 5)] == [15, 19, 17, 19, 22]
    assert [f(c, 1, 0) for c in (1, 2, 3, 4, 5)] == [13, 17, 15, 16, 18]
    assert f(4, 1, 1) == 39 and f(4, 1, 2) == 47
    assert f(9, 1, 0) == -1


def _pp(env, ns=4, nt=3, **over):
    # arrays stay small: every call here must be rejected before any element is read
    a = np.ones(8)
    args = dict(ns=ns, sx=a, sy=a, ss=a, sr=a, nt=nt, tx=a, ty=a, tr=None, tu=a.copy(),
                tv=a.copy(), tw=None, te=None)
    args.update(over)
    p = lambda x: None if x is None else x.ctypes.data_as(ctypes.c_void_p)
    return _cabi.lib().o2d_hip_points_on_points_d(
        env, args["ns"], p(args["sx"]), p(args["sy"]), p(args["ss"]), p(args["sr"]), args["nt"],
        p(args["tx"]), p(args["ty"]), p(args["tr"]), p(args["tu"]), p(args["tv"]), p(args["tw"]),
        p(args["te"]))


def test_argument_errors_are_reported_not_raised():
    ok = _cabi.make_env("rm", "fp64", "vel")
    assert _pp(None) == -1
    assert _pp(ctypes.byref(_cabi.make_env(7, "fp64", "vel"))) == -3
    assert _pp(ctypes.byref(_cabi.make_env("rm", 9, "vel"))) == -5
    assert _pp(ctypes.byref(_cabi.make_env("rm", "fp64", 2))) == -4   # velandgrad unsupported
    assert _pp(ctypes.byref(_cabi.make_env("rm", "fp64", 3))) == -4   # psi unsupported
    assert _pp(ctypes.byref(ok), ns=-1) == -2
    assert _pp(ctypes.byref(ok), nt=1 << 31) == -2
    assert _pp(ctypes.byref(ok), sx=None) == -1
    assert _pp(ctypes.byref(ok), tu=None) == -1
    # vorticity requested without its output array
    assert _pp(ctypes.byref(_cabi.make_env("rm", "fp64", "vort"))) == -1
    # the split form also needs the vorticity array
    assert _pp(ctypes.byref(_cabi.make_env("rm", "split", "vort"))) == -1
    msg = _cabi.lib().o2d_hip_last_error_string().decode()
    assert "needs tw" in msg


def test_panel_index_out_of_range_rejected_on_host():
    L = _cabi.lib()
    env = _cabi.make_env("v2", "fp64", "vel")
    nx = np.zeros(3)
    idx = np.array([0, 1, 1, 3], dtype=np.uint32)  # node 3 does not exist
    one = np.ones(2)
    out = np.zeros(2)
    p = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    rc = L.o2d_hip_panels_on_points_d(ctypes.byref(env), 3, p(nx), p(nx), 2, p(idx), p(one), None,
                                      2, p(one), p(one), p(out), p(out.copy()))
    assert rc == -2
    assert "index" in L.o2d_hip_last_error_string().decode()


def test_bem_and_rk_argument_errors():
    L = _cabi.lib()
    env = _cabi.make_env("v2", "fp64", "vel")
    z = np.zeros(8)
    p = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    idx = np.array([0, 1, 1, 0], dtype=np.uint32)
    # nunk must be 1 or 2
    assert L.o2d_hip_bem_assemble_d(ctypes.byref(env), 2, p(z), p(z), 2, p(idx), 3, 0, None, None,
                                    0.0, p(z), 2) == -4
    # lda smaller than n
    assert L.o2d_hip_bem_assemble_d(ctypes.byref(env), 2, p(z), p(z), 2, p(idx), 2, 1, p(z), p(z),
                                    0.0, p(z), 4) == -2
    # augmented without the unit-rotation velocities
    assert L.o2d_hip_bem_assemble_d(ctypes.byref(env), 2, p(z), p(z), 2, p(idx), 1, 1, None, None,
                                    0.0, p(z), 3) == -1
    w = np.ones(1)
    ptrs = (ctypes.c_void_p * 1)(z.ctypes.data)
    assert L.o2d_hip_rk_move_d(ctypes.byref(env), 8, 0.1, 4, p(w), ptrs, ptrs, p(z), p(z)) == -4
    # the adjoint is fp64 only
    senv = _cabi.make_env("gauss", "split", "vel")
    assert L.o2d_hip_points_on_points_adjoint_d(ctypes.byref(senv), 1, *([p(z)] * 4), 1, p(z), p(z),
                                                p(z), p(z), *([p(z)] * 4)) == -5


def test_timing_read_rejects_unknown_class():
    L = _cabi.lib()
    ms, n = ctypes.c_double(), ctypes.c_int64()
    assert L.o2d_hip_timing_read(0, ctypes.byref(ms), ctypes.byref(n)) == -1
    assert L.o2d_hip_timing_read(99, ctypes.byref(ms), ctypes.byref(n)) == -1


def test_python_binding_raises_on_error():
    import omega2d_amd as o2d
    with pytest.raises(o2d.O2DError) as ei:
        o2d.points_on_points(np.ones(2), np.ones(2), np.ones(2), np.ones(2), np.ones(2),
                             np.ones(2), core="rm", precision="split", results=3)
    assert ei.value.code == -4  # streamfunction is not a result of this path
    with pytest.raises(TypeError):
        o2d.points_on_points(np.ones(2, np.int64), np.ones(2), np.ones(2), np.ones(2), np.ones(2),
                             np.ones(2))
