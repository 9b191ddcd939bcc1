"""ctypes declarations of the C ABI in include/omega2d_hip.h.

The library is loaded from omega2d_amd/lib/libomega2d_hip.so (built in-tree by
``python -m omega2d_amd.build``). There is no fallback: a missing library raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# O2D_HIP_LIB selects another build of the same library (used for launch-shape tuning)
LIB_PATH = os.environ.get("O2D_HIP_LIB") or os.path.join(_HERE, "lib", "libomega2d_hip.so")

# values from include/omega2d_hip.h
O2D_OK = 0
ERRORS = {-1: "O2D_ERR_ARG", -2: "O2D_ERR_SIZE", -3: "O2D_ERR_CORE", -4: "O2D_ERR_RESULTS",
          -5: "O2D_ERR_PRECISION", -6: "O2D_ERR_HIP", -7: "O2D_ERR_STATE"}
CORE_RM, CORE_GAUSSIAN, CORE_WL, CORE_V2, CORE_V3 = 1, 2, 3, 4, 5
VELONLY, VELANDVORT, VELANDSHEAR = 1, 4, 5
PLAIN_FP64, SPLIT_FP32 = 1, 2
KT_NBODY_SPLIT, KT_NBODY_PLAIN, KT_PANELS, KT_PANEL_RHS, KT_BEM_ASSEMBLE, KT_ADJOINT, \
    KT_SPLIT_PREP, KT_NBODY_SPLIT_PT, KT_NBODY_PLAIN_PT, KT_NEAR, KT_VORT, KT_TREE_BUILD, \
    KT_TREE_EVAL = 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13

CORES = {"rm": CORE_RM, "gauss": CORE_GAUSSIAN, "gaussian": CORE_GAUSSIAN, "wl": CORE_WL,
         "v2": CORE_V2, "v3": CORE_V3}
RESULTS = {"vel": VELONLY, "vort": VELANDVORT, "shear": VELANDSHEAR}
PRECISIONS = {"fp64": PLAIN_FP64, "plain": PLAIN_FP64, "split": SPLIT_FP32, "fp32": SPLIT_FP32}


class Env(ctypes.Structure):
    """o2d_env: core, precision, results, device, stream."""
    _fields_ = [("core", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("results", ctypes.c_int32), ("device", ctypes.c_int32),
                ("stream", ctypes.c_void_p)]


class O2DError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} returned {ERRORS.get(code, code)}: {msg}")
        self.code = code


class TreeOpts(ctypes.Structure):
    """o2d_tree_opts: theta, order, pad, core_tol."""
    _fields_ = [("theta", ctypes.c_double), ("order", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("core_tol", ctypes.c_double)]


_i64, _i32, _dbl, _vp = ctypes.c_int64, ctypes.c_int32, ctypes.c_double, ctypes.c_void_p
_envp = ctypes.POINTER(Env)
_optp = ctypes.POINTER(TreeOpts)

# name -> argtypes (every function returns int unless listed in _RESTYPES)
SIGNATURES = {
    "o2d_hip_init": [_i32],
    "o2d_hip_release": [],
    "o2d_hip_last_error_string": [],
    "o2d_hip_abi_version": [],
    "o2d_hip_flops_per_pair": [_i32, _i32, _i32],
    "o2d_hip_timing_enable": [_i32],
    "o2d_hip_timing_read": [_i32, ctypes.POINTER(_dbl), ctypes.POINTER(_i64)],
    "o2d_hip_points_on_points_d": [_envp, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp,
                                   _vp, _vp],
    "o2d_hip_points_on_points_f": [_envp, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp,
                                   _vp, _vp],
    "o2d_hip_points_on_points_dev_d": [_envp, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp,
                                       _vp, _vp, _vp],
    "o2d_hip_points_on_points_noalloc_dev_d": [_envp, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp,
                                               _vp, _vp, _vp, _vp, _vp],
    "o2d_hip_panels_on_points_d": [_envp, _i64, _vp, _vp, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp,
                                   _vp],
    "o2d_hip_panels_on_points_dev_d": [_envp, _i64, _vp, _vp, _i64, _vp, _vp, _vp, _i64, _vp, _vp,
                                       _vp, _vp],
    "o2d_hip_points_on_panels_d": [_envp, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _i64, _vp, _vp, _vp],
    "o2d_hip_points_on_panels_dev_d": [_envp, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _i64, _vp, _vp,
                                       _vp],
    "o2d_hip_bem_assemble_d": [_envp, _i64, _vp, _vp, _i64, _vp, _i32, _i32, _vp, _vp, _dbl, _vp,
                               _i64],
    "o2d_hip_bem_assemble_dev_d": [_envp, _i64, _vp, _vp, _i64, _vp, _i32, _i32, _vp, _vp, _dbl,
                                   _vp, _i64],
    "o2d_hip_bem_assemble_pair_d": [_envp, _i64, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _i64, _vp,
                                    _i32, _i32, _i32, _vp, _vp, _vp, _i64],
    "o2d_hip_bem_assemble_pair_dev_d": [_envp, _i64, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _i64,
                                        _vp, _i32, _i32, _i32, _vp, _vp, _vp, _i64],
    "o2d_hip_bem_factor_d": [_envp, _i64, _vp, _i64],
    "o2d_hip_bem_factor_dev_d": [_envp, _i64, _vp, _i64],
    "o2d_hip_bem_factors_d": [_envp, _i64, _vp, _vp],
    "o2d_hip_bem_solve_d": [_envp, _i64, _vp, _vp],
    "o2d_hip_bem_solve_dev_d": [_envp, _i64, _vp, _vp],
    "o2d_hip_rk_move_d": [_envp, _i64, _dbl, _i32, _vp, _vp, _vp, _vp, _vp],
    "o2d_hip_rk_move_dev_d": [_envp, _i64, _dbl, _i32, _vp, _vp, _vp, _vp, _vp],
    "o2d_hip_points_on_points_adjoint_d": [_envp, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp,
                                           _vp, _vp, _vp, _vp, _vp],
    "o2d_hip_points_on_points_adjoint_dev_d": [_envp, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp,
                                               _vp, _vp, _vp, _vp, _vp, _vp],
    "o2d_hip_morton_order_dev_d": [_envp, _i64, _vp, _vp, _dbl, _dbl, _dbl, _vp, _vp],
    "o2d_hip_points_on_points_tree_d": [_envp, _optp, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp,
                                        _vp, _vp, _vp],
    "o2d_hip_points_on_points_tree_dev_d": [_envp, _optp, _i64, _vp, _vp, _vp, _vp, _i64, _vp,
                                            _vp, _vp, _vp, _vp],
    "o2d_hip_points_vorticity_d": [_envp, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp],
    "o2d_hip_points_vorticity_dev_d": [_envp, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp],
    "o2d_hip_nearest_distances_d": [_envp, _i64, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _vp],
    "o2d_hip_nearest_distances_dev_d": [_envp, _i64, _vp, _vp, _i64, _vp, _i64, _vp, _vp, _vp],
    "o2d_hip_reflect_d": [_envp, _i64, _vp, _vp, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp],
    "o2d_hip_reflect_dev_d": [_envp, _i64, _vp, _vp, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp],
    "o2d_hip_clear_inner_d": [_envp, _i32, _dbl, _dbl, _i64, _vp, _vp, _i64, _vp, _vp, _vp, _i64,
                              _vp, _vp, _vp, _vp, _vp, _vp],
    "o2d_hip_clear_inner_dev_d": [_envp, _i32, _dbl, _dbl, _i64, _vp, _vp, _i64, _vp, _vp, _vp,
                                  _i64, _vp, _vp, _vp, _vp, _vp, _vp],
}
# the reference's external solver hook (counts by pointer, return value = flop count)
_ip = ctypes.POINTER(ctypes.c_int)
EXTERNAL_SIGNATURES = {
    "external_vel_solver_f_": [_ip, _vp, _vp, _vp, _vp, _ip, _vp, _vp, _vp, _vp],
    "external_vel_solver_tr_f_": [_ip, _vp, _vp, _vp, _vp, _ip, _vp, _vp, _vp, _vp, _vp],
    "external_vel_solver_d_": [_ip, _vp, _vp, _vp, _vp, _ip, _vp, _vp, _vp, _vp],
}
_RESTYPES = {"o2d_hip_last_error_string": ctypes.c_char_p, "o2d_hip_abi_version": _i32,
             "o2d_hip_flops_per_pair": _i32, "external_vel_solver_f_": ctypes.c_float,
             "external_vel_solver_tr_f_": ctypes.c_float, "external_vel_solver_d_": ctypes.c_float}

_lib = None


def lib():
    """Load the HIP library (once). Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} "
                               "(run python -m omega2d_amd.build)")
        # torch ships its own HIP runtime and asks for it by an unversioned name, so it cannot
        # reuse a runtime this library loaded first: two runtimes would then share one HSA
        # runtime and torch would find no GPU. Loading torch first makes this library bind to
        # torch's runtime (same soname), so the process has exactly one.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, args in {**SIGNATURES, **EXTERNAL_SIGNATURES}.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = L
    return _lib


def call(name, *args):
    """Call an entry point and raise O2DError on a non-zero status."""
    rc = getattr(lib(), name)(*args)
    if rc != O2D_OK:
        msg = lib().o2d_hip_last_error_string()
        raise O2DError(name, rc, msg.decode() if msg else "")
    return rc


def make_env(core="v2", precision="fp64", results="vel", device=0, stream=None):
    return Env(CORES[core] if isinstance(core, str) else int(core),
               PRECISIONS[precision] if isinstance(precision, str) else int(precision),
               RESULTS[results] if isinstance(results, str) else int(results),
               int(device), stream)


def timing_enable(on=True):
    call("o2d_hip_timing_enable", int(bool(on)))


def timing_read(kernel_class):
    """(total milliseconds, launches) recorded for a kernel class since the last read."""
    ms, n = ctypes.c_double(), ctypes.c_int64()
    call("o2d_hip_timing_read", int(kernel_class), ctypes.byref(ms), ctypes.byref(n))
    return ms.value, n.value
