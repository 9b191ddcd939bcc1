"""The no-allocation particle sum (o2d_hip_points_on_points_noalloc_dev_d) against the entry
point it varies (o2d_hip_points_on_points_dev_d), through the C ABI on device pointers.

Two calls in a row, the second smaller, must each agree with the buffer-pool entry to the
precision form's bar (fp64 <= 1e-12, split <= 1e-6 relative L2: the variant may cut the source
range into fewer chunks, which changes the fp64 summation grouping only), and sizes beyond the
stated maximum must be refused with O2D_ERR_SIZE.
"""
import ctypes

import numpy as np
import pytest
import torch

from omega2d_amd import _cabi
from omega2d_amd.device import env_for
from helpers import rel_l2

pytestmark = pytest.mark.gpu

BAR = {"fp64": 1e-12, "split": 1e-6}
MAX_POINTS = 262144


def _blobs(rng, n):
    c = rng.uniform(-2, 2, (6, 2))
    b = rng.integers(0, 6, n)
    return (c[b, 0] + 0.15 * rng.normal(size=n), c[b, 1] + 0.15 * rng.normal(size=n),
            rng.normal(0, 1e-3, n), rng.uniform(0.005, 0.02, n))


def _sum(entry, core, precision, results, src, tgt, tr):
    dev = torch.device("cuda", 0)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    sx, sy, ss, sr = (T(a) for a in src)
    tx, ty = T(tgt[0]), T(tgt[1])
    trd = None if tr is None else T(tr)
    nt = tx.numel()
    tu = torch.zeros(nt, dtype=torch.float64, device=dev)
    tv = torch.zeros_like(tu)
    aux = torch.zeros_like(tu)
    env = env_for(dev, core, precision, results)
    p = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())  # noqa: E731
    tw = aux if results == "vort" else None
    te = aux if results == "shear" else None
    _cabi.call(entry, ctypes.byref(env), sx.numel(), p(sx), p(sy), p(ss), p(sr), nt, p(tx),
               p(ty), p(trd), p(tu), p(tv), p(tw), p(te))
    torch.cuda.synchronize()
    out = [tu.cpu().numpy(), tv.cpu().numpy()]
    if results != "vel":
        out.append(aux.cpu().numpy())
    return out


@pytest.mark.parametrize("core,results", [("gauss", "vel"), ("v2", "vort"), ("rm", "shear")])
@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_two_calls_match_the_buffer_pool_entry(core, results, precision):
    rng = np.random.default_rng(17)
    # first call: particles on themselves at the largest size; second: fewer particles on
    # separate point targets (the array keeps the first call's contents)
    cases = []
    x, y, s, r = _blobs(rng, MAX_POINTS)
    cases.append(((x, y, s, r), (x, y), r))
    x2, y2, s2, r2 = _blobs(rng, 50000)
    tx, ty, _, _ = _blobs(rng, 7000)
    cases.append(((x2, y2, s2, r2), (tx, ty), None))
    for src, tgt, tr in cases:
        got = _sum("o2d_hip_points_on_points_noalloc_dev_d", core, precision, results, src, tgt,
                   tr)
        ref = _sum("o2d_hip_points_on_points_dev_d", core, precision, results, src, tgt, tr)
        for g, e in zip(got, ref):
            assert np.all(np.isfinite(g))
            assert rel_l2(g, e) <= BAR[precision], (len(src[0]), rel_l2(g, e))


def test_sizes_beyond_the_maximum_are_refused():
    dev = torch.device("cuda", 0)
    a = torch.zeros(MAX_POINTS + 1, dtype=torch.float64, device=dev)
    env = env_for(dev, "rm", "fp64", "vel")
    p = ctypes.c_void_p(a.data_ptr())
    L = _cabi.lib()
    for ns, nt in ((MAX_POINTS + 1, 10), (10, MAX_POINTS + 1)):
        rc = L.o2d_hip_points_on_points_noalloc_dev_d(ctypes.byref(env), ns, p, p, p, p, nt, p, p,
                                                      None, p, p, None, None)
        assert rc == -2  # O2D_ERR_SIZE
        assert b"at most 262144" in L.o2d_hip_last_error_string()
