"""The reference's external solver hook (src/Influence.h:31-38, 74-98) on the GPU: the same sums as
o2d_hip_points_on_points_f/_d with the default configuration (V2 core; split form for float
arrays, fp64 for double), accumulated unscaled, with the reference's flop count returned."""
import ctypes

import numpy as np
import pytest

import oracle as O
import omega2d_amd as o2d
from omega2d_amd import _cabi
from helpers import particles, rel_l2

pytestmark = pytest.mark.gpu


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _counts(ns, nt):
    return ctypes.byref(ctypes.c_int(ns)), ctypes.byref(ctypes.c_int(nt))


@pytest.mark.parametrize("particle_targets", [False, True])
def test_float_hook_matches_entry_point_and_oracle(particle_targets):
    L = _cabi.lib()
    rng = np.random.default_rng(21)
    sx, sy, ss, sr = (a.astype(np.float32) for a in particles(rng, 5000, rad=(0.01, 0.05)))
    tx, ty, _, tr = (a.astype(np.float32) for a in particles(rng, 3000, rad=(0.01, 0.05)))
    tu, tv = np.zeros(3000, np.float32), np.zeros(3000, np.float32)
    pns, pnt = _counts(5000, 3000)
    trr = tr if particle_targets else None

    def hook():
        if particle_targets:
            return L.external_vel_solver_tr_f_(pns, _p(sx), _p(sy), _p(ss), _p(sr), pnt, _p(tx),
                                               _p(ty), _p(tr), _p(tu), _p(tv))
        return L.external_vel_solver_f_(pns, _p(sx), _p(sy), _p(ss), _p(sr), pnt, _p(tx),
                                        _p(ty), _p(tu), _p(tv))

    flops = hook()
    per = L.o2d_hip_flops_per_pair(4, 0 if particle_targets else 1, 0)
    assert flops == pytest.approx(3000.0 * (2.0 + per * 5000.0), rel=1e-6)
    f64 = lambda a: None if a is None else a.astype(np.float64)
    ru, rv = O.points_on_points("v2", *(f64(a) for a in (sx, sy, ss, sr, tx, ty)), f64(trr))
    assert rel_l2(tu.astype(np.float64), ru) <= 1e-6
    assert rel_l2(tv.astype(np.float64), rv) <= 1e-6
    # a second call accumulates exactly as the entry point does
    eu, ev = o2d.points_on_points(sx, sy, ss, sr, tx, ty, trr, core="v2", precision="split",
                                  tu=tu.copy(), tv=tv.copy())
    hook()
    assert np.array_equal(tu, eu) and np.array_equal(tv, ev)


def test_double_hook_is_the_fp64_sum():
    L = _cabi.lib()
    rng = np.random.default_rng(22)
    sx, sy, ss, sr = particles(rng, 4000, rad=(0.01, 0.05))
    tx, ty, _, _ = particles(rng, 2500)
    tu, tv = np.zeros(2500), np.zeros(2500)
    pns, pnt = _counts(4000, 2500)
    flops = L.external_vel_solver_d_(pns, _p(sx), _p(sy), _p(ss), _p(sr), pnt, _p(tx), _p(ty),
                                     _p(tu), _p(tv))
    assert flops > 0
    ru, rv = O.points_on_points("v2", sx, sy, ss, sr, tx, ty)
    assert rel_l2(tu, ru) <= 1e-12 and rel_l2(tv, rv) <= 1e-12
    # accumulates: a second call doubles the sums
    L.external_vel_solver_d_(pns, _p(sx), _p(sy), _p(ss), _p(sr), pnt, _p(tx), _p(ty), _p(tu),
                             _p(tv))
    assert rel_l2(tu, 2 * ru) <= 1e-12 and rel_l2(tv, 2 * rv) <= 1e-12


def test_particle_target_hook_needs_radii():
    L = _cabi.lib()
    a = np.ones(4, dtype=np.float32)
    pns, pnt = _counts(4, 4)
    assert L.external_vel_solver_tr_f_(pns, _p(a), _p(a), _p(a), _p(a), pnt, _p(a), _p(a), None,
                                       _p(a.copy()), _p(a.copy())) < 0
