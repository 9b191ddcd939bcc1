"""The drop-in, seen from the reference's side: its own points_affect_points<float,double>, built
with EXTERNAL_VEL_SOLVE and linked against libomega2d_hip.so (oracle/ref_dropin.cpp, built by
oracle/Makefile), runs with ExecEnv(false, direct, cpu_x86), so src/Influence.h:74-98 hands each
sum to the library's external_vel_solver_f_ / _tr_f_ with no change to the reference's source.
The steps around the calls are find_vels' (src/Convection.h:185-213): zero the targets, visit
every source collection through InfluenceVisitor, finalize.

The hook reads its configuration from the environment once per process; these tests use its
defaults, the reference's compiled-in Vatistas n=2 core and the split form for float arrays.
Bar: 1e-6 relative L2 against the fp64 oracle of the same float inputs.
"""
import numpy as np
import pytest

import oracle as O
from helpers import particles, rel_l2

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not O.have_dropin(),
                                 reason="oracle/_ref/libo2d_ref_dropin.so not built")]


def _f32(*a):
    return [np.asarray(x, dtype=np.float32) for x in a]


def _oracle_sum(sets, tx, ty, tr):
    f64 = lambda a: None if a is None else np.asarray(a, dtype=np.float64)
    u = np.zeros(len(tx))
    v = np.zeros(len(tx))
    for st in sets:
        a, b = O.points_on_points("v2", *(f64(x) for x in st), f64(tx), f64(ty), f64(tr))
        u += a
        v += b
    return u, v


@pytest.mark.parametrize("particle_targets", [False, True])
def test_reference_find_vels_through_the_hook(particle_targets):
    rng = np.random.default_rng(61)
    # two vortex collections accumulating into one target collection, as find_vels sums them
    s1 = _f32(*particles(rng, 7000, rad=(0.01, 0.04)))
    s2 = _f32(*particles(rng, 3001, rad=(0.02, 0.05)))
    tx, ty, _, tr = _f32(*particles(rng, 2500, rad=(0.01, 0.04)))
    trr = tr if particle_targets else None
    u, v = O.dropin_find_vels([s1, s2], tx, ty, trr)
    ru, rv = _oracle_sum([s1, s2], tx, ty, trr)
    assert rel_l2(u.astype(np.float64), ru) <= 1e-6
    assert rel_l2(v.astype(np.float64), rv) <= 1e-6


def test_reference_find_vels_self_and_finalize():
    """A collection acting on itself plus a second one, finalized with a freestream:
    u = fs + u / 2 pi (src/ElementBase.h:215-220)."""
    rng = np.random.default_rng(62)
    s1 = _f32(*particles(rng, 5000, rad=(0.01, 0.03)))
    s2 = _f32(*particles(rng, 1200, rad=(0.01, 0.03)))
    fs = (1.0, -0.25)
    u, v = O.dropin_find_vels([s1, s2], s1[0], s1[1], s1[3], finalize=True, fs=fs)
    ru, rv = _oracle_sum([s1, s2], s1[0], s1[1], s1[3])
    fu, fv = O.finalize(ru, rv, fs)
    assert rel_l2(u.astype(np.float64), fu) <= 1e-6
    assert rel_l2(v.astype(np.float64), fv) <= 1e-6
