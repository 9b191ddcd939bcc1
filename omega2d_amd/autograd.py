"""torch.autograd.Function over the adjoint entry point: a test and optimisation harness.

    u, v = probe_velocity(x, y, strength, radius, px, py, core="gauss")
    J = (u**2 + v**2).sum(); J.backward()

Forward: the fp64 velocity at point targets (probes) of the particles, with the reference's
1/(2 pi) factor (src/ElementBase.h:215-220) and zero freestream. Backward: the hand-written
adjoint kernel (o2d_hip_points_on_points_adjoint_dev_d), which recomputes the interaction
instead of storing the N x M matrix. Gradients flow to particle positions, strengths and core
radii; the probes are treated as fixed.
"""
import math

import torch

from . import device as dv

_SCALE = 0.5 / math.pi


class ProbeVelocity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, y, strength, radius, px, py, core):
        tu, tv = dv.points_on_points(x, y, strength, radius, px, py, None, core=core,
                                     precision="fp64")
        ctx.save_for_backward(x, y, strength, radius, px, py)
        ctx.core = core
        return tu * _SCALE, tv * _SCALE

    @staticmethod
    def backward(ctx, gu, gv):
        x, y, strength, radius, px, py = ctx.saved_tensors
        gu = (gu * _SCALE).contiguous()
        gv = (gv * _SCALE).contiguous()
        gx, gy, gs, gr = dv.points_on_points_adjoint(x, y, strength, radius, px, py, gu, gv,
                                                     core=ctx.core)
        return gx, gy, gs, gr, None, None, None


def probe_velocity(x, y, strength, radius, px, py, core="gauss"):
    """Velocity (u, v) at probes (px, py); differentiable in x, y, strength and radius."""
    return ProbeVelocity.apply(x, y, strength, radius, px, py, core)
