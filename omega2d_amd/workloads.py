"""Synthetic inputs of the five benchmark configurations (numpy, seeded, fp64 unless stated).

Every choice a configuration leaves open is fixed here and named in the docstring.

C1  10 000 particles in a disc of radius 0.5, Gaussian vorticity, Rosenhead-Moore, RK2
C2  2^20 particles in 64 Lamb-Oseen blobs + 65 536 tracers, Gaussian core (the headline)
C3  circle of 2 048 panels + 200 000 wake particles, Vatistas n=2
C4  250 000 Gaussian-core particles, 4 096 probes on a circle (adjoint)
C5  2^22 particles in a perturbed shear layer, Rosenhead-Moore, RK3
"""
import numpy as np


def c1_disc(n=10000, seed=0):
    """Disc of radius 0.5, positions uniform in area (r = 0.5 sqrt(U), theta = 2 pi U'),
    strengths exp(-r^2/0.05) dA normalised to total circulation 1, sigma = 0.02, RM core,
    targets = sources, freestream (0, 0), one Ralston RK2 step with dt = 0.01."""
    rng = np.random.default_rng(seed)
    r = 0.5 * np.sqrt(rng.uniform(size=n))
    th = 2.0 * np.pi * rng.uniform(size=n)
    x, y = r * np.cos(th), r * np.sin(th)
    g = np.exp(-r * r / 0.05) * (np.pi * 0.25 / n)
    g *= 1.0 / np.sum(g)
    return dict(name="C1-disc-10k", core="rm", x=x, y=y, s=g, r=np.full(n, 0.02), dt=0.01,
                fs=(0.0, 0.0), order=2)


def c2_blobs(n=1 << 20, ntracer=1 << 16, nblob=64, seed=2, sigma=None):
    """64 Lamb-Oseen blobs: centres U[-4,4]^2, radii R_b U[0.1,0.5], blob circulation
    Gamma_b U[-1,1]; n/64 particles per blob at c_b + (R_b/sqrt 2) N(0, I) (vorticity
    ~ exp(-r^2/R_b^2)); strengths Gamma_b/(n/64) + N(0, 1e-4); core radius
    0.01 sqrt(64/n) (= 7.8125e-5 at n = 2^20, the literal reading of the configuration);
    Gaussian (exponential) core with its compact cutoff; 65 536 tracers U[-5,5]^2."""
    rng = np.random.default_rng(seed)
    per = n // nblob
    cx = rng.uniform(-4, 4, nblob)
    cy = rng.uniform(-4, 4, nblob)
    rb = rng.uniform(0.1, 0.5, nblob)
    gb = rng.uniform(-1, 1, nblob)
    x = np.empty(n)
    y = np.empty(n)
    s = np.empty(n)
    for b in range(nblob):
        sl = slice(b * per, (b + 1) * per)
        x[sl] = cx[b] + rb[b] / np.sqrt(2.0) * rng.normal(size=per)
        y[sl] = cy[b] + rb[b] / np.sqrt(2.0) * rng.normal(size=per)
        s[sl] = gb[b] / per
    s += rng.normal(0.0, 1e-4, n)
    sig = 0.01 * np.sqrt(64.0 / n) if sigma is None else sigma
    tx = rng.uniform(-5, 5, ntracer)
    ty = rng.uniform(-5, 5, ntracer)
    return dict(name="C2-blobs-1M", core="gauss", x=x, y=y, s=s, r=np.full(n, sig), tx=tx, ty=ty)


def c3_circle_wake(npan=2048, nwake=200000, seed=3):
    """Unit-diameter circle, nodes (0.5 cos(2 pi i/P), -0.5 sin(2 pi i/P)) stored as float
    (clockwise, src/BoundaryFeature.cpp:504-513), ground body (reactive, augmented);
    freestream (1, 0); wake r = sqrt(U(0.52^2, 3^2)), theta U(-pi/2, pi/2) (the downstream
    half-plane), strengths N(0, 2e-4), core radius 0.005, Vatistas n=2 core."""
    rng = np.random.default_rng(seed)
    i = np.arange(npan, dtype=np.float32)
    th = 2.0 * np.pi * i.astype(np.float64) / np.float64(np.float32(npan))
    nx = (0.5 * np.cos(th)).astype(np.float32).astype(np.float64)
    ny = (-0.5 * np.sin(th)).astype(np.float32).astype(np.float64)
    idx = np.stack([np.arange(npan), (np.arange(npan) + 1) % npan], 1).astype(np.uint32).ravel()
    r = np.sqrt(rng.uniform(0.52**2, 3.0**2, nwake))
    t = rng.uniform(-np.pi / 2, np.pi / 2, nwake)
    return dict(name="C3-circle-wake", core="v2", nx=nx, ny=ny, idx=idx, x=r * np.cos(t),
                y=r * np.sin(t), s=rng.normal(0, 2e-4, nwake), r=np.full(nwake, 0.005),
                fs=(1.0, 0.0))


def c4_adjoint(n=250000, nprobe=4096, seed=5):
    """Particles U[-1,1]^2, strengths N(0, 1e-3), core radii U[0.005, 0.02], Gaussian core;
    probes on the circle of radius 1.5 (point targets); J = sum_k |u_k|^2 with the 1/(2 pi)
    factor and zero freestream."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, n)
    y = rng.uniform(-1, 1, n)
    s = rng.normal(0, 1e-3, n)
    r = rng.uniform(0.005, 0.02, n)
    k = np.arange(nprobe)
    return dict(name="C4-adjoint-250k", core="gauss", x=x, y=y, s=s, r=r,
                tx=1.5 * np.cos(2 * np.pi * k / nprobe), ty=1.5 * np.sin(2 * np.pi * k / nprobe))


def c5_shear_layer(n=1 << 22, seed=4):
    """x U[0,8], y = 0.01 sin(2 pi x) + 0.05 (U - 0.5) (a uniform slab of thickness 0.05),
    strengths 8/n + N(0, 1e-6), RM core radius 0.002, Ralston RK3 with dt = 0.01."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 8, n)
    y = 0.01 * np.sin(2 * np.pi * x) + 0.05 * (rng.uniform(size=n) - 0.5)
    s = 8.0 / n + rng.normal(0, 1e-6, n)
    return dict(name="C5-shear-4M", core="rm", x=x, y=y, s=s, r=np.full(n, 0.002), dt=0.01,
                fs=(0.0, 0.0), order=3)


# Ralston schemes of src/Convection.h:390-598: stage offsets and final weights
RK_STAGES = {
    1: dict(c=[0.0], w=[1.0]),
    2: dict(c=[0.0, 2.0 / 3.0], w=[0.25, 0.75]),
    3: dict(c=[0.0, 0.5, 0.75], w=[2.0 / 9.0, 3.0 / 9.0, 4.0 / 9.0]),
}
