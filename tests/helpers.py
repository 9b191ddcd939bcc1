"""Small shared helpers for the tests (no test logic)."""
import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def circle_panels(npan, diam=1.0, cx=0.0, cy=0.0, external=True):
    """Node layout of SolidCircle::init_elements (src/BoundaryFeature.cpp:489-530): clockwise
    from theta = 0, nodes rounded to float as the reference stores them."""
    i = np.arange(npan, dtype=np.float32)
    th = 2.0 * np.pi * i.astype(np.float64) / np.float64(np.float32(npan))
    nx = (cx + 0.5 * diam * np.cos(th)).astype(np.float32)
    ny = (cy - 0.5 * diam * np.sin(th)).astype(np.float32)
    idx = np.stack([np.arange(npan), (np.arange(npan) + 1) % npan], 1).astype(np.uint32)
    if not external:
        idx = idx[:, ::-1].copy()
    return nx, ny, idx.ravel()


def particles(rng, n, lo=-1.0, hi=1.0, rad=(0.01, 0.05), smean=0.0, sstd=1.0):
    x = rng.uniform(lo, hi, n)
    y = rng.uniform(lo, hi, n)
    s = rng.normal(smean, sstd, n)
    r = rng.uniform(rad[0], rad[1], n)
    return x, y, s, r
