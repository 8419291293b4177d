"""Shared numeric helpers of the test-suite."""
import numpy as np


def rel_err(a, b):
    """max |a - b| / max |b| (normwise relative, infinity norm)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / den


def l2_rms(a, b):
    return float(np.sqrt(np.mean((np.asarray(a) - np.asarray(b)) ** 2)))


def smooth_fields(nx, ny, seed=0, nbump=6):
    """Seeded smooth test fields: phi in [0,1] with curved interfaces, T in [0,0.6]."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:ny, 0:nx].astype(np.float64)
    phi = np.zeros((ny, nx))
    for _ in range(nbump):
        cx, cy = rng.uniform(0, nx), rng.uniform(0, ny)
        r0 = rng.uniform(0.08, 0.2) * min(nx, ny)
        ang = np.arctan2(y - cy, x - cx)
        r = np.hypot(x - cx, y - cy) * (1 + 0.15 * np.cos(rng.integers(3, 7) * ang))
        phi = np.maximum(phi, 0.5 * (1 + np.tanh((r0 - r) / 2.5)))
    T = 0.3 + 0.3 * np.sin(2 * np.pi * x / nx) * np.cos(2 * np.pi * y / ny) * rng.uniform(0.5, 1)
    return np.ascontiguousarray(phi), np.ascontiguousarray(T)
