"""Small float32 fields for the 2-bit adaptive tree tests (seeded, closed forms)."""
import numpy as np


def sphere_box(shape=(40, 33, 27), pitch=1.0):
    x = np.indices(shape).astype(np.float64) * pitch
    c = np.array(shape, np.float64) * pitch * 0.45
    d_s = np.sqrt(((x - c[:, None, None, None]) ** 2).sum(0)) - 0.3 * min(shape) * pitch
    q = np.abs(x - (c * 1.3)[:, None, None, None]) - 0.15 * min(shape) * pitch
    d_b = np.sqrt((np.maximum(q, 0) ** 2).sum(0)) + np.minimum(q.max(0), 0)
    return np.minimum(d_s, d_b).astype(np.float32)


def noisy(shape=(19, 23, 17), seed=5):
    rng = np.random.default_rng(seed)
    base = sphere_box(shape)
    return (base + rng.normal(0, 0.8, shape)).astype(np.float32)


def planes(shape=(16, 16, 16)):
    x = np.indices(shape).astype(np.float32)
    return (x[0] - 7.5).astype(np.float32)
