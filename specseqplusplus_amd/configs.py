"""Synthetic inputs of the five workload configurations (seeded, numpy PCG64).

c1: 512 points i.i.d. U[0,1]^3, Rips r=0.25, dim<=2.
c2: 4096 Haar-random rotations (unit quaternions -> 3x3 matrices flattened to R^9), Rips
    r=1.0, dim<=2 (reduced with model=3, max_iteration=10, ss_block_size=102400,
    growth=1.05, queue_size=65536, level1/level2=1000/8000).
c4/c5 clouds: uniform cube, noisy circle (sigma 0.05), noisy torus (R=1, r=0.3, sigma 0.03).
"""
import numpy as np


def c1_points(seed=0):
    return np.random.default_rng(seed).random((512, 3)).astype(np.float32)


def haar_rotations(n, seed=0):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([
        1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
        2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
        2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], axis=1)
    return R.astype(np.float32)


def c2_points(seed=0):
    return haar_rotations(4096, seed)


C2_PARAMS = dict(dualize=0, model=3, max_iteration=10, ss_block_size=102400, growth=1.05, queue_size=65536,
                 level1=1000, level2=8000)
C1_PARAMS = dict(dualize=0, model=0, max_iteration=10, ss_block_size=102400, growth=1.05, queue_size=65536,
                 level1=1000, level2=2000)


def cloud(rng, kind, n):
    """kind 0: uniform cube; 1: noisy circle; 2: noisy torus."""
    if kind == 0:
        return rng.random((n, 3))
    if kind == 1:
        t = rng.random(n) * 2 * np.pi
        return np.stack([np.cos(t), np.sin(t), np.zeros(n)], 1) + rng.normal(0, 0.05, (n, 3))
    u = rng.random(n) * 2 * np.pi
    v = rng.random(n) * 2 * np.pi
    return np.stack([(1 + 0.3 * np.cos(v)) * np.cos(u), (1 + 0.3 * np.cos(v)) * np.sin(u), 0.3 * np.sin(v)], 1) + \
        rng.normal(0, 0.03, (n, 3))
