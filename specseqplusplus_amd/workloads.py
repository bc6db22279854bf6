"""Synthetic inputs of the batched and served streams (seeded, so every run sees the same data).

Point clouds follow phat-highclear/addons/rips.cpp:49-87 semantics once built: float32 points
in R^3, edges of length <= r, cliques up to dimension 2.  Three distributions, used in turn:
uniform in the unit cube, a noisy unit circle (sigma 0.05) and a noisy torus (R = 1, r = 0.3,
sigma 0.03).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np


def make_cloud(rng, kind, n):
    """n points of distribution `kind` (0 cube, 1 circle, 2 torus), float64."""
    if kind == 0:
        return rng.random((n, 3))
    if kind == 1:
        t = rng.random(n) * 2 * np.pi
        return np.stack([np.cos(t), np.sin(t), np.zeros(n)], 1) + rng.normal(0, 0.05, (n, 3))
    u = rng.random(n) * 2 * np.pi
    v = rng.random(n) * 2 * np.pi
    return np.stack([(1 + 0.3 * np.cos(v)) * np.cos(u), (1 + 0.3 * np.cos(v)) * np.sin(u), 0.3 * np.sin(v)], 1) + \
        rng.normal(0, 0.03, (n, 3))


def percentile_threshold(X, q):
    """The q-th percentile of the cloud's pairwise distances (float64)."""
    from scipy.spatial.distance import pdist
    return float(np.percentile(pdist(X.astype(np.float64)), q))


def batched_stream(count=2048, nmin=64, nmax=2048, seed=2024, q=30.0):
    """The batched stream: `count` clouds, point counts log-uniform in [nmin, nmax], the three
    distributions in turn, each cloud's threshold its q-th distance percentile.  Returns (list of
    float32 point arrays, list of thresholds)."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(nmin), np.log(nmax), count)).astype(int)
    pts = [make_cloud(rng, i % 3, int(ns[i])).astype(np.float32) for i in range(count)]
    with ThreadPoolExecutor(8) as ex:
        rs = list(ex.map(lambda X: percentile_threshold(X, q), pts))
    return pts, rs


def served_requests(seed, n_clouds, n_mats, mean_degree=20.0, nmin=64, nmax=4096, cmin=1e4, cmax=1e6):
    """The request pool of the served stream.  Clouds: nmin-nmax points (log-uniform, the three
    distributions), threshold the 30th distance percentile, lowered for large clouds so that the
    mean vertex degree is at most `mean_degree` (a 4096-point cloud at the 30th percentile has
    about 7e8 triangles, seconds of device time per request, which no 200-5000 req/s sweep can
    serve).  Matrices: alpha complexes of cmin-cmax cells (log-uniform) at the density of the
    100,000-point alpha configuration.  Returns ([(points, r)], [Matrix])."""
    from .alpha import alpha_complex
    rng = np.random.default_rng(seed)
    pts = []
    for i in range(n_clouds):
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        X = make_cloud(rng, i % 3, n).astype(np.float32)
        q = min(30.0, 100.0 * mean_degree / (n - 1))
        pts.append((X, percentile_threshold(X, q)))
    mats = []
    for i in range(n_mats):
        cols = np.exp(rng.uniform(np.log(cmin), np.log(cmax)))
        npts = max(64, int(cols / 28.5))
        P = rng.random((npts, 3)) * 100.0 * (npts / 1e5) ** (1 / 3)
        mats.append(alpha_complex(P, 90.0))
    return pts, mats
