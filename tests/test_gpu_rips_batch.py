"""GPU parity of the batched Rips cohomology path (points in, pairs out; the coboundary
matrices are built straight into the engine's dual layout): every cloud's pairs and essential
cells equal the oracle's twist reduction of the host-built Rips matrix.  Integer results:
bit-exact."""
import os

import numpy as np
import pytest

from oracle import oracle as O

import specseqplusplus_amd as S
from specseqplusplus_amd import configs as CF

pytestmark = pytest.mark.gpu
P = dict(dualize=1, ss_block_size=1 << 30)


@pytest.fixture(scope="module")
def h():
    hd = S.Handle(0)
    yield hd
    hd.close()


def oracle_of(X, r):
    m = S.vr_build(X, r, 2)
    mo = O.Matrix(m.colptr, m.rows, m.dims)
    pairs = O.twist_pairs(mo)
    ess = np.setdiff1d(np.arange(m.n), pairs.ravel())
    return m, pairs, ess


def pct(X, q):
    from scipy.spatial.distance import pdist
    return float(np.percentile(pdist(X.astype(np.float64)), q)) if len(X) > 1 else 0.5


def lattice(shape, jitter=0.0, seed=0):
    g = np.stack(np.meshgrid(*[np.arange(k) for k in shape], indexing="ij"), -1).reshape(-1, len(shape))
    g = g.astype(np.float64)
    if jitter:
        g += np.random.default_rng(seed).normal(0, jitter, g.shape)
    return g.astype(np.float32)


def mixed_clouds(seed):
    rng = np.random.default_rng(seed)
    out = []
    for i, n in enumerate([0, 1, 2, 3, 37, 90, 150, 64, 200, 5, 120, 80, 33, 170]):
        X = CF.cloud(rng, i % 3, n).astype(np.float32) if n else np.zeros((0, 3), np.float32)
        out.append((X, pct(X, 30)))
    out.append((lattice((4, 4, 3)), 1.5))            # runs of equal lengths (1, sqrt 2)
    out.append((lattice((5, 4, 3)), 2.0))            # longer runs, several per filtration
    rX = np.round(rng.random((60, 2)) * 4).astype(np.float32)
    out.append((np.concatenate([rX, np.zeros((60, 1), np.float32)], 1), 2.5))  # repeated points
    out.append((rng.random((40, 3)).astype(np.float32), 0.0))                 # no edges
    out.append((rng.random((30, 3)).astype(np.float32), 10.0))                # complete: 4060 triangles
    return out


def test_batched_cohomology_matches_oracle(h):
    clouds = mixed_clouds(7)
    outs = h.batch_reduce_points(clouds, 2, S.params(**P))
    for (X, r), (pairs, ess) in zip(clouds, outs):
        m, ref, ref_ess = oracle_of(X, r)
        assert pairs.shape == ref.shape and np.array_equal(pairs, ref), (len(X), r)
        assert np.array_equal(ess, ref_ess), (len(X), r)
        assert 2 * len(pairs) + len(ess) == m.n


def test_batched_cohomology_equals_general_path(h, monkeypatch):
    """The same clouds through the general path (primal build, sort-based anti-transpose)."""
    clouds = mixed_clouds(8)[4:12]
    fast = h.batch_reduce_points(clouds, 2, S.params(**P))
    monkeypatch.setenv("SSQ_RIPS_GENERAL", "1")
    slow = h.batch_reduce_points(clouds, 2, S.params(**P))
    for (a, ae), (b, be) in zip(fast, slow):
        assert np.array_equal(a, b) and np.array_equal(ae, be)


@pytest.mark.parametrize("budget", ["1", "3000000"])
def test_groups_split_by_memory_budget(h, monkeypatch, budget):
    """A small device-memory budget splits the batch into many groups (one cloud each at
    budget 1); pairs are unchanged."""
    clouds = mixed_clouds(9)
    ref = h.batch_reduce_points(clouds, 2, S.params(**P))
    monkeypatch.setenv("SSQ_RIPS_GROUP_BYTES", budget)
    got = h.batch_reduce_points(clouds, 2, S.params(**P))
    for (a, ae), (b, be) in zip(ref, got):
        assert np.array_equal(a, b) and np.array_equal(ae, be)


@pytest.mark.parametrize("streams", ["2", "16"])
def test_groups_on_many_streams(h, monkeypatch, streams):
    """The groups spread over several streams (each with its own engine and device region),
    one cloud per group: pairs are unchanged.  A handle of its own, so that the extra engines
    are released at its end."""
    clouds = mixed_clouds(11)
    ref = h.batch_reduce_points(clouds, 2, S.params(**P))
    monkeypatch.setenv("SSQ_RIPS_GROUP_BYTES", "1")
    monkeypatch.setenv("SSQ_RIPS_STREAMS", streams)
    h2 = S.Handle(0)
    try:
        got = h2.batch_reduce_points(clouds, 2, S.params(**P))
    finally:
        h2.close()
    for (a, ae), (b, be) in zip(ref, got):
        assert np.array_equal(a, b) and np.array_equal(ae, be)


def test_device_resident_points_and_pairs(h):
    """Points resident in HBM, pairs written to HBM with per-cloud offsets; the counts entry
    gives each cloud's edges and triangles."""
    clouds = mixed_clouds(10)
    pbase = np.concatenate([[0], np.cumsum([len(X) for X, _ in clouds])]).astype(np.int64)
    allp = np.concatenate([X for X, _ in clouds]).astype(np.float32)
    d_pts = h.device_array(allp)
    rad = np.array([r for _, r in clouds], np.float32)
    e, t = h.rips_count(d_pts, pbase, rad)
    cap = int(pbase[-1] + e.sum())
    d_pairs = h.device_empty(2 * cap, np.int64)
    off = h.batch_reduce_points_device(d_pts, pbase, rad, d_pairs, cap)
    allpairs = d_pairs.numpy(2 * off[-1]).reshape(-1, 2)
    for i, (X, r) in enumerate(clouds):
        m, ref, _ = oracle_of(X, r)
        assert int((m.dims == 1).sum()) == e[i] and int((m.dims == 2).sum()) == t[i]
        assert np.array_equal(allpairs[off[i]:off[i + 1]], ref), i


def test_large_clouds_near_the_size_limit(h):
    """Clouds of 2000-4096 points (the largest the path takes), sparse enough for the oracle."""
    rng = np.random.default_rng(11)
    clouds = []
    for i, n in enumerate([4096, 2048, 3000]):
        X = CF.cloud(rng, i % 3, n).astype(np.float32)
        clouds.append((X, pct(X, 1.5)))
    outs = h.batch_reduce_points(clouds, 2, S.params(**P))
    for (X, r), (pairs, ess) in zip(clouds, outs):
        m, ref, ref_ess = oracle_of(X, r)
        assert np.array_equal(pairs, ref) and np.array_equal(ess, ref_ess), len(X)


def test_oversized_cloud_uses_general_path(h):
    """A cloud above 4096 points goes through the general path (same pairs)."""
    rng = np.random.default_rng(12)
    X = rng.random((4100, 3)).astype(np.float32)
    r = 0.05
    (pairs, ess), = h.batch_reduce_points([(X, r)], 2, S.params(**P))
    _, ref, ref_ess = oracle_of(X, r)
    assert np.array_equal(pairs, ref) and np.array_equal(ess, ref_ess)
