"""The batched and served streams at the specified sizes.

Batched: 2048 Rips complexes, point counts log-uniform in [64, 2048] (cube / circle / torus,
each at its 30th distance percentile, dim <= 2), reduced from points in HBM in one call.  Every
cloud's pairs are checked at full size against invariants that hold whatever the order of
reduction: births strictly increasing, every cell in at most one pair, birth < death < cells,
at most one pair per vertex and edge, and the number of vertex births (rank of the edge
boundary) equal to n minus the number of connected components (an independent graph count).  A
sample, including the cloud with the most triangles and clouds near 2048 points of each
distribution, is compared bit for bit with the oracle's cohomology twist reduction.

Served: a Poisson stream of point clouds (64-4096 points) and alpha boundary matrices (1e4-1e6
cells) through the server; every result is compared with the oracle's pairs for its request."""
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle as O

import specseqplusplus_amd as S
from specseqplusplus_amd import workloads as W
from specseqplusplus_amd.server import Server

pytestmark = pytest.mark.gpu

BATCH = dict(dualize=1, model=0, ss_block_size=1 << 30, growth=1.05, queue_size=65536, level1=64, level2=1024)
SERVED_PTS = dict(dualize=1, model=0, ss_block_size=1 << 30, growth=1.05, queue_size=16384, level1=64, level2=1024)
SERVED_MAT = dict(dualize=0, model=0, ss_block_size=1 << 30, growth=1.05, queue_size=65536, level1=1000, level2=16000)


def components(X, r):
    """Connected components of the graph of edges with float32 length <= r (the builder's rule:
    squared differences summed in double, square root, rounded to float32)."""
    from scipy.sparse.csgraph import connected_components
    from scipy.spatial.distance import pdist, squareform
    if len(X) < 2:
        return len(X)
    D = squareform(pdist(X.astype(np.float64))).astype(np.float32)
    return connected_components(D <= np.float32(r), directed=False)[0]


def say(msg):
    print(f"[workloads] {msg}", flush=True)


def test_batched_stream_full_size():
    t = time.time()
    pts, rs = W.batched_stream(2048, 64, 2048, 2024)
    say(f"2048 clouds generated in {time.time() - t:.1f} s")
    h = S.Handle(0)
    try:
        pbase = np.concatenate([[0], np.cumsum([len(X) for X in pts])]).astype(np.int64)
        d_pts = h.device_array(np.concatenate(pts).astype(np.float32))
        rad = np.array(rs, np.float32)
        edges, tris = h.rips_count(d_pts, pbase, rad)
        assert int(tris.sum()) > 10 ** 10  # the specified size: ~1.9e10 triangles
        cap = int(pbase[-1] + edges.sum())
        d_pairs = h.device_empty(2 * cap, np.int64)
        t = time.time()
        off, st = h.batch_reduce_points_device(d_pts, pbase, rad, d_pairs, cap, S.params(**BATCH), stats=True)
        say(f"reduced {int(pbase[-1] + edges.sum() + tris.sum())} cells in {time.time() - t:.1f} s")
        allp = d_pairs.numpy(2 * int(off[-1])).reshape(-1, 2)
    finally:
        h.close()
    per = [allp[off[i]:off[i + 1]] for i in range(len(pts))]

    def invariants(i):
        n, pr = len(pts[i]), per[i]
        cells = n + int(edges[i]) + int(tris[i])
        if len(pr) == 0:
            return n <= 1
        b, d = pr[:, 0], pr[:, 1]
        ok = bool((np.diff(b) > 0).all() and (b < d).all() and d.max() < cells and b.min() >= 0)
        ok &= len(np.unique(pr)) == 2 * len(pr)
        ok &= len(pr) <= n + int(edges[i])  # births are vertices or edges (dim <= 2)
        ok &= cells - 2 * len(pr) >= 1  # N = 2 #pairs + #essential, with at least one class
        ok &= int((b < n).sum()) == n - components(pts[i], rs[i])  # #pairs_0 = rank of the edge boundary
        return ok

    with ThreadPoolExecutor(16) as ex:
        bad = [i for i, ok in enumerate(ex.map(invariants, range(len(pts)))) if not ok]
    assert not bad, f"clouds failing the invariants: {bad[:10]}"
    say("invariants hold for all 2048 clouds")

    # oracle sample: the most triangles, the largest cloud of each distribution, random ones
    ns = np.array([len(X) for X in pts])
    sample = {int(np.argmax(tris))}
    for k in range(3):
        idx = np.arange(k, len(pts), 3)
        sample.add(int(idx[np.argmax(ns[idx])]))
    sample |= {int(i) for i in np.random.default_rng(3).choice(len(pts), 4, replace=False)}
    assert max(ns[list(sample)]) > 1900
    hb = S.Handle(0)
    try:
        mats = {i: hb.vr_build_gpu(pts[i], rs[i], 2) for i in sorted(sample, key=lambda i: tris[i])}
    finally:
        hb.close()

    def oracle(i):
        m = mats[i]
        assert m.n == len(pts[i]) + edges[i] + tris[i]
        return i, O.twist_pairs(O.Matrix(m.colptr, m.rows, m.dims), dual=True)

    t = time.time()
    with ThreadPoolExecutor(len(sample)) as ex:
        for i, ref in ex.map(oracle, sorted(sample)):
            assert np.array_equal(per[i], ref), f"cloud {i} ({ns[i]} points) differs from the oracle"
    say(f"oracle sample {sorted(sample)} ({int(sum(mats[i].n for i in sample))} cells) bit-exact in "
        f"{time.time() - t:.1f} s")


def test_served_stream_every_result_against_oracle():
    pts, mats = W.served_requests(seed=31, n_clouds=64, n_mats=6)
    assert max(len(X) for X, _ in pts) > 3000 and max(m.n for m in mats) > 300_000

    def ref_pts(c):
        m = S.vr_build(c[0], c[1], 2)
        return O.twist_pairs(O.Matrix(m.colptr, m.rows, m.dims), dual=True)

    def ref_mat(m):
        return O.twist_pairs(O.Matrix(m.colptr, m.rows, m.dims))

    with ThreadPoolExecutor(16) as ex:
        refs = (list(ex.map(ref_pts, pts)), list(ex.map(ref_mat, mats)))
    rng = np.random.default_rng(5)
    nreq, qps = 3000, 1500.0
    kinds = rng.integers(0, 2, nreq)
    picks = rng.integers(0, 1 << 30, nreq)
    gaps = rng.exponential(1.0 / qps, nreq)
    with Server(streams=4, max_batch=64, max_batch_cells=2_000_000, p=S.params(**SERVED_MAT),
                points_p=S.params(**SERVED_PTS)) as srv:
        tickets = []
        nxt = time.perf_counter()
        for k in range(nreq):
            nxt += gaps[k]
            while time.perf_counter() < nxt:
                time.sleep(0.0003)
            if kinds[k] == 0:
                j = int(picks[k] % len(pts))
                tickets.append((0, j, srv.submit_points(pts[j][0], pts[j][1], 2)))
            else:
                j = int(picks[k] % len(mats))
                tickets.append((1, j, srv.submit_matrix(mats[j])))
        srv.wait_all()
        lat = []
        for kind, j, tk in tickets:
            pairs, ess, ms = srv.poll(tk)
            assert np.array_equal(pairs, refs[kind][j]), (kind, j)
            lat.append(ms)
    say(f"{nreq} requests at {qps} req/s offered: p50 {np.percentile(lat, 50):.1f} ms, "
        f"p99 {np.percentile(lat, 99):.1f} ms, all bit-exact")
