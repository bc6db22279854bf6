"""Serving: each result is the same alone and in the mix (bit-exact)."""
import ctypes
import time

import numpy as np
import pytest

import complexes as C
from oracle import oracle as O

import specseqplusplus_amd as S
from specseqplusplus_amd import _lib as L
from specseqplusplus_amd.server import Server

pytestmark = pytest.mark.gpu


def test_server_results_equal_alone():
    clouds, mats = [], []
    for seed in range(24):
        X = np.asarray(C.uniform_points(30 + 5 * seed, 3, 700 + seed), dtype=np.float32)
        clouds.append(X)
        mats.append(S.vr_build(X, 0.32, 2))
    with Server(streams=3, max_batch=8) as srv:
        tickets = []
        for i, (X, m) in enumerate(zip(clouds, mats)):
            if i % 2:
                tickets.append(srv.submit_points(X, 0.32, 2))
            else:
                tickets.append(srv.submit_matrix(m))
        srv.wait_all()
        for t, m in zip(tickets, mats):
            pairs, ess, lat = srv.poll(t)
            assert np.array_equal(pairs, O.twist_pairs(O.Matrix(m.colptr, m.rows, m.dims)))
            assert lat > 0


def test_server_mixed_dimensions_and_parameters():
    """Point clouds of two ambient dimensions and of simplex dimension 2 (GPU-built) and 3
    (host-built), interleaved with matrices, under a small cell cap and small batches: every
    batch holds one kind and one shape, and every result equals the oracle's pairs."""
    rng = np.random.default_rng(5)
    reqs = []
    for i in range(30):
        kind = i % 3
        if kind == 0:
            X = rng.random((20 + i, 3)).astype(np.float32)
            reqs.append(("pts", X, 0.35, 2))
        elif kind == 1:
            X = rng.random((12 + i // 2, 4)).astype(np.float32)
            reqs.append(("pts", X, 0.7, 3 if i % 2 else 2))
        else:
            X = rng.random((25 + i, 3)).astype(np.float32)
            reqs.append(("mat", S.vr_build(X, 0.3, 2), None, None))
    with Server(streams=2, max_batch=5, max_batch_cells=20000, p=S.params(),
                points_p=S.params(dualize=1, ss_block_size=1 << 30)) as srv:
        tickets = []
        for kind, a, r, md in reqs:
            tickets.append(srv.submit_points(a, r, md) if kind == "pts" else srv.submit_matrix(a))
        srv.wait_all()
        for (kind, a, r, md), t in zip(reqs, tickets):
            m = S.vr_build(a, r, md) if kind == "pts" else a
            pairs, ess, lat = srv.poll(t)
            assert np.array_equal(pairs, O.twist_pairs(O.Matrix(m.colptr, m.rows, m.dims))), (kind, md)


def test_server_bad_matrix_fails_alone():
    """A malformed matrix among good requests: the host checks reject it at submission (bad
    offsets), or the device check fails its batch (row out of range), which is then reduced
    request by request, so only the bad request reports the error; the others equal the
    oracle."""
    rng = np.random.default_rng(9)
    good = [S.vr_build(rng.random((30 + 3 * i, 3)).astype(np.float32), 0.35, 2) for i in range(8)]
    bad = S.Matrix(good[0].colptr.copy(), good[0].rows.copy(), good[0].dims.copy())
    bad.rows[-1] = bad.n + 5  # row past the end: caught on the device
    with Server(streams=1, max_batch=16) as srv:
        offs = S.Matrix(good[1].colptr.copy(), good[1].rows, good[1].dims)
        offs.colptr[-1] += 1  # offsets end past nnz: rejected at submission
        with pytest.raises(L.SSQError):
            srv.submit_matrix(offs)
        tickets = [srv.submit_matrix(m) for m in good[:4]] + [srv.submit_matrix(bad)] + \
                  [srv.submit_matrix(m) for m in good[4:]]
        srv.wait_all()
        results = []
        for t in tickets:
            try:
                results.append(srv.poll(t)[0])
            except L.SSQError as e:
                results.append(e)
    assert isinstance(results[4], L.SSQError) and results[4].status == L.SSQ_EFORMAT
    for m, r in zip(good, results[:4] + results[5:]):
        assert np.array_equal(r, O.twist_pairs(O.Matrix(m.colptr, m.rows, m.dims)))


def test_server_dense_cloud_under_a_small_budget():
    """A dense cloud (about 50 neighbours per point) among sparse ones with a device budget far
    below its size: admission counts its edges and triangles exactly at submission, it runs on
    its own, and every result equals the oracle's."""
    rng = np.random.default_rng(10)
    reqs = [(rng.random((40 + i, 3)).astype(np.float32), 0.2) for i in range(6)]
    Xd = rng.random((400, 3)).astype(np.float32)
    reqs.insert(3, (Xd, 0.35))
    with Server(streams=2, max_batch=8, mem_budget_bytes=1 << 20, p=S.params(),
                points_p=S.params(dualize=1, ss_block_size=1 << 30)) as srv:
        tickets = [srv.submit_points(X, r, 2) for X, r in reqs]
        srv.wait_all()
        for (X, r), t in zip(reqs, tickets):
            m = S.vr_build(X, r, 2)
            pairs = srv.poll(t)[0]
            assert np.array_equal(pairs, O.twist_pairs(O.Matrix(m.colptr, m.rows, m.dims), dual=True)), len(X)


def test_server_registered_matrices():
    """Matrices registered once and kept in device memory, submitted by id among host matrices
    and point clouds: every result equals the oracle's; a malformed matrix is refused at
    registration, an unknown or unregistered id at submission, and the ids still registered
    stay usable after another is unregistered."""
    rng = np.random.default_rng(12)
    mats = [S.vr_build(rng.random((40 + 9 * i, 3)).astype(np.float32), 0.3, 2) for i in range(6)]
    refs = [O.twist_pairs(O.Matrix(m.colptr, m.rows, m.dims)) for m in mats]
    clouds = [rng.random((30 + i, 3)).astype(np.float32) for i in range(6)]
    with Server(streams=2, max_batch=6, p=S.params(), points_p=S.params(dualize=1, ss_block_size=1 << 30)) as srv:
        bad = S.Matrix(mats[0].colptr.copy(), mats[0].rows.copy(), mats[0].dims.copy())
        bad.rows[-1] = bad.n + 3
        with pytest.raises(L.SSQError) as e:
            srv.register_matrix(bad)
        assert e.value.status == L.SSQ_EFORMAT
        ids = [srv.register_matrix(m) for m in mats]
        reqs = []
        for k in range(60):
            j = int(rng.integers(0, len(mats)))
            if k % 3 == 0:
                reqs.append(("reg", j, srv.submit_registered(ids[j])))
            elif k % 3 == 1:
                reqs.append(("mat", j, srv.submit_matrix(mats[j])))
            else:
                reqs.append(("pts", j, srv.submit_points(clouds[j], 0.3, 2)))
        srv.wait_all()
        for kind, j, t in reqs:
            pairs = srv.poll(t)[0]
            if kind == "pts":
                m = S.vr_build(clouds[j], 0.3, 2)
                ref = O.twist_pairs(O.Matrix(m.colptr, m.rows, m.dims), dual=True)
            else:
                ref = refs[j]
            assert np.array_equal(pairs, ref), (kind, j)
        with pytest.raises(L.SSQError):
            srv.submit_registered(10 ** 6)
        t = srv.submit_registered(ids[1])
        srv.unregister_matrix(ids[1])  # waits for that request
        assert np.array_equal(srv.poll(t)[0], refs[1])
        with pytest.raises(L.SSQError):
            srv.submit_registered(ids[1])
        with pytest.raises(L.SSQError):
            srv.unregister_matrix(ids[1])
        t = srv.submit_registered(ids[2])
        srv.wait_all()
        assert np.array_equal(srv.poll(t)[0], refs[2])
