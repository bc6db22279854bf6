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
