"""Python handle on the libssq continuous-batching server (ssq_server_*)."""
import ctypes

import numpy as np

from . import _lib as L
from .api import params as make_params


class Server:
    def __init__(self, device=0, streams=2, max_batch=64, mem_budget_bytes=0, p=None, max_batch_cells=0,
                 points_p=None):
        """p: parameters for submitted matrices; points_p: for point clouds (Rips complexes,
        cohomology by default)."""
        cfg = L.ssq_server_config()
        cfg.device, cfg.streams, cfg.max_batch, cfg.mem_budget_bytes = device, streams, max_batch, mem_budget_bytes
        cfg.max_batch_cells = max_batch_cells
        cfg.params = p or make_params()
        cfg.points_params = points_p or make_params(dualize=1, ss_block_size=1 << 30, queue_size=16384, level1=64,
                                                     level2=1024)
        self._s = ctypes.c_void_p()
        L.check(L.lib().ssq_server_create(ctypes.byref(cfg), ctypes.byref(self._s)), "ssq_server_create")

    def submit_matrix(self, m):
        t = ctypes.c_int64()
        ms = m.c_struct()
        L.check(L.lib().ssq_server_submit_matrix(self._s, ctypes.byref(ms), ctypes.byref(t)))
        return t.value

    def register_matrix(self, m):
        """Keep m resident in device memory for repeated requests; returns its id."""
        mid = ctypes.c_int64()
        ms = m.c_struct()
        L.check(L.lib().ssq_server_register_matrix(self._s, ctypes.byref(ms), ctypes.byref(mid)), "register")
        return mid.value

    def submit_registered(self, matrix_id):
        t = ctypes.c_int64()
        L.check(L.lib().ssq_server_submit_registered(self._s, int(matrix_id), ctypes.byref(t)))
        return t.value

    def unregister_matrix(self, matrix_id):
        """Waits for the matrix's outstanding requests, then frees its device copy."""
        L.check(L.lib().ssq_server_unregister_matrix(self._s, int(matrix_id)), "unregister")

    def submit_points(self, pts, r, maxdim=2):
        pts = np.ascontiguousarray(pts, dtype=np.float32)
        t = ctypes.c_int64()
        L.check(L.lib().ssq_server_submit_points(self._s, pts.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                 pts.shape[0], pts.shape[1], float(r), int(maxdim), ctypes.byref(t)))
        return t.value

    def poll(self, ticket):
        """(pairs, essential, latency_ms) or None while pending."""
        pr = L.ssq_pairs()
        lat = ctypes.c_double()
        rc = L.lib().ssq_server_poll(self._s, ticket, ctypes.byref(pr), ctypes.byref(lat))
        if rc == L.SSQ_EBUSY:
            return None
        L.check(rc, f"request {ticket}")
        n = pr.np
        pairs = np.ctypeslib.as_array(pr.pairs, shape=(max(2 * n, 1),))[:2 * n].copy().reshape(-1, 2)
        ess = np.ctypeslib.as_array(pr.essential, shape=(max(pr.ne, 1),))[:pr.ne].copy()
        L.lib().ssq_pairs_free(ctypes.byref(pr))
        return pairs, ess, lat.value

    def stats(self):
        """Per batch kind (host matrices, point clouds, registered matrices): batches, requests,
        cells, mean and maximum batch service time in ms."""
        st = L.ssq_server_stats()
        L.check(L.lib().ssq_server_get_stats(self._s, ctypes.byref(st)), "stats")
        out = {}
        for k, name in enumerate(("matrices", "clouds", "registered")):
            b = st.batches[k]
            out[name] = {"batches": b, "requests": st.requests[k], "cells": st.cells[k],
                         "mean_service_ms": round(st.service_ms[k] / b, 3) if b else None,
                         "max_service_ms": round(st.max_service_ms[k], 3)}
        return out

    def wait_all(self):
        L.check(L.lib().ssq_server_wait_all(self._s), "wait_all")

    def close(self):
        if self._s:
            L.lib().ssq_server_destroy(self._s)
            self._s = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
