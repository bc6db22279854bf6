"""Parameter autotuner for batched reductions, one entry per size class.

Every parameter set gives the same pairs (the pairing of a filtered complex does not depend
on the reduction schedule), so the tuner only times candidates: homology with guided
clearing or cohomology (dualize), page width, additions in flight and slot capacities.
Results are cached per size class (log2 of the mean cells per complex) in a JSON file.
"""
import json
import math
import os
import time

import numpy as np

from .api import Matrix, params

CANDIDATES = [
    dict(dualize=1, model=0, ss_block_size=1 << 30, growth=1.05, queue_size=65536, level1=64, level2=1024),
    dict(dualize=1, model=0, ss_block_size=1 << 30, growth=1.05, queue_size=16384, level1=64, level2=1024),
    dict(dualize=1, model=0, ss_block_size=1 << 18, growth=2.0, queue_size=65536, level1=64, level2=1024),
    dict(dualize=0, model=1, ss_block_size=1 << 30, growth=1.05, queue_size=65536, level1=64, level2=1024),
    dict(dualize=0, model=0, ss_block_size=102400, growth=1.05, queue_size=65536, level1=1000, level2=2000),
]


def block_diagonal(mats):
    base = np.cumsum([0] + [m.n for m in mats])
    nzb = np.cumsum([0] + [m.nnz for m in mats])
    colptr = np.empty(base[-1] + 1, np.int64)
    rows = np.empty(nzb[-1], np.int32)
    dims = np.empty(base[-1], np.int32)
    for i, m in enumerate(mats):
        colptr[base[i]:base[i + 1] + 1] = m.colptr + nzb[i]
        rows[nzb[i]:nzb[i + 1]] = m.rows + base[i]
        dims[base[i]:base[i + 1]] = m.dims
    return Matrix(colptr, rows, dims), base


def size_class(mats):
    mean = max(1.0, float(np.mean([m.n for m in mats])))
    return int(math.log2(mean))


def tune(handle, mats, candidates=None, repeats=2, cache_path=None):
    """Return (params dict, timings) for this batch's size class."""
    key = str(size_class(mats))
    if cache_path and os.path.exists(cache_path):
        cache = json.load(open(cache_path))
        if key in cache:
            return cache[key]["params"], cache[key]["timings"]
    candidates = candidates or CANDIDATES
    bd, _ = block_diagonal(mats)
    dm = handle.upload(bd)
    out = handle.device_empty(2 * max(bd.n, 1), np.int64)
    timings = []
    for c in candidates:
        p = params(**c)
        handle.reduce_resident(dm, out, p)  # warm
        best = float("inf")
        for _ in range(repeats):
            handle.synchronize()
            t = time.perf_counter()
            handle.reduce_resident(dm, out, p)
            handle.synchronize()
            best = min(best, time.perf_counter() - t)
        timings.append((best, c))
    timings.sort(key=lambda x: x[0])
    chosen = timings[0][1]
    table = [{"ms": round(t * 1e3, 3), "params": c} for t, c in timings]
    if cache_path:
        cache = json.load(open(cache_path)) if os.path.exists(cache_path) else {}
        cache[key] = {"params": chosen, "timings": table}
        json.dump(cache, open(cache_path, "w"), indent=1)
    return chosen, table
