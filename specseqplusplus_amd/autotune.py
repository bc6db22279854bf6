"""Parameter autotuner for batched reductions, one entry per size class.

Every parameter set gives the same pairs (the pairing of a filtered complex does not depend
on the reduction schedule), so the tuner only times candidates, in two stages: first the
reduction (homology with guided clearing or cohomology, page width, additions in flight, slot
capacities), then, on the winner, the additions' workgroup shape (wavefronts per short-addition
workgroup, the long-column grid, the LDS merge limit, rounds between worklist reads).  The
reference picks these per dataset by hand (README.md:79-101).  Results are cached per size
class (log2 of the mean cells per complex) in a JSON file.
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


# Workgroup shapes tried on the stage-one winner ({} = the engine's defaults: 4 wavefronts per
# workgroup, 2 long-column workgroups per CU, 1024-entry LDS merges, 3 rounds per read).
SHAPES = [
    dict(),
    dict(add_waves=2),
    dict(add_waves=8),
    dict(short_max=256),
    dict(short_max=512),
    dict(long_grid=128),
    dict(long_grid=1024),
    dict(rounds_per_sync=1),
    dict(rounds_per_sync=6),
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


def _time(handle, dm, out, c, repeats):
    p = params(**c)
    handle.reduce_resident(dm, out, p)  # warm
    best = float("inf")
    for _ in range(repeats):
        handle.synchronize()
        t = time.perf_counter()
        handle.reduce_resident(dm, out, p)
        handle.synchronize()
        best = min(best, time.perf_counter() - t)
    return best


def tune(handle, mats, candidates=None, shapes=None, repeats=2, cache_path=None):
    """Return (params dict, timings) for this batch's size class: the fastest candidate, then
    the fastest workgroup shape on it."""
    key = str(size_class(mats))
    if cache_path and os.path.exists(cache_path):
        cache = json.load(open(cache_path))
        if key in cache:
            return cache[key]["params"], cache[key]["timings"]
    candidates = candidates or CANDIDATES
    shapes = SHAPES if shapes is None else shapes
    bd, _ = block_diagonal(mats)
    dm = handle.upload(bd)
    out = handle.device_empty(2 * max(bd.n, 1), np.int64)
    timings = sorted(((_time(handle, dm, out, c, repeats), c) for c in candidates), key=lambda x: x[0])
    base = timings[0][1]
    shaped = sorted(((_time(handle, dm, out, dict(base, **sh), repeats), dict(base, **sh)) for sh in shapes),
                    key=lambda x: x[0])
    timings = sorted(timings + shaped, key=lambda x: x[0])
    chosen = timings[0][1]
    table = [{"ms": round(t * 1e3, 3), "params": c} for t, c in timings]
    if cache_path:
        cache = json.load(open(cache_path)) if os.path.exists(cache_path) else {}
        cache[key] = {"params": chosen, "timings": table}
        json.dump(cache, open(cache_path, "w"), indent=1)
    return chosen, table
