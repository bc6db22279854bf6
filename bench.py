"""Benchmark: batched persistent homology of many Vietoris-Rips complexes on one MI355X.

Headline workload (the batched stream as specified): 2048 independent point clouds, point
counts log-uniform in [64, 2048], clouds 1/3 each uniform cube, noisy circle (sigma 0.05) and
noisy torus (R=1, r=0.3, sigma 0.03) in R^3, each with its own threshold at the 30th
percentile of its pairwise distances, simplices up to dimension 2, float32 filtration, int32
indices: 1.9e8 edges and 1.9e10 triangles per step.  The points are resident in HBM before the
timed region; one step turns every cloud's points into its persistence pairs in HBM (sizing,
Rips coboundary matrices built straight into the engine's layout, cohomology reduction, pairs).

Also measured: bit-exact parity of a sample against the CPU oracle (and structural checks of
every cloud's pairs), the oracle's own rate on that sample (cpu_baseline), the plain
per-complex GPU reduction (one engine call per cloud), a served stream (20,000 Poisson
requests per offered load, p50/p99 latency, every result checked against the oracle), the
large single-complex configuration C2 (model 3), and the roofline of the dominant kernel.

  python bench.py [--steps K] [--warmup W] [--complexes C] [--nmax 2048] [--no-cpu] [--served N]
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

os.environ.setdefault("SSQ_KERNEL_TIMING", "1")
# sixteen hardware queues per process (HIP's default is 4; SSQ_BENCH_HW_QUEUES changes it), read
# when HIP starts: the batched Rips path runs one stream per queue, and its groups' reductions
# overlap on them
os.environ["GPU_MAX_HW_QUEUES"] = os.environ.get("SSQ_BENCH_HW_QUEUES", "16")
ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from specseqplusplus_amd import workloads as W  # noqa: E402

METRIC = "pairs bit-exact; diagrams/s + p50/p99: batched server vs plain per-complex GPU"
HBM_PEAK_GBS = 8000.0
BATCH_PARAMS = dict(dualize=1, model=0, ss_block_size=1 << 30, growth=1.05, queue_size=65536, level1=64, level2=1024)
# Device bytes per group of clouds in the batched leg (SSQ_RIPS_GROUP_BYTES, unless set): sixteen
# groups in flight, one per stream, fit the GPU's memory with the streams' device regions.
# Fixed, so that the one-stream roofline step and its rocprofv3 runs reduce the same groups.
GROUP_BYTES = 9_000_000_000
SERVED_PARAMS = dict(dualize=1, model=0, ss_block_size=1 << 30, growth=1.05, queue_size=16384, level1=64, level2=1024)
SERVED_MATRIX_PARAMS = dict(dualize=0, model=0, ss_block_size=1 << 30, growth=1.05, queue_size=65536, level1=1000,
                            level2=16000)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------------------
# roofline
# ---------------------------------------------------------------------------------------
# Column-addition kernels: (name, time field, algorithmic-bytes field).  The engine sums the
# algorithmic bytes per kernel.  A merge a <- a xor b of sorted lists (k_add, and the tiled
# merges) costs 4*(|a| + |b| + |a xor b|) + 24 bytes: both columns read, the result written, the
# pivot lookup and column metadata.  An addition into a bit set kept on chip (k_add_long, the
# window of k_add_huge) costs 4*|b| + 24: only the pivot's rows are read.
ADD_KERNELS = [("k_add", "t_add_kernel_ms", "add_bytes"),
               ("k_add_long", "t_add_long_kernel_ms", "dense_bytes"),
               ("k_add_huge", "t_add_huge_kernel_ms", "huge_bytes")]
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_c4.txt")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel`, from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (lines 'traffic <kernel> <bytes per launch>'); None when absent."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                t = line.split()
                if len(t) == 3 and t[0] == "traffic" and t[1] == kernel:
                    return float(t[2])
    except OSError:
        pass
    return None


def roofline_record(st):
    """Roofline of the column-addition kernel with the largest measured time in one step on one
    stream: its algorithmic bytes over its summed launch durations (HIP events on the stream it
    is launched on)."""
    per = [(st[tf], st[bf], name) for name, tf, bf in ADD_KERNELS]
    ms, nbytes, name = max(per)
    launches = max(1, int(st["n_add_launches"]))  # each add kernel launches once per round
    achieved = (nbytes / 1e9) / (ms / 1e3) if ms > 0 else 0.0
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": measured_traffic(name),
            "kernel": name, "launches": launches, "avg_launch_us": round(ms * 1e3 / launches, 2),
            "bytes_per_launch": round(nbytes / launches, 1),
            "kernels_ms_per_step": {n: round(t, 3) for t, _, n in per}}


# ---------------------------------------------------------------------------------------
# legs
# ---------------------------------------------------------------------------------------
def rips_streams():
    """Streams of the batched Rips path, by libssq's rule: SSQ_RIPS_STREAMS, else one per
    hardware queue (GPU_MAX_HW_QUEUES, HIP's default 4), at most 16."""
    n = int(os.environ.get("SSQ_RIPS_STREAMS") or os.environ.get("GPU_MAX_HW_QUEUES") or 4)
    return max(1, min(16, n))


def batched_leg(h, pts, rs, args):
    """The headline: every cloud's pairs, points resident in HBM, one stream per hardware
    queue."""
    import specseqplusplus_amd as S
    pbase = np.concatenate([[0], np.cumsum([len(X) for X in pts])]).astype(np.int64)
    d_pts = h.device_array(np.concatenate(pts).astype(np.float32))
    rad = np.array(rs, np.float32)
    edges, tris = h.rips_count(d_pts, pbase, rad)
    cap = int(pbase[-1] + edges.sum())
    d_pairs = h.device_empty(2 * cap, np.int64)
    p = S.params(**BATCH_PARAMS)
    os.environ.setdefault("SSQ_RIPS_GROUP_BYTES", str(GROUP_BYTES))
    group_bytes = int(os.environ["SSQ_RIPS_GROUP_BYTES"])
    log(f"workload: {len(pts)} clouds, {pbase[-1]} points, {edges.sum()} edges, {tris.sum()} triangles")
    for i in range(args.warmup):
        t = time.perf_counter()
        h.batch_reduce_points_device(d_pts, pbase, rad, d_pairs, cap, p)
        log(f"warmup {i}: {time.perf_counter() - t:.2f} s")
    h.synchronize()
    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        off, st = h.batch_reduce_points_device(d_pts, pbase, rad, d_pairs, cap, p, stats=True)
        stats.append(st)
    h.synchronize()
    dt = time.perf_counter() - t0
    log(f"batched: {dt * 1e3 / args.steps:.1f} ms/step, {len(pts) * args.steps / dt:.1f} diagrams/s")
    allpairs = d_pairs.numpy(2 * int(off[-1])).reshape(-1, 2)
    per = [allpairs[off[i]:off[i + 1]] for i in range(len(pts))]
    # one more step on one stream for the roofline: there, kernel time per step is at most the
    # step's time
    os.environ["SSQ_RIPS_STREAMS"] = "1"
    t = time.perf_counter()
    _, st1 = h.batch_reduce_points_device(d_pts, pbase, rad, d_pairs, cap, p, stats=True)
    one_stream_ms = (time.perf_counter() - t) * 1e3
    del os.environ["SSQ_RIPS_STREAMS"]
    del os.environ["SSQ_RIPS_GROUP_BYTES"]
    for a in (d_pts, d_pairs):
        a.free()
    return {"ms_per_step": dt * 1e3 / args.steps, "diagrams_per_s": len(pts) * args.steps / dt, "stats": stats,
            "per": per, "edges": edges, "tris": tris, "pbase": pbase, "one_stream_ms": one_stream_ms,
            "one_stream_stats": st1, "group_bytes": group_bytes}


def structural_checks(pts, per, edges, tris):
    """Every cloud's pairs: births strictly increasing (so sorted), each cell in at most one
    pair, birth < death < cells, one pair per vertex birth except one per connected component
    (H0: n - #components pairs), and at most one pair per vertex and edge."""
    from scipy.sparse.csgraph import connected_components
    from scipy.spatial.distance import pdist, squareform

    def one(i):
        X, pr = pts[i], per[i]
        n = len(X)
        cells = n + int(edges[i]) + int(tris[i])
        ok = len(pr) <= n + int(edges[i])
        if len(pr):
            b, d = pr[:, 0], pr[:, 1]
            ok &= bool((np.diff(b) > 0).all() and (b < d).all() and d.max() < cells and b.min() >= 0)
            ok &= len(np.unique(pr)) == 2 * len(pr)
        return ok, n, int((pr[:, 0] < n).sum()) if len(pr) else 0

    res = list(ThreadPoolExecutor(16).map(one, range(len(pts))))
    ok = all(r[0] for r in res)
    return ok, res


def components(X, r):
    from scipy.sparse.csgraph import connected_components
    from scipy.spatial.distance import pdist, squareform
    if len(X) < 2:
        return len(X)
    D = squareform(pdist(X.astype(np.float64))).astype(np.float32)  # float32 distances, as built
    A = D <= np.float32(r)
    return connected_components(A, directed=False)[0]


def parity_leg(h, pts, rs, per, args):
    """Oracle (cohomology twist reduction of the full Rips matrix) on a sample: the largest
    cloud and random ones until the time budget is spent, one cloud per host thread.  The
    matrices are built by the general GPU Rips builder (itself checked cell for cell against
    the host builder in the tests); the oracle's time is the CPU baseline."""
    from concurrent.futures import FIRST_COMPLETED, wait
    from oracle import oracle as O
    O.build()
    workers = max(1, min(16, os.cpu_count() or 1))
    order = [int(np.argmax([len(X) for X in pts]))] + \
        [int(i) for i in np.random.default_rng(1).permutation(len(pts)) if i != np.argmax([len(X) for X in pts])]
    it = iter(order)
    done, busy, exact, cells, oracle_bytes = 0, 0.0, True, 0, 0.0
    oracle_adds, oracle_chain, oracle_depth, sample = 0.0, 0.0, 0.0, []
    t_start = time.perf_counter()

    def work(i, m):
        mo = O.Matrix(m.colptr, m.rows, m.dims)
        tt = time.perf_counter()
        ref, ost = O.twist_pairs(mo, dual=True, stats=True)
        return i, ref, time.perf_counter() - tt, m.n, ost

    with ThreadPoolExecutor(workers) as ex:
        futs = set()
        while True:
            while len(futs) < workers and time.perf_counter() - t_start < args.cpu_seconds:
                i = next(it, None)
                if i is None:
                    break
                futs.add(ex.submit(work, i, h.vr_build_gpu(pts[i], rs[i], 2)))
            if not futs:
                break
            fin, futs = wait(futs, return_when=FIRST_COMPLETED)
            for f in fin:
                i, ref, dt, n, ost = f.result()
                busy += dt
                done += 1
                cells += n
                oracle_bytes += ost["bytes"]
                oracle_adds += ost["additions"]
                oracle_chain = max(oracle_chain, ost["max_chain"])
                oracle_depth = max(oracle_depth, ost["depth"])
                sample.append(i)
                exact &= bool(ref.shape == per[i].shape and np.array_equal(ref, per[i]))
    wall = time.perf_counter() - t_start
    log(f"parity: {done} clouds ({cells} cells) bit-exact={exact}; oracle {done / busy:.2f}/s per core")
    # the engine's additions on the same clouds, one batched call: no schedule of the engine
    # may add more columns than the sequential reduction does (the GPU cannot inflate B_alg)
    import specseqplusplus_amd as S
    sp = [pts[i] for i in sample]
    pb = np.concatenate([[0], np.cumsum([len(X) for X in sp])]).astype(np.int64)
    d_sp = h.device_array(np.concatenate(sp).astype(np.float32))
    srad = np.array([rs[i] for i in sample], np.float32)
    se, _ = h.rips_count(d_sp, pb, srad)
    scap = int(pb[-1] + se.sum())
    d_sp_pairs = h.device_empty(2 * scap, np.int64)
    _, est = h.batch_reduce_points_device(d_sp, pb, srad, d_sp_pairs, scap, S.params(**BATCH_PARAMS), stats=True)
    for a in (d_sp, d_sp_pairs):
        a.free()
    adds = {"oracle_additions": oracle_adds, "engine_additions": float(est["n_additions"]),
            "engine_over_oracle": round(float(est["n_additions"]) / max(oracle_adds, 1.0), 4),
            "oracle_max_chain": oracle_chain, "engine_max_chain": float(est["max_chain"]),
            "oracle_dependency_depth": oracle_depth,
            "oracle_bytes": oracle_bytes}
    log(f"additions on the sample: engine {adds['engine_additions']:.0f}, oracle {oracle_adds:.0f} "
        f"(ratio {adds['engine_over_oracle']}); longest chain engine {adds['engine_max_chain']:.0f}, "
        f"oracle {oracle_chain:.0f}")
    cpu = {"value": round(done / wall, 3), "unit": "diagrams/s", "cores": workers, "kind": "port",
           "single_core_value": round(done / busy, 3),
           "sample": f"{done} clouds of this stream (the largest, then random ones), {cells} cells; oracle twist "
                     f"reduction in cohomology (anti-transpose, bit-tree pivot column, clearing), one cloud per "
                     f"thread on {workers} threads, {wall:.1f} s wall (matrix builds not timed)"}
    return {"checked": done, "cells": cells, "bit_exact": exact, "additions_on_sample": adds}, cpu


def plain_leg(pts, rs, args):
    """The plain per-complex GPU reduction: one engine call per cloud (the general path: GPU Rips
    build, anti-transpose, single-complex reduction), on a random sample; and the batched path
    called with one cloud at a time."""
    import specseqplusplus_amd as S
    idx = np.random.default_rng(2).permutation(len(pts))[:args.plain]
    p = S.params(**BATCH_PARAMS)
    out = {}
    for name, env in (("plain_gpu", "1"), ("one_cloud_per_call", None)):
        if env:
            os.environ["SSQ_RIPS_GENERAL"] = env
        h = S.Handle(0)
        big = int(idx[np.argmax([len(pts[i]) for i in idx])])
        h.batch_reduce_points([(pts[big], rs[big])], 2, p)  # size the buffers
        t = time.perf_counter()
        for i in idx:
            h.batch_reduce_points([(pts[i], rs[i])], 2, p)
        out[name] = len(idx) / (time.perf_counter() - t)
        h.close()
        os.environ.pop("SSQ_RIPS_GENERAL", None)
        log(f"{name}: {out[name]:.2f} diagrams/s on {len(idx)} random clouds")
    return out, len(idx)


def served_pool(seed, n_clouds, n_mats):
    """Requests of the served stream (workloads.served_requests: clouds of 64-4096 points, large
    ones at a mean degree of at most 20; alpha matrices of 1e4-1e6 cells) and their reference
    pairs from the oracle."""
    import specseqplusplus_amd as S
    from oracle import oracle as O
    pts, mats = W.served_requests(seed, n_clouds, n_mats)

    def ref_pts(c):
        m = S.vr_build(c[0], c[1], 2)
        return O.twist_pairs(O.Matrix(m.colptr, m.rows, m.dims), dual=True)

    def ref_mat(m):
        return O.twist_pairs(O.Matrix(m.colptr, m.rows, m.dims))

    with ThreadPoolExecutor(16) as ex:
        rp = list(ex.map(ref_pts, pts))
        rm = list(ex.map(ref_mat, mats))
    return pts, mats, (rp, rm)


def served_leg(args):
    import specseqplusplus_amd as S
    from specseqplusplus_amd.server import Server
    t = time.perf_counter()
    pool_pts, pool_mats, pool_ref = served_pool(args.seed + 7, args.pool_clouds, args.pool_mats)
    log(f"served pool: {len(pool_pts)} clouds, {len(pool_mats)} matrices "
        f"({min(m.n for m in pool_mats)}-{max(m.n for m in pool_mats)} cells), oracle references in "
        f"{time.perf_counter() - t:.1f} s")
    served = {"requests_per_load": args.served, "pool": {"clouds": len(pool_pts), "matrices": len(pool_mats)},
              "reference": "oracle", "loads": []}
    # plain per-request GPU: the same mix one request at a time (host data in, pairs out)
    h = S.Handle(0)
    rng = np.random.default_rng(99)
    n_plain = max(1, min(args.served, 400))
    kinds = rng.integers(0, 2, n_plain)
    picks = rng.integers(0, 1 << 30, n_plain)
    pp_pts, pp_mat = S.params(**SERVED_PARAMS), S.params(**SERVED_MATRIX_PARAMS)
    h.reduce(max(pool_mats, key=lambda m: m.n), pp_mat)
    Xb, rb = max(pool_pts, key=lambda c: len(c[0]))
    h.batch_reduce_points([(Xb, rb)], 2, pp_pts)
    t = time.perf_counter()
    for k in range(n_plain):
        if kinds[k] == 0:
            X, r = pool_pts[picks[k] % len(pool_pts)]
            h.batch_reduce_points([(X, r)], 2, pp_pts)
        else:
            h.reduce(pool_mats[picks[k] % len(pool_mats)], pp_mat)
    served["plain_per_request_diagrams_per_s"] = round(n_plain / (time.perf_counter() - t), 1)
    # the same, with the matrices already in device memory (one resident matrix per call)
    dms = [h.to_device(m) for m in pool_mats]
    t = time.perf_counter()
    for k in range(n_plain):
        if kinds[k] == 0:
            X, r = pool_pts[picks[k] % len(pool_pts)]
            h.batch_reduce_points([(X, r)], 2, pp_pts)
        else:
            h.batch_reduce_resident([dms[picks[k] % len(pool_mats)]], pp_mat)
    served["plain_per_request_resident_diagrams_per_s"] = round(n_plain / (time.perf_counter() - t), 1)
    for dm in dms:
        h.free_device_matrix(dm)
    h.close()
    log(f"served mix, plain one request at a time: {served['plain_per_request_diagrams_per_s']} diagrams/s "
        f"(matrices from the host), {served['plain_per_request_resident_diagrams_per_s']} (matrices resident)")
    served["matrices"] = "registered (resident in HBM)" if args.served_matrices == "registered" else "host"
    for qps in args.qps:
        rng = np.random.default_rng(int(qps))
        kinds = rng.integers(0, 2, args.served)
        picks = rng.integers(0, 1 << 30, args.served)
        gaps = rng.exponential(1.0 / qps, args.served)
        lat = []
        with Server(streams=args.server_streams, max_batch=args.server_batch, max_batch_cells=args.server_cells,
                    p=S.params(**SERVED_MATRIX_PARAMS), points_p=S.params(**SERVED_PARAMS)) as srv:
            mids = [srv.register_matrix(m) for m in pool_mats] if args.served_matrices == "registered" else None
            tickets = []
            ts = time.perf_counter()
            nxt = ts
            for k in range(args.served):
                nxt += gaps[k]
                while True:
                    now = time.perf_counter()
                    if now >= nxt:
                        break
                    time.sleep(min(nxt - now, 0.0005))
                j = int(picks[k] % (len(pool_pts) if kinds[k] == 0 else len(pool_mats)))
                if kinds[k] == 0:
                    X, r = pool_pts[j]
                    tickets.append((0, j, srv.submit_points(X, r, 2)))
                else:
                    tickets.append((1, j, srv.submit_registered(mids[j]) if mids else srv.submit_matrix(pool_mats[j])))
            srv.wait_all()
            te = time.perf_counter()
            sstats = srv.stats()
            exact = True
            kind_of = []
            for kind, j, tk in tickets:
                res = srv.poll(tk)
                lat.append(res[2])
                kind_of.append(kind)
                exact &= bool(np.array_equal(res[0], pool_ref[kind][j]))
        lat = np.array(lat)
        kind_of = np.array(kind_of)
        rate = args.served / (te - ts)
        rec = {"qps_offered": qps, "diagrams_per_s": round(rate, 1),
               "over_plain": round(rate / served["plain_per_request_diagrams_per_s"], 2),
               "over_plain_resident": round(rate / served["plain_per_request_resident_diagrams_per_s"], 2),
               "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
               "p99_ms_clouds": round(float(np.percentile(lat[kind_of == 0], 99)), 3) if (kind_of == 0).any() else None,
               "p99_ms_matrices": round(float(np.percentile(lat[kind_of == 1], 99)), 3) if (kind_of == 1).any() else None,
               "bit_exact_vs_oracle": exact, "batches": sstats}
        served["loads"].append(rec)
        log(f"served {rec}")
    return served


def single_config(args, which=None, quiet=False):
    """One single-complex configuration: the GPU reduction with the matrix resident in HBM, the
    oracle on one host core, and B_alg / t_gpu, where B_alg = the oracle's addition bytes (same
    sub-complex) + 4*nnz + 12*N for the input pass + 16*#pairs for the output."""
    import specseqplusplus_amd as S
    from specseqplusplus_amd import configs as CF
    from oracle import oracle as O
    O.build()
    which = which or args.single
    t0 = time.time()
    if which == "c1":
        m, kw, what = S.vr_build(CF.c1_points(), 0.25, 2), dict(CF.C1_PARAMS), "C1: Rips of 512 uniform points, r=0.25, dim<=2"
    elif which == "c2":
        m, kw = S.vr_build(CF.c2_points(), 1.0, 2), dict(CF.C2_PARAMS)
        what = "C2: Rips of 4096 Haar rotations in R^9, r=1.0, dim<=2, model 3"
    else:
        from specseqplusplus_amd.alpha import alpha_complex, c3_points
        m, kw = alpha_complex(c3_points(), 90.0), dict(model=0, queue_size=20000, level1=1000, level2=16000)
        what = "C3: alpha complex of 1e5 uniform points in a 100-cube, alpha^2<=90, dim<=3"
    if args.dualize and not quiet:
        kw["dualize"] = 1
        what += ", cohomology (dualize)"
    log(f"{what}: {m.n} cells, {m.nnz} nonzeros (built in {time.time() - t0:.1f}s)")
    h = S.Handle(0)
    dm = h.upload(m)
    out = h.device_empty(2 * m.n, np.int64)
    p = S.params(**kw)
    steps = args.single_steps if quiet else args.steps
    for _ in range(args.warmup):
        h.reduce_resident(dm, out, p)
    h.synchronize()
    sts = []
    t = time.perf_counter()
    for _ in range(steps):
        sts.append(h.reduce_resident(dm, out, p))
    h.synchronize()
    ms = (time.perf_counter() - t) * 1e3 / steps
    n_pairs, st = sts[-1]
    pairs = out.numpy(2 * n_pairs).reshape(-1, 2)
    mo = O.Matrix(m.colptr, m.rows, m.dims)
    runs = []
    for _ in range(5):  # the median of five runs
        tt = time.perf_counter()
        skip = O.collapse_skip(mo) if kw.get("model", 0) & 2 else None  # the CPU path collapses too
        ref, ost = O.twist_pairs(mo, skip=skip, stats=True)
        runs.append(time.perf_counter() - tt)
    cpu_s = float(np.median(runs))
    b_alg = ost["bytes"] + 4.0 * m.nnz + 12.0 * m.n + 16.0 * len(ref)
    achieved = b_alg / (ms / 1e3) / 1e9
    rec = {"metric": f"single-complex reduction time ({which})", "value": round(ms, 3), "unit": "ms",
           "higher_is_better": False, "n_gpus": 1, "steps": steps, "warmup": args.warmup,
           "dtype": "int32", "data": "synthetic (seeded)",
           "config": {"workload": what, "cells": int(m.n), "nnz": int(m.nnz), "params": kw},
           "bit_exact": bool(np.array_equal(pairs, ref)), "pairs": int(len(ref)),
           "engine": {k: st[k] for k in ("t_init_ms", "t_collapse_ms", "t_reduce_ms", "n_additions", "n_apparent",
                                         "n_cleared_hd", "n_kept_edges", "n_pages", "n_rounds", "max_chain")},
           "b_alg_roofline": {"b_alg_bytes": b_alg, "oracle_additions": ost["additions"], "achieved": round(achieved, 2),
                              "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6)},
           "cpu_baseline": {"value": round(cpu_s * 1e3, 1), "unit": "ms", "cores": 1, "kind": "port",
                            "sample": "the whole matrix, median of 5 runs; oracle twist reduction (bit-tree pivot "
                                      "column, clearing)"
                                      + (", after the oracle's sequential GUDHI edge collapse (timed)"
                                         if skip is not None else "")}}
    h.close()
    if quiet:
        return rec
    print(json.dumps(rec))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--complexes", type=int, default=2048)
    ap.add_argument("--nmin", type=int, default=64)
    ap.add_argument("--nmax", type=int, default=2048)
    ap.add_argument("--plain", type=int, default=48, help="random clouds timed one call each (plain GPU)")
    ap.add_argument("--served", type=int, default=20000, help="served requests per offered load (0 = skip)")
    ap.add_argument("--qps", type=float, nargs="+", default=[1000.0, 1800.0, 2000.0, 2500.0])
    ap.add_argument("--pool-clouds", type=int, default=192)
    ap.add_argument("--pool-mats", type=int, default=24)
    ap.add_argument("--server-streams", type=int, default=4)
    ap.add_argument("--server-cells", type=int, default=4_000_000)
    ap.add_argument("--server-batch", type=int, default=32, help="requests per served batch at most")
    ap.add_argument("--served-matrices", choices=["registered", "host"], default="registered",
                    help="matrix requests by the id of a matrix registered in HBM, or as host matrices")
    ap.add_argument("--cpu-seconds", type=float, default=45.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c2", action="store_true", help="skip the C2 sub-record")
    ap.add_argument("--single-steps", type=int, default=5)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--single", choices=["c1", "c2", "c3"],
                    help="time one single-complex configuration instead of the batched stream")
    ap.add_argument("--dualize", action="store_true", help="with --single: reduce in cohomology")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: every complex is independent, so more GPUs would each run "
              "their own copy of this stream; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.single:
        single_config(args)
        return

    import specseqplusplus_amd as S
    t = time.time()
    pts, rs = W.batched_stream(args.complexes, args.nmin, args.nmax, args.seed)
    log(f"generated {len(pts)} clouds in {time.time() - t:.1f} s")
    h = S.Handle(0)
    B = batched_leg(h, pts, rs, args)
    h.close()
    N = int(B["pbase"][-1] + B["edges"].sum() + B["tris"].sum())
    NNZ = int(2 * B["edges"].sum() + 3 * B["tris"].sum())
    npairs = int(sum(len(p) for p in B["per"]))

    ok_struct, res = structural_checks(pts, B["per"], B["edges"], B["tris"])
    with ThreadPoolExecutor(16) as ex:
        comps = list(ex.map(lambda i: components(pts[i], rs[i]), range(len(pts))))
    h0_ok = all(res[i][2] == len(pts[i]) - comps[i] for i in range(len(pts)))
    log(f"structural checks of all {len(pts)} clouds: {ok_struct}; H0 pair counts = n - components: {h0_ok}")

    parity, cpu = ({"checked": 0}, None)
    if not args.no_cpu:
        h = S.Handle(0)
        parity, cpu = parity_leg(h, pts, rs, B["per"], args)
        h.close()
    parity["all_clouds_structural"] = bool(ok_struct and h0_ok)

    plain, n_plain = ({}, 0)
    if args.plain > 0:
        plain, n_plain = plain_leg(pts, rs, args)
    served = served_leg(args) if args.served > 0 else None
    c2 = None if args.no_c2 else single_config(args, "c2", quiet=True)

    st1 = B["one_stream_stats"]
    roof = roofline_record(st1)
    roof["one_stream_step_ms"] = round(B["one_stream_ms"], 1)
    # whole-step algorithmic bytes that no schedule can inflate: one pass over the input rows and
    # cell records and the pairs written; the additions are left out (so this is a lower bound)
    b_step = 4.0 * NNZ + 12.0 * N + 16.0 * npairs
    step_gbs = b_step / (B["ms_per_step"] / 1e3) / 1e9
    last = B["stats"][-1]
    out = {
        "metric": METRIC, "value": round(B["diagrams_per_s"], 2), "unit": "diagrams/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(B["ms_per_step"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (seeded point clouds; every step builds the Vietoris-Rips coboundary matrices on the GPU)",
        "config": {"workload": f"batched: {args.complexes} Rips complexes, n log-uniform [{args.nmin},{args.nmax}], "
                               f"cube/circle/torus, r = 30th percentile, dim<=2; points in HBM -> pairs in HBM",
                   "points": int(B["pbase"][-1]), "edges": int(B["edges"].sum()), "triangles": int(B["tris"].sum()),
                   "cells": N, "nnz": NNZ, "pairs": npairs, "streams": rips_streams(),
                   "hw_queues": int(os.environ["GPU_MAX_HW_QUEUES"]), "group_bytes": B["group_bytes"],
                   "groups_per_step": int(last["n_pages"]),
                   "params": BATCH_PARAMS},
        "plain_gpu_diagrams_per_s": round(plain["plain_gpu"], 3) if plain else None,
        "batched_over_plain": round(B["diagrams_per_s"] / plain["plain_gpu"], 2) if plain else None,
        "one_cloud_per_call_diagrams_per_s": round(plain["one_cloud_per_call"], 3) if plain else None,
        "plain_sample": n_plain,
        "served": served, "parity": parity,
        "engine": {"pending": last["n_pending"], "apparent": last["n_apparent"], "additions": last["n_additions"],
                   "huge_additions": last["n_huge_additions"], "rounds": last["n_rounds"],
                   "max_chain": last["max_chain"], "max_addition_len": last["max_addition_len"],
                   "reduce_ms_summed_over_streams": round(last["t_reduce_ms"], 1)},
        "roofline": roof,
        "step_roofline": {"b_alg_bytes": b_step, "achieved": round(step_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(step_gbs / HBM_PEAK_GBS, 6),
                          "note": "4*nnz + 12*cells + 16*pairs per step over the step time (additions excluded)"},
        "cpu_baseline": cpu,
        "c2": c2,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
