"""Benchmark: batched persistent-homology reduction on one MI355X.

Workload (the metric's batched stream, sizes scaled to fit a bench run): many independent
Vietoris-Rips complexes, point counts log-uniform in [nmin, nmax], clouds 1/3 each uniform
cube, noisy circle (sigma 0.05) and noisy torus (R=1, r=0.3, sigma 0.03) in R^3, each with
its own threshold at the 30th percentile of its pairwise distances, dimension <= 2, float32
filtration, int32 indices.  The boundary matrices are resident in HBM (block-diagonal)
before the timed region; one step reduces all of them to persistence pairs in HBM.

Also measured: the plain per-complex GPU reduction (one engine call per complex, in
sequence), a served stream (Poisson arrivals, p50/p99 latency), bit-exact parity of a sample
against the CPU oracle, the CPU oracle's own rate on that sample (cpu_baseline), and the
column-addition kernel's achieved algorithmic bandwidth (roofline).

  python bench.py [--steps K] [--warmup W] [--complexes C] [--nmax 256] [--no-cpu]
"""
import argparse
import json
import os
import sys
import threading
import time
from concurrent.futures import ThreadPoolExecutor

os.environ.setdefault("SSQ_KERNEL_TIMING", "1")
ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "pairs bit-exact; diagrams/s + p50/p99: batched server vs plain per-complex GPU"
HBM_PEAK_GBS = 8000.0


def make_cloud(rng, kind, n):
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
    from scipy.spatial.distance import pdist
    return float(np.percentile(pdist(X.astype(np.float64)), q))


def build_stream(count, nmin, nmax, seed, q=30.0):
    import specseqplusplus_amd as S
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(nmin), np.log(nmax), count)).astype(int)
    specs = []
    for i in range(count):
        X = make_cloud(rng, i % 3, int(ns[i])).astype(np.float32)
        specs.append((X, percentile_threshold(X, q)))
    with ThreadPoolExecutor(8) as ex:
        mats = list(ex.map(lambda s: S.vr_build(s[0], s[1], 2), specs))
    return specs, mats


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
    return colptr, rows, dims, base


def split_pairs(pairs, base):
    out = []
    cut = np.searchsorted(pairs[:, 0], base)
    for i in range(len(base) - 1):
        out.append(pairs[cut[i]:cut[i + 1]] - base[i])
    return out


BATCH_PARAMS = dict(dualize=1, model=0, ss_block_size=1 << 30, growth=1.05, queue_size=65536, level1=64, level2=1024)
SERVED_PARAMS = dict(dualize=1, model=0, ss_block_size=1 << 30, growth=1.05, queue_size=16384, level1=64, level2=1024)
SERVED_MATRIX_PARAMS = dict(dualize=0, model=0, ss_block_size=1 << 30, growth=1.05, queue_size=65536, level1=1000,
                            level2=16000)


def build_served_pool(seed, n_clouds=48, n_mats=12):
    """Requests for the served stream and their reference pairs (computed alone on the GPU
    once here; the oracle checks a sample of them in the parity leg)."""
    import specseqplusplus_amd as S
    from specseqplusplus_amd.alpha import alpha_complex
    rng = np.random.default_rng(seed)
    pts = []
    for i in range(n_clouds):
        n = int(np.exp(rng.uniform(np.log(64), np.log(4096))))
        X = make_cloud(rng, i % 3, n).astype(np.float32)
        q = min(30.0, 100.0 * 20.0 / (n - 1))  # 30th percentile, capped at mean degree 20
        pts.append((X, percentile_threshold(X, q)))
    mats = []
    for i in range(n_mats):
        cols = np.exp(rng.uniform(np.log(1e4), np.log(1e6)))
        npts = max(64, int(cols / 28.5))
        P = rng.random((npts, 3)) * 100.0 * (npts / 1e5) ** (1 / 3)
        mats.append(alpha_complex(P, 90.0))
    h = S.Handle(0)
    p = S.params(**SERVED_PARAMS)
    ref_pts = [h.reduce(S.vr_build(X, r, 2), p)[0] for X, r in pts]
    ref_mats = [h.reduce(m, S.params(**SERVED_MATRIX_PARAMS))[0] for m in mats]
    h.close()
    return pts, mats, (ref_pts, ref_mats)


# Column-addition kernels: (name, time field, algorithmic-bytes field).  Algorithmic bytes of
# one addition a <- a xor b are 4*(|a| + |b| + |a xor b|) + 24 (read both columns, write the
# result, plus pivot lookups and column metadata); the engine sums them per kernel.
ADD_KERNELS = [("k_add", "t_add_kernel_ms", "add_bytes"),
               ("k_add_long", "t_add_long_kernel_ms", "dense_bytes")]
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_batched_rips.txt")


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


def roofline_record(stats):
    """Roofline of the column-addition kernel with the largest measured time: algorithmic bytes
    per launch over its average launch duration (HIP events on the engine's stream)."""
    launches = sum(s["n_add_launches"] for s in stats)  # k_add and k_add_long launch once per round each
    per = []
    for name, tf, bf in ADD_KERNELS:
        per.append((sum(s[tf] for s in stats), sum(s[bf] for s in stats), name))
    ms, nbytes, name = max(per)
    achieved = (nbytes / 1e9) / (ms / 1e3) if ms > 0 else 0.0
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": measured_traffic(name),
            "kernel": name, "launches": int(launches), "avg_launch_us": round(ms * 1e3 / max(launches, 1), 2),
            "bytes_per_launch": round(nbytes / max(launches, 1), 1),
            "kernels_ms": {n: round(t, 3) for t, _, n in per}}


def single_config(args):
    """One single-complex configuration: the GPU reduction with the matrix resident in HBM,
    the oracle on one host core, and B_alg / t_gpu, where B_alg = the oracle's addition bytes
    (same sub-complex) + 4*nnz + 12*N for the input pass + 16*#pairs for the output."""
    import specseqplusplus_amd as S
    from specseqplusplus_amd import configs as CF
    from oracle import oracle as O
    O.build()
    t0 = time.time()
    if args.single == "c1":
        m, kw, what = S.vr_build(CF.c1_points(), 0.25, 2), dict(CF.C1_PARAMS), "C1: Rips of 512 uniform points, r=0.25, dim<=2"
    elif args.single == "c2":
        m, kw = S.vr_build(CF.c2_points(), 1.0, 2), dict(CF.C2_PARAMS)
        what = "C2: Rips of 4096 Haar rotations in R^9, r=1.0, dim<=2, model 3"
    else:
        from specseqplusplus_amd.alpha import alpha_complex, c3_points
        m, kw = alpha_complex(c3_points(), 90.0), dict(model=0, queue_size=20000, level1=1000, level2=16000)
        what = "C3: alpha complex of 1e5 uniform points in a 100-cube, alpha^2<=90, dim<=3"
    if args.dualize:
        kw["dualize"] = 1
        what += ", cohomology (dualize)"
    print(f"[bench] {what}: {m.n} cells, {m.nnz} nonzeros (built in {time.time() - t0:.1f}s)",
          file=sys.stderr, flush=True)
    h = S.Handle(0)
    dm = h.upload(m)
    out = h.device_empty(2 * m.n, np.int64)
    p = S.params(**kw)
    for _ in range(args.warmup):
        h.reduce_resident(dm, out, p)
    h.synchronize()
    sts = []
    t = time.perf_counter()
    for _ in range(args.steps):
        sts.append(h.reduce_resident(dm, out, p))
    h.synchronize()
    ms = (time.perf_counter() - t) * 1e3 / args.steps
    n_pairs, st = sts[-1]
    pairs = out.numpy(2 * n_pairs).reshape(-1, 2)
    mo = O.Matrix(m.colptr, m.rows, m.dims)
    tt = time.perf_counter()
    skip = O.collapse_skip(mo) if kw.get("model", 0) & 2 else None  # the CPU path collapses too
    ref, ost = O.twist_pairs(mo, skip=skip, stats=True)
    cpu_s = time.perf_counter() - tt
    b_alg = ost["bytes"] + 4.0 * m.nnz + 12.0 * m.n + 16.0 * len(ref)
    achieved = b_alg / (ms / 1e3) / 1e9
    rec = {"metric": f"single-complex reduction time ({args.single})", "value": round(ms, 3), "unit": "ms",
           "higher_is_better": False, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "dtype": "int32", "data": "synthetic (seeded)",
           "config": {"workload": what, "cells": int(m.n), "nnz": int(m.nnz), "params": kw},
           "bit_exact": bool(np.array_equal(pairs, ref)), "pairs": int(len(ref)),
           "engine": {k: st[k] for k in ("t_init_ms", "t_collapse_ms", "t_reduce_ms", "n_additions", "n_apparent",
                                         "n_cleared_hd", "n_kept_edges", "n_pages", "n_rounds", "max_chain")},
           "b_alg_roofline": {"b_alg_bytes": b_alg, "oracle_additions": ost["additions"], "achieved": round(achieved, 2),
                              "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6)},
           "cpu_baseline": {"value": round(cpu_s * 1e3, 1), "unit": "ms", "cores": 1, "kind": "port",
                            "sample": "the whole matrix, oracle twist reduction (bit-tree pivot column, clearing)"
                                      + (", after the oracle's sequential GUDHI edge collapse (timed)"
                                         if skip is not None else "")}}
    print(json.dumps(rec))
    h.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--complexes", type=int, default=2048)
    ap.add_argument("--nmin", type=int, default=64)
    ap.add_argument("--nmax", type=int, default=256)
    ap.add_argument("--plain", type=int, default=512, help="complexes timed one call each (plain GPU)")
    ap.add_argument("--served", type=int, default=2000, help="served requests per offered load (0 = skip)")
    ap.add_argument("--qps", type=float, nargs="+", default=[200.0, 1000.0, 2000.0, 2500.0, 4000.0])
    ap.add_argument("--server-streams", type=int, default=4)
    ap.add_argument("--server-cells", type=int, default=2_000_000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--single", choices=["c1", "c2", "c3"],
                    help="time one single-complex configuration instead of the batched stream")
    ap.add_argument("--dualize", action="store_true", help="with --single: reduce in cohomology")
    ap.add_argument("--streams", type=int, default=3,
                    help="concurrent groups (one HIP stream each) the batched stream is split into")
    ap.add_argument("--no-tune", action="store_true",
                    help="use the default batched parameters instead of tuning each size class")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: every complex is independent, so more GPUs would each run "
              "their own copy of this stream; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.single:
        single_config(args)
        return

    import specseqplusplus_amd as S
    from specseqplusplus_amd.server import Server

    t0 = time.time()
    specs, mats = build_stream(args.complexes, args.nmin, args.nmax, args.seed)
    N, NNZ = sum(m.n for m in mats), sum(m.nnz for m in mats)
    print(f"[bench] workload: {args.complexes} complexes, {N} cells, {NNZ} nonzeros "
          f"(generated in {time.time() - t0:.1f}s)", file=sys.stderr, flush=True)

    # size classes (by point count), each tuned on a sample and reduced as one batch
    from specseqplusplus_amd.autotune import block_diagonal as bdiag, tune
    h = S.Handle(0)
    ns = np.array([len(sp[0]) for sp in specs])
    split = int(np.sqrt(args.nmin * args.nmax))
    classes = [np.where(ns < split)[0], np.where(ns >= split)[0]]
    classes = [c for c in classes if len(c)]
    # each size class is split into interleaved groups; every group is one block-diagonal
    # batch on its own handle (HIP stream), and a step reduces all groups concurrently
    total_cells = sum(m.n for m in mats)
    class_params = []
    for ci, idx in enumerate(classes):
        if args.no_tune:
            chosen, table = dict(BATCH_PARAMS), []
        else:
            chosen, table = tune(h, [mats[i] for i in idx[:min(len(idx), 96)]])
        class_params.append((chosen, table))
    # Each handle owns one HIP stream, and HIP puts a stream on one of GPU_MAX_HW_QUEUES hardware
    # queues (4 by default) when it is created: a stream created while four others are open
    # shares a queue and serialises with another busy stream, even after the others close
    # (4 groups: 32.6 ms/step created with 4 streams open, 46.6 ms with 5).  So every leg
    # closes its handles before the next one opens its own.
    h.close()
    plans = []
    for ci, idx in enumerate(classes):
        chosen, table = class_params[ci]
        cells = sum(mats[i].n for i in idx)
        parts = max(1, int(round(args.streams * cells / total_cells)))
        for g in range(parts):
            gidx = idx[g::parts]
            hg = S.Handle(0)
            bd, cbase = bdiag([mats[i] for i in gidx])
            plans.append({"idx": gidx, "h": hg, "dm": hg.upload(bd), "out": hg.device_empty(2 * bd.n, np.int64),
                          "p": S.params(**chosen), "chosen": chosen, "table": table, "base": cbase,
                          "cells": bd.n, "nnz": bd.nnz, "cls": ci})
            del bd
        print(f"[bench] class {ci}: {len(idx)} complexes, {cells} cells in {parts} group(s), tuned {chosen} "
              f"({[t['ms'] for t in table]} ms)", file=sys.stderr, flush=True)
    for pl in plans:
        pl["h"].synchronize()
    pool = ThreadPoolExecutor(len(plans))

    def passes(pl, k):
        # k passes over one group, back to back on its stream
        out = [pl["h"].reduce_resident(pl["dm"], pl["out"], pl["p"]) for _ in range(k)]
        pl["h"].synchronize()
        return out

    def step():
        return [f.result()[0] for f in [pool.submit(passes, pl, 1) for pl in plans]]

    # A step is one pass over every group.  The timed region brackets all K steps (every group
    # synchronised before and after), and inside it each group runs its K passes without
    # waiting for the others, so one group's bandwidth-bound init overlaps another's
    # latency-bound add rounds (29.6 ms/step against 32.4 with the groups in lockstep).
    for f in [pool.submit(passes, pl, args.warmup) for pl in plans]:
        f.result()
    stats, last = [], []  # every call's stats; each group's last call
    t = time.perf_counter()
    futs = [pool.submit(passes, pl, args.steps) for pl in plans]
    for f in futs:
        res = f.result()
        stats.extend(st for _, st in res)
        last.append(res[-1][1])
    dt = time.perf_counter() - t
    ms_per_step = dt * 1e3 / args.steps
    diagrams_per_s = args.complexes * args.steps / dt
    print(f"[bench] batched: {ms_per_step:.2f} ms/step, {diagrams_per_s:.0f} diagrams/s", file=sys.stderr, flush=True)
    per = [None] * args.complexes
    for pl, (n_pairs, _) in zip(plans, step()):
        pp = pl["out"].numpy(2 * n_pairs).reshape(-1, 2)
        for k, arr in enumerate(split_pairs(pp, pl["base"])):
            per[pl["idx"][k]] = arr
    class_of = np.zeros(args.complexes, int)
    for pl in plans:
        class_of[pl["idx"]] = pl["cls"]
    pool.shutdown()
    for pl in plans:
        for a in (pl["dm"].colptr, pl["dm"].rows, pl["dm"].dims, pl["out"]):
            a.free()
        pl["h"].close()
    h = S.Handle(0) if args.plain > 0 or args.served > 0 else None

    # parity sample against the oracle (CPU); also the CPU baseline rate
    parity = {"checked": 0, "bit_exact": True}
    cpu = None
    if not args.no_cpu:
        from concurrent.futures import FIRST_COMPLETED, wait
        from oracle import oracle as O
        O.build()
        # one complex per core (the oracle is re-entrant and ctypes drops the GIL during the
        # call), random order, until the time budget is spent
        workers = max(1, min(16, os.cpu_count() or 1))
        order = iter(np.random.default_rng(1).permutation(args.complexes))

        def work(i):
            m = mats[i]
            tt = time.perf_counter()
            ref = O.twist_pairs(O.Matrix(m.colptr, m.rows, m.dims))
            return i, ref, time.perf_counter() - tt

        done, busy = 0, 0.0
        t_start = time.perf_counter()
        with ThreadPoolExecutor(workers) as ex:
            futs = set()
            while True:
                while len(futs) < 2 * workers and time.perf_counter() - t_start < args.cpu_seconds:
                    i = next(order, None)
                    if i is None:
                        break
                    futs.add(ex.submit(work, i))
                if not futs:
                    break
                fin, futs = wait(futs, return_when=FIRST_COMPLETED)
                for f in fin:
                    i, ref, dt = f.result()
                    busy += dt
                    done += 1
                    ok = ref.shape == per[i].shape and np.array_equal(ref, per[i])
                    parity["bit_exact"] &= bool(ok)
        wall = time.perf_counter() - t_start
        parity["checked"] = done
        cpu = {"value": done / wall, "unit": "diagrams/s", "cores": workers, "kind": "port",
               "single_core_value": done / busy,
               "sample": f"{done} complexes of this stream (random order), oracle twist reduction "
                         f"(bit-tree pivot column, clearing), one complex per thread on {workers} threads, "
                         f"{wall:.1f}s wall"}
        print(f"[bench] parity {parity}; cpu {cpu['value']:.1f} diagrams/s on {workers} threads "
              f"({cpu['single_core_value']:.1f}/s per core)", file=sys.stderr, flush=True)

    # plain per-complex GPU: one engine call per complex, matrices already resident
    plain = None
    if args.plain > 0:
        idx = np.random.default_rng(2).permutation(args.complexes)[:args.plain]
        cls_p = [S.params(**c) for c, _ in class_params]
        dev_mats = [(h.upload(mats[i]), cls_p[class_of[i]]) for i in idx]
        d_pairs = h.device_empty(2 * max(mats[i].n for i in idx), np.int64)
        big = int(np.argmax([mats[i].n for i in idx]))  # size the engine's buffers first
        h.reduce_resident(dev_mats[big][0], d_pairs, dev_mats[big][1])
        h.synchronize()
        tt = time.perf_counter()
        for dm, pp in dev_mats:
            h.reduce_resident(dm, d_pairs, pp)
        h.synchronize()
        plain = len(dev_mats) / (time.perf_counter() - tt)
        del dev_mats, d_pairs
        print(f"[bench] plain per-complex GPU: {plain:.0f} diagrams/s", file=sys.stderr, flush=True)

    # served stream: Poisson arrivals; half point clouds (64-4096 points, Rips built by the
    # server), half prebuilt alpha-complex matrices (1e4-1e6 columns); several offered loads
    served = None
    if args.served > 0:
        served = {"requests_per_load": args.served, "loads": []}
        pool_pts, pool_mats, pool_ref = build_served_pool(args.seed + 7)
        print(f"[bench] served pool: {len(pool_pts)} clouds, {len(pool_mats)} matrices "
              f"({min(m.n for m in pool_mats)}-{max(m.n for m in pool_mats)} cells)", file=sys.stderr, flush=True)
        # plain per-request GPU: the same request mix, one request at a time (Rips build on the
        # GPU for clouds, then a single-complex reduction), host data in and pairs out
        rng = np.random.default_rng(99)
        n_plain = max(1, min(args.served, 400))
        kinds = rng.integers(0, 2, n_plain)
        picks = rng.integers(0, 1 << 30, n_plain)
        pp_pts, pp_mat = S.params(**SERVED_PARAMS), S.params(**SERVED_MATRIX_PARAMS)
        # size the engine's buffers with the largest request of each kind first
        h.reduce(max(pool_mats, key=lambda m: m.n), pp_mat)
        Xb, rb = max(pool_pts, key=lambda c: len(c[0]))
        h.reduce(h.vr_build_gpu(Xb, rb, 2), pp_pts)
        tt = time.perf_counter()
        for k in range(n_plain):
            if kinds[k] == 0:
                X, r = pool_pts[picks[k] % len(pool_pts)]
                h.reduce(h.vr_build_gpu(X, r, 2), pp_pts)
            else:
                h.reduce(pool_mats[picks[k] % len(pool_mats)], pp_mat)
        served["plain_per_request_diagrams_per_s"] = round(n_plain / (time.perf_counter() - tt), 1)
        print(f"[bench] served mix, plain one request at a time: "
              f"{served['plain_per_request_diagrams_per_s']} diagrams/s", file=sys.stderr, flush=True)
        h.close()  # the server's workers open their own streams
        for qps in args.qps:
            rng = np.random.default_rng(int(qps))
            kinds = rng.integers(0, 2, args.served)
            picks = rng.integers(0, 1 << 30, args.served)
            gaps = rng.exponential(1.0 / qps, args.served)
            lat = []
            with Server(streams=args.server_streams, max_batch=64, max_batch_cells=args.server_cells,
                        p=S.params(**SERVED_MATRIX_PARAMS), points_p=S.params(**SERVED_PARAMS)) as srv:
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
                    if kinds[k] == 0:
                        X, r = pool_pts[picks[k] % len(pool_pts)]
                        tickets.append((0, picks[k] % len(pool_pts), srv.submit_points(X, r, 2)))
                    else:
                        tickets.append((1, picks[k] % len(pool_mats), srv.submit_matrix(pool_mats[picks[k] % len(pool_mats)])))
                srv.wait_all()
                te = time.perf_counter()
                exact = True
                for kind, j, tk in tickets:
                    res = srv.poll(tk)
                    lat.append(res[2])
                    exact &= bool(np.array_equal(res[0], pool_ref[kind][j]))
            lat = np.array(lat)
            rate = args.served / (te - ts)
            rec = {"qps_offered": qps, "diagrams_per_s": round(rate, 1),
                   "over_plain": round(rate / served["plain_per_request_diagrams_per_s"], 2),
                   "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                   "bit_exact": exact}
            served["loads"].append(rec)
            print(f"[bench] served {rec}", file=sys.stderr, flush=True)

    st = {k: sum(x[k] for x in last) for k in last[0] if isinstance(last[0][k], (int, float))}
    roofline = roofline_record(stats)
    out = {
        "metric": METRIC, "value": round(diagrams_per_s, 2), "unit": "diagrams/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (seeded point clouds -> Vietoris-Rips boundary matrices)",
        "config": {"workload": f"batched: {args.complexes} Rips complexes, n log-uniform [{args.nmin},{args.nmax}], "
                               f"cube/circle/torus, r = 30th percentile, dim<=2, block-diagonal in HBM",
                   "cells": int(N), "nnz": int(NNZ),
                   "streams": args.streams,
                   "groups": [{"size_class": pl["cls"], "complexes": int(len(pl["idx"])), "cells": int(pl["cells"]),
                               "nnz": int(pl["nnz"]), "params": pl["chosen"],
                               "tuning_ms": [t["ms"] for t in pl["table"]]} for pl in plans]},
        "plain_gpu_diagrams_per_s": round(plain, 2) if plain else None,
        "batched_over_plain": round(diagrams_per_s / plain, 2) if plain else None,
        "served": served, "parity": parity,
        "engine": {"pending": st["n_pending"], "apparent": st["n_apparent"], "cleared_hd": st["n_cleared_hd"],
                   "additions": st["n_additions"], "rounds": st["n_rounds"], "pages": st["n_pages"],
                   "tiled_additions": st["n_tiled_additions"],
                   "max_chain": max(x["max_chain"] for x in last),
                   "max_addition_len": max(x["max_addition_len"] for x in last),
                   "t_init_ms": round(st["t_init_ms"], 3), "t_reduce_ms": round(st["t_reduce_ms"], 3)},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
