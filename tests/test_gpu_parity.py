"""GPU parity: the HIP engine (through the C ABI) against the oracle and the golden
fixtures.  Integer work: every comparison is bit-exact."""
import os
import subprocess

import numpy as np
import pytest

import complexes as C
import fixtures as F
from oracle import oracle as O

import specseqplusplus_amd as S
from specseqplusplus_amd import _lib

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "specseqplusplus_amd", "bin", "specseq++")


@pytest.fixture(scope="module")
def h():
    hd = S.Handle(0)
    yield hd
    hd.close()


def as_product(mo):
    return S.Matrix(mo.colptr, mo.rows, mo.dims)


def oracle_pairs(mo, model=0):
    return O.twist_pairs(mo, skip=O.collapse_skip(mo) if model & 2 else None)


def check(h, mo, **kw):
    pairs, ess, st = h.reduce(as_product(mo), S.params(**kw), stats=True)
    ref = oracle_pairs(mo, kw.get("model", 0))
    assert pairs.shape == ref.shape and np.array_equal(pairs, ref), kw
    # every cell is a birth, a death or essential (removed cells excluded)
    skip = O.collapse_skip(mo) if kw.get("model", 0) & 2 else np.zeros(mo.n, np.uint8)
    assert 2 * len(pairs) + len(ess) == int((skip == 0).sum())
    return pairs, ess, st


def c1_matrix():
    X = np.asarray(C.uniform_points(512, 3, 0), dtype=np.float32)
    m = S.vr_build(X, 0.25, 2)
    return O.Matrix(m.colptr, m.rows, m.dims)


def test_single_triangle_kat_via_cli(tmp_path):
    for fmt, ext in [("ascii", "dat"), ("binary", "bin")]:
        out = tmp_path / f"o.{ext}"
        r = subprocess.run([CLI, f"{F.GOLDEN}/single_triangle.{ext}", str(out), fmt, fmt, "0", "0", "0", "102400",
                            "1.05", "65536", "1000", "2000"], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        assert out.read_bytes() == open(f"{F.GOLDEN}/single_triangle_persistence_pairs.{ext}", "rb").read()


@pytest.mark.parametrize("tri", [C.TORUS_7, C.RP2_6, C.SPHERE_TET, C.BALL_3])
@pytest.mark.parametrize("seed", [None, 1, 5])
def test_closed_forms(h, tri, seed):
    check(h, C.closure_matrix(tri, seed))


@pytest.mark.parametrize("seed", range(16))
def test_random_vr_small(h, seed):
    X = C.uniform_points(30 + 3 * seed, 3, seed)
    check(h, C.vr_matrix(X, 0.35 + 0.01 * seed, 2 + seed % 2))


@pytest.mark.parametrize("name", F.DATA_FIXTURES)
def test_shipped_fixtures(h, name):
    m = S.load_phat(F.fixture_path(name))
    pairs, _ = h.reduce(m)
    assert np.array_equal(pairs, F.expected_pairs(name))


@pytest.mark.parametrize("kw", [
    {}, {"level1": 1, "level2": 2}, {"level1": 4, "level2": 8}, {"queue_size": 8}, {"ss_block_size": 64},
    {"ss_block_size": 1000, "growth": 1.0}, {"ss_block_size": 1 << 30}, {"model": 1}, {"model": 1, "max_iteration": 0},
    {"model": 2}, {"model": 3}, {"model": 3, "level1": 2, "level2": 3}, {"dualize": 1}])
def test_config1_variants(h, kw):
    check(h, c1_matrix(), **kw)


def test_guided_clearing_is_exact_and_clears(h):
    mo = c1_matrix()
    p0, _, s0 = check(h, mo)
    p1, _, s1 = check(h, mo, model=1, max_iteration=50)
    assert np.array_equal(p0, p1) and s1["n_cleared_hd"] > 0 and s1["n_pending"] < s0["n_pending"]


def test_model2_uses_reference_kept_edges(h):
    # the oracle's collapse is pinned to the reference's GUDHI build; the engine's sub-complex
    # pairs must equal the oracle's pairs of that sub-complex
    for seed in range(4):
        X = C.uniform_points(60, 3, 300 + seed)
        check(h, C.vr_matrix(X, 0.4, 2), model=2)


def test_repeat_calls_deterministic(h):
    mo = c1_matrix()
    first = h.reduce(as_product(mo))[0]
    for _ in range(3):
        assert np.array_equal(h.reduce(as_product(mo))[0], first)


@pytest.mark.parametrize("kw", [{"dualize": 1}, {"dualize": 1, "model": 2}, {"dualize": 1, "model": 3},
                                {"dualize": 1, "level1": 1, "level2": 2, "ss_block_size": 100}])
def test_dualize_variants_small(h, kw):
    for seed in range(3):
        X = C.uniform_points(45, 3, 500 + seed)
        check(h, C.vr_matrix(X, 0.4, 3), **kw)


def test_dualize_device_resident_and_batch(h):
    mats = []
    for seed in range(6):
        X = np.asarray(C.uniform_points(50 + 9 * seed, 3, 60 + seed), dtype=np.float32)
        mats.append(S.vr_build(X, 0.3, 2))
    p = S.params(dualize=1)
    res = h.batch_reduce(mats, p)
    for m, (pairs, ess) in zip(mats, res):
        mo = O.Matrix(m.colptr, m.rows, m.dims)
        ref = O.twist_pairs(mo)
        assert np.array_equal(pairs, ref)
        assert 2 * len(pairs) + len(ess) == m.n
    m = mats[-1]
    dm = h.upload(m)
    out = h.device_empty(2 * m.n, np.int64)
    np_, _ = h.reduce_resident(dm, out, p)
    assert np.array_equal(out.numpy(2 * np_).reshape(-1, 2), O.twist_pairs(O.Matrix(m.colptr, m.rows, m.dims)))


def test_batch_equals_individual(h):
    mats = []
    for seed in range(12):
        X = np.asarray(C.uniform_points(40 + 7 * seed, 3, 40 + seed), dtype=np.float32)
        mats.append(S.vr_build(X, 0.3, 2))
    res = h.batch_reduce(mats)
    for m, (pairs, ess) in zip(mats, res):
        alone, ess_alone = h.reduce(m)
        assert np.array_equal(pairs, alone) and np.array_equal(ess, ess_alone)
        assert np.array_equal(pairs, O.twist_pairs(O.Matrix(m.colptr, m.rows, m.dims)))


def test_resident_batch_equals_host_batch(h):
    """ssq_batch_reduce_resident on matrices placed once with ssq_matrix_to_device: each
    block's pairs and essentials equal the host batch's and the oracle's, with an empty matrix
    between two others and one matrix used twice in the batch, in homology and cohomology."""
    mats = []
    for seed in range(9):
        X = np.asarray(C.uniform_points(30 + 11 * seed, 3, 90 + seed), dtype=np.float32)
        mats.append(S.vr_build(X, 0.3, 2))
    mats.insert(4, S.Matrix(np.zeros(1, np.int64), np.zeros(0, np.int32), np.zeros(0, np.int32)))
    dms = [h.to_device(m) for m in mats]
    try:
        order = list(range(len(mats))) + [2]
        for kw in ({}, {"dualize": 1}):
            res = h.batch_reduce_resident([dms[i] for i in order], S.params(**kw))
            host = h.batch_reduce([mats[i] for i in order], S.params(**kw))
            for i, (pairs, ess), (hp, he) in zip(order, res, host):
                m = mats[i]
                assert np.array_equal(pairs, hp) and np.array_equal(ess, he), (i, kw)
                assert np.array_equal(pairs, O.twist_pairs(O.Matrix(m.colptr, m.rows, m.dims))), (i, kw)
        assert h.batch_reduce_resident([]) == []
    finally:
        for dm in dms:
            h.free_device_matrix(dm)


def test_resident_placement_checks_the_matrix(h):
    with pytest.raises(S.SSQError) as e:
        h.to_device(_malformed("unsorted"))
    assert e.value.status == _lib.SSQ_EFORMAT
    good = as_product(c1_matrix())
    dm = h.to_device(good)
    try:
        assert np.array_equal(h.batch_reduce_resident([dm])[0][0], O.twist_pairs(c1_matrix()))
    finally:
        h.free_device_matrix(dm)


def test_device_resident_entry_point(h):
    mo = c1_matrix()
    dm = h.upload(as_product(mo))
    out = h.device_empty(2 * mo.n, np.int64)
    np_, st = h.reduce_resident(dm, out)
    pairs = out.numpy(2 * np_).reshape(-1, 2)
    assert np.array_equal(pairs, O.twist_pairs(mo))


def test_medium_vr_properties(h):
    X = np.asarray(C.uniform_points(3000, 3, 11), dtype=np.float32)
    m = S.vr_build(X, 0.1, 2)
    pairs, ess = h.reduce(m, S.params(model=1))
    mo = O.Matrix(m.colptr, m.rows, m.dims)
    assert np.array_equal(pairs, O.twist_pairs(mo))
    assert (pairs[:, 0] < pairs[:, 1]).all() and (np.diff(pairs[:, 0]) > 0).all()
    assert (m.dims[pairs[:, 1]] == m.dims[pairs[:, 0]] + 1).all()


def test_invalid_parameters_are_rejected(h):
    mo = c1_matrix()
    for kw in [{"ss_block_size": 0}, {"growth": 0.5}, {"queue_size": 0}, {"level1": 0}, {"model": 4}]:
        with pytest.raises(S.SSQError):
            h.reduce(as_product(mo), S.params(**kw))


def _malformed(kind):
    """A small valid complex with one planted defect (rows and offsets checked on the device)."""
    mo = c1_matrix()
    cp, rows, dims = mo.colptr.copy(), mo.rows.copy(), mo.dims.copy()
    c = int(np.flatnonzero(np.diff(cp) >= 2)[5])  # a column with at least two rows
    if kind == "row_out_of_range":
        rows[cp[c + 1] - 1] = len(dims)
    elif kind == "negative_row":
        rows[cp[c]] = -1
    elif kind == "unsorted":
        rows[cp[c]], rows[cp[c] + 1] = rows[cp[c] + 1], rows[cp[c]]
    elif kind == "duplicate":
        rows[cp[c] + 1] = rows[cp[c]]
    elif kind == "offsets_not_monotone":
        cp[c + 1] = cp[c] - 1
    elif kind == "last_offset":
        cp[-1] -= 1
    return S.Matrix(cp, rows, dims)


@pytest.mark.parametrize("kind", ["row_out_of_range", "negative_row", "unsorted", "duplicate",
                                  "offsets_not_monotone", "last_offset"])
def test_malformed_matrix_is_rejected(h, kind):
    bad = _malformed(kind)
    good = as_product(c1_matrix())
    ref = h.reduce(good)[0]
    with pytest.raises(S.SSQError) as e:
        h.reduce(bad)
    assert e.value.status == _lib.SSQ_EFORMAT
    with pytest.raises(S.SSQError) as e:  # the defect is found inside a batch too
        h.batch_reduce([good, bad, good])
    assert e.value.status == _lib.SSQ_EFORMAT
    # the handle stays usable and exact
    assert np.array_equal(h.reduce(good)[0], ref)
    outs = h.batch_reduce([good, good])
    assert all(np.array_equal(o[0], ref) for o in outs)


def _circle(n, seed):
    rng = np.random.default_rng(seed)
    t = rng.random(n) * 2 * np.pi
    X = np.stack([np.cos(t), np.sin(t), np.zeros(n)], 1) + rng.normal(0, 0.05, (n, 3))
    from scipy.spatial.distance import pdist
    return X.astype(np.float32), float(np.percentile(pdist(X.astype(np.float32).astype(np.float64)), 30))


@pytest.mark.parametrize("env", [{}, {"SSQ_TEST_SHORT_MAX": "16"},
                                 {"SSQ_TEST_SHORT_MAX": "16", "SSQ_TEST_WINDOW_BITS": "256"},
                                 {"SSQ_TEST_WINDOW_BITS": "4096"},
                                 {"SSQ_TEST_WINDOW_BITS": "256", "SSQ_TEST_GLOBAL_BITS": "2048"},
                                 {"SSQ_TEST_SHORT_MAX": "16", "SSQ_TEST_WINDOW_BITS": "64",
                                  "SSQ_TEST_GLOBAL_BITS": "64"}])
def test_long_column_paths(env, monkeypatch):
    """Noisy circles in cohomology grow long cocycle columns: the LDS bit set, the device-memory
    bit set for columns that leave the LDS window, the tiled merges past that one, each forced
    by smaller limits."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    hd = S.Handle(0)
    try:
        for seed in range(2):
            X, r = _circle(150, seed)
            m = S.vr_build(X, r, 2)
            mo = O.Matrix(m.colptr, m.rows, m.dims)
            for kw in [{"dualize": 1}, {"dualize": 0, "model": 1}]:
                pairs, ess, st = hd.reduce(m, S.params(**kw), stats=True)
                assert np.array_equal(pairs, O.twist_pairs(mo)), (env, kw, seed)
    finally:
        hd.close()


@pytest.mark.parametrize("case", range(6))
def test_gpu_rips_builder_matches_host(h, case):
    from specseqplusplus_amd import configs as CF
    rng = np.random.default_rng(case)
    if case < 3:
        X = CF.cloud(rng, case, 180 + 40 * case).astype(np.float32)
        from scipy.spatial.distance import pdist
        r = float(np.percentile(pdist(X.astype(np.float64)), 30))
    elif case == 3:
        X, r = CF.haar_rotations(600, 3), 1.0
    elif case == 4:
        X, r = rng.random((1, 3)).astype(np.float32), 0.5   # a single point
    else:
        X, r = rng.random((50, 3)).astype(np.float32), 0.0  # no edges
    for maxdim in (0, 1, 2):
        a = S.vr_build(X, r, maxdim)
        b = h.vr_build_gpu(X, r, maxdim)
        assert np.array_equal(a.dims, b.dims) and np.array_equal(a.colptr, b.colptr), (case, maxdim)
        assert np.array_equal(a.rows, b.rows), (case, maxdim)


@pytest.mark.parametrize("maxdim", [1, 2])
def test_batch_reduce_points_equals_alone(h, maxdim):
    """Clouds built on the GPU as one block-diagonal batch and reduced together give, per
    cloud, the pairs of its own host-built Rips matrix reduced alone (and the oracle's)."""
    from scipy.spatial.distance import pdist
    from specseqplusplus_amd import configs as CF
    rng = np.random.default_rng(40 + maxdim)
    clouds = []
    for i in range(13):
        n = [0, 1, 2, 37, 90, 150, 64, 200, 5, 120, 80, 33, 170][i]
        X = CF.cloud(rng, i % 3, n).astype(np.float32) if n else np.zeros((0, 3), np.float32)
        r = float(np.percentile(pdist(X.astype(np.float64)), 25)) if n > 1 else 0.5
        clouds.append((X, r))
    for p in [S.params(dualize=1, ss_block_size=1 << 30), S.params()]:
        outs = h.batch_reduce_points(clouds, maxdim, p)
        for (X, r), (pairs, ess) in zip(clouds, outs):
            m = S.vr_build(X, r, maxdim)
            ref, ref_ess = h.reduce(m, p)
            assert np.array_equal(pairs, ref) and np.array_equal(ess, ref_ess), (len(X), maxdim)
            mo = O.Matrix(m.colptr, m.rows, m.dims)
            assert np.array_equal(pairs, O.twist_pairs(mo))


DUAL_PATHS = {"sort": {}, "atomic": {"SSQ_TEST_NO_SORT_DUAL": "1"}}


@pytest.mark.parametrize("path", list(DUAL_PATHS))
def test_dual_transpose_paths(monkeypatch, path):
    """Cohomology through each anti-transpose (radix sort, the default; atomic counts and
    fills, forced by the test hook) gives the oracle's pairs, batched and device-resident."""
    from scipy.spatial.distance import pdist
    from specseqplusplus_amd import configs as CF
    for k, v in DUAL_PATHS[path].items():
        monkeypatch.setenv(k, v)
    hd = S.Handle(0)
    try:
        rng = np.random.default_rng(77)
        mats = []
        for i, n in enumerate([3, 60, 1, 140, 0, 90, 200, 17, 110]):
            X = CF.cloud(rng, i % 3, n).astype(np.float32) if n else np.zeros((0, 3), np.float32)
            r = float(np.percentile(pdist(X.astype(np.float64)), 30)) if n > 1 else 0.5
            mats.append(S.vr_build(X, r, 2))
        p = S.params(dualize=1, ss_block_size=1 << 30)
        outs = hd.batch_reduce(mats, p)
        for m, (pairs, ess) in zip(mats, outs):
            mo = O.Matrix(m.colptr, m.rows, m.dims)
            assert np.array_equal(pairs, O.twist_pairs(mo)), (path, m.n)
        # the same batch resident in HBM (the path of a device-resident caller)
        from specseqplusplus_amd.autotune import block_diagonal
        bd, base = block_diagonal(mats)
        dm = hd.upload(bd)
        out = hd.device_empty(2 * bd.n, np.int64)
        npairs, _ = hd.reduce_resident(dm, out, p)
        allp = out.numpy(2 * npairs).reshape(-1, 2)
        assert np.array_equal(allp, O.twist_pairs(O.Matrix(bd.colptr, bd.rows, bd.dims)))
    finally:
        hd.close()


@pytest.mark.parametrize("path", list(DUAL_PATHS))
def test_dual_long_coface_row(monkeypatch, path):
    """A vertex on 3000 edges: one 3000-entry dual column next to short ones; each
    anti-transpose stays exact, alone and next to ordinary complexes in a batch."""
    for k, v in DUAL_PATHS[path].items():
        monkeypatch.setenv(k, v)
    k = 3000
    nv = k + 1
    colptr = [0] * (nv + 1)
    rows = []
    for i in range(1, k + 1):  # edges (0, i), then the triangles (0, i, i+1) for the first few i
        rows += [0, i]
        colptr.append(len(rows))
    ntri = 40
    for i in range(1, ntri + 1):
        # edge ids: edge (0, i) is column nv + i - 1; add edge (i, i+1) first
        rows += [i, i + 1]
        colptr.append(len(rows))
    base_e2 = nv + k
    for i in range(1, ntri + 1):
        rows += sorted([nv + i - 1, nv + i, base_e2 + i - 1])
        colptr.append(len(rows))
    dims = [0] * nv + [1] * k + [1] * ntri + [2] * ntri
    m = S.Matrix(np.array(colptr, np.int64), np.array(rows, np.int32), np.array(dims, np.int32))
    mo = O.Matrix(m.colptr, m.rows, m.dims)
    ref = O.twist_pairs(mo)
    hd = S.Handle(0)
    try:
        p = S.params(dualize=1, ss_block_size=1 << 30)
        assert np.array_equal(hd.reduce(m, p)[0], ref)
        small = as_product(c1_matrix())
        outs = hd.batch_reduce([small, m, small], p)
        assert np.array_equal(outs[1][0], ref)
        assert np.array_equal(outs[0][0], outs[2][0])
    finally:
        hd.close()


@pytest.mark.parametrize("n", [4095, 4096, 4097, 8191, 8193, 12289])
def test_cell_counts_at_scan_tile_boundaries(h, n):
    """Exactly n cells around multiples of the 4096-count scan tile (the column pointers of the
    anti-transpose and the compactions scan over n): a path on v vertices and isolated
    vertices to pad; homology and cohomology against the oracle."""
    v = (n + 1) // 2
    pad = n - (2 * v - 1)
    colptr, rows, dims = [0], [], []
    for _ in range(v + pad):  # vertices 0..v-1 on the path, then the isolated ones
        colptr.append(len(rows))
        dims.append(0)
    edge0 = v + pad
    for i in range(v - 1):  # edge (i, i+1) is column edge0 + i
        rows += [i, i + 1]
        colptr.append(len(rows))
        dims.append(1)
    mo = O.Matrix(np.array(colptr, np.int64), np.array(rows, np.int32), np.array(dims, np.int32))
    assert mo.n == n
    for kw in [{}, {"dualize": 1}]:
        check(h, mo, **kw)


def test_dual_sort_nine_bit_digits(h):
    """A dualized batch of 2^16..2^18 cells: its sort keys have 17-18 bits, where the
    anti-transpose sorts with 9-bit digits (two passes instead of three 8-bit ones).  Every
    complex's pairs equal the oracle's."""
    from scipy.spatial.distance import pdist
    from specseqplusplus_amd import configs as CF
    rng = np.random.default_rng(2718)
    mats, total, i = [], 0, 0
    while total < 100_000:
        X = CF.cloud(rng, i % 3, 90 + (i % 5) * 10).astype(np.float32)
        r = float(np.percentile(pdist(X.astype(np.float64)), 30))
        m = S.vr_build(X, r, 2)
        mats.append(m)
        total += m.n
        i += 1
    assert (1 << 16) <= total < (1 << 18)
    outs = h.batch_reduce(mats, S.params(dualize=1, ss_block_size=1 << 30))
    for m, (pairs, ess) in zip(mats, outs):
        assert np.array_equal(pairs, O.twist_pairs(O.Matrix(m.colptr, m.rows, m.dims)))


def test_two_phases_equal_one_call(h):
    """ssq_init + ssq_compute (the reference's GPU_Init / GPU_Compute) give ssq_reduce's pairs in
    every mode; compute without an initialized matrix, or after another reduction discarded it,
    is refused."""
    mo = c1_matrix()
    m = as_product(mo)
    for kw in (dict(), dict(dualize=1), dict(model=1), dict(model=3), dict(level1=1, level2=2, ss_block_size=64)):
        p = S.params(**kw)
        st_i = h.init(m, p)
        pairs, ess, st_c = h.compute(stats=True)
        ref, ref_ess = h.reduce(m, p)
        assert np.array_equal(pairs, ref) and np.array_equal(ess, ref_ess), kw
        assert st_i["t_init_ms"] > 0 and st_c["t_reduce_ms"] > 0 and st_c["n_pending"] == st_i["n_pending"]
    with pytest.raises(_lib.SSQError):
        h.compute()
    h.init(m)
    h.reduce(m)
    with pytest.raises(_lib.SSQError):
        h.compute()


def test_reference_class_mirror():
    from specseqplusplus_amd.specseq import SpecSeqPlusPlus
    mo = c1_matrix()
    cols = [mo.rows[mo.colptr[c]:mo.colptr[c + 1]].tolist() for c in range(mo.n)]
    ss = SpecSeqPlusPlus(cols, mo.dims.tolist(), [0, 1000, 2000, 0], dualize=0, model=1)
    t_init = ss.GPU_Init()
    t_comp = ss.GPU_Compute()
    assert t_init > 0 and t_comp > 0
    got = np.array(ss.GetPersistencePairs(), dtype=np.int64).reshape(-1, 2)
    assert np.array_equal(got, O.twist_pairs(mo))
    assert 2 * len(got) + len(ss.GetEssential()) == mo.n


SHAPE_VARIANTS = [dict(add_waves=2), dict(add_waves=8), dict(short_max=64), dict(short_max=256),
                  dict(long_grid=1), dict(long_grid=4096), dict(rounds_per_sync=1), dict(rounds_per_sync=8),
                  dict(add_waves=2, short_max=64, long_grid=7, rounds_per_sync=5)]


@pytest.mark.parametrize("shape", SHAPE_VARIANTS, ids=lambda s: ",".join(f"{k}={v}" for k, v in s.items()))
def test_workgroup_shapes_are_exact(h, shape):
    """The additions' workgroup shape never changes the pairs: homology and cohomology of C1,
    and the batched Rips path (where long columns appear)."""
    mo = c1_matrix()
    for base in (dict(), dict(dualize=1), dict(model=1, level1=1, level2=2)):
        pairs, ess = h.reduce(as_product(mo), S.params(**base, **shape))
        assert np.array_equal(pairs, O.twist_pairs(mo)), (base, shape)
    rng = np.random.default_rng(3)
    clouds = [(rng.random((n, 3)).astype(np.float32), 0.45) for n in (120, 300, 500)]
    outs = h.batch_reduce_points(clouds, 2, S.params(dualize=1, ss_block_size=1 << 30, **shape))
    for (X, r), (pairs, _ess) in zip(clouds, outs):
        m = S.vr_build(X, r, 2)
        assert np.array_equal(pairs, O.twist_pairs(O.Matrix(m.colptr, m.rows, m.dims))), shape


def test_shape_out_of_range_is_refused(h):
    mo = c1_matrix()
    for bad in (dict(add_waves=3), dict(short_max=32), dict(short_max=2048), dict(long_grid=-1),
                dict(rounds_per_sync=9)):
        with pytest.raises(_lib.SSQError):
            h.reduce(as_product(mo), S.params(**bad))


def test_autotuner_chooses_params_and_shape(h):
    from specseqplusplus_amd.autotune import tune
    mats = []
    for seed in range(6):
        X = np.asarray(C.uniform_points(60 + 20 * seed, 3, 300 + seed), dtype=np.float32)
        mats.append(S.vr_build(X, 0.3, 2))
    chosen, table = tune(h, mats, repeats=1)
    assert len(table) >= 9 and table[0]["params"] == chosen
    res = h.batch_reduce(mats, S.params(**chosen))
    for m, (pairs, _ess) in zip(mats, res):
        assert np.array_equal(pairs, O.twist_pairs(O.Matrix(m.colptr, m.rows, m.dims)))
