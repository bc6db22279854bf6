"""Pins the CPU oracle: KATs, closed forms, independent algorithms, the reference's own
collapser, and planted mistakes that the checks must catch."""
import numpy as np
import pytest

import complexes as C
import fixtures as F
from oracle import oracle as O


def test_single_triangle_kat_ascii_and_binary(tmp_path):
    for name, binary in [("single_triangle.dat", False), ("single_triangle.bin", True)]:
        m = O.load_phat(f"{F.GOLDEN}/{name}", binary)
        pairs = O.twist_pairs(m)
        assert pairs.tolist() == [[1, 3], [2, 4], [5, 6]]
        out = tmp_path / ("p.bin" if binary else "p.dat")
        O.save_pairs(str(out), pairs, binary)
        ref = f"{F.GOLDEN}/single_triangle_persistence_pairs." + ("bin" if binary else "dat")
        assert out.read_bytes() == open(ref, "rb").read()


@pytest.mark.parametrize("name,tri,betti", [
    ("torus", C.TORUS_7, [1, 2, 1]), ("rp2", C.RP2_6, [1, 1, 1]),
    ("sphere", C.SPHERE_TET, [1, 0, 1]), ("ball", C.BALL_3, [1, 0, 0, 0])])
@pytest.mark.parametrize("seed", [None, 1, 2, 3])
def test_closed_form_betti(name, tri, betti, seed):
    m = C.closure_matrix(tri, seed)
    p = O.twist_pairs(m)
    assert C.betti_from_pairs(m, p) == betti
    assert (p == O.dense_pairs(m)).all()
    assert (p == O.rank_pairs(m)).all()


@pytest.mark.parametrize("seed", range(12))
def test_twist_matches_rank_formula_and_standard(seed):
    X = C.uniform_points(14 + seed % 5, 3, seed)
    m = C.vr_matrix(X, 0.6, 3)
    if m.n > 128:
        m = C.vr_matrix(X, 0.45, 2)
    p = O.twist_pairs(m)
    assert np.array_equal(p, O.dense_pairs(m))
    if m.n <= 128:
        assert np.array_equal(p, O.rank_pairs(m))


@pytest.mark.parametrize("seed", range(6))
def test_twist_matches_standard_medium(seed):
    X = C.uniform_points(60, 3, 50 + seed)
    m = C.vr_matrix(X, 0.35, 2)
    assert np.array_equal(O.twist_pairs(m), O.dense_pairs(m))


def test_pair_invariants_on_fixtures():
    for name in F.DATA_FIXTURES:
        path = F.fixture_path(name)
        m = O.load_phat(path)
        p = O.twist_pairs(m)
        assert np.array_equal(p, F.expected_pairs(name))
        assert (p[:, 0] < p[:, 1]).all()
        assert (np.diff(p[:, 0]) > 0).all()
        assert (m.dims[p[:, 1]] == m.dims[p[:, 0]] + 1).all()
        assert len(np.unique(p)) == 2 * len(p)


def test_collapse_restatement_matches_reference_fixtures():
    for u, v, kept in F.kept_edge_cases():
        assert np.array_equal(O.collapse_edges(u, v), kept)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref (reference collapser) not built here")
@pytest.mark.parametrize("seed", range(10))
def test_collapse_restatement_matches_reference_build(seed):
    X = C.uniform_points(40 + 10 * seed, 3, 900 + seed)
    m = C.vr_matrix(X, 0.3 + 0.03 * seed, 1)
    e = np.where(m.dims == 1)[0]
    u, v = m.rows[m.colptr[e]], m.rows[m.colptr[e] + 1]
    assert np.array_equal(O.collapse_edges(u, v), O.ref_collapse_edges(u, v))


def test_collapse_subcomplex_invariants():
    X = C.uniform_points(50, 3, 7)
    m = C.vr_matrix(X, 0.45, 2)
    skip = O.collapse_skip(m)
    e = np.where(m.dims == 1)[0]
    assert not skip[m.dims == 0].any()
    assert skip[e[0]] == 0  # the first edge is always kept
    # removed cells never appear as faces of kept cells
    for c in np.where(skip == 0)[0]:
        assert not skip[m.rows[m.colptr[c]:m.colptr[c + 1]]].any()
    # collapse preserves H0 of the whole flag complex
    p_full = O.twist_pairs(m)
    p_sub = O.twist_pairs(m, skip=skip)
    ess_full = C.betti_from_pairs(m, p_full)[0]
    paired = set(p_sub.ravel().tolist())
    ess_sub0 = sum(1 for c in np.where(m.dims == 0)[0] if c not in paired)
    assert ess_full == ess_sub0


def _low_betti(m, pairs, skip):
    paired = np.zeros(m.n, bool)
    paired[pairs[:, 0]] = True
    paired[pairs[:, 1]] = True
    alive = ~paired & (skip == 0)
    return [int((alive & (m.dims == k)).sum()) for k in (0, 1)]


def _collapse_cases():
    rng = np.random.default_rng(1)
    for _ in range(40):
        n = int(rng.integers(20, 90))
        yield C.vr_matrix(rng.random((n, 3)), float(rng.uniform(0.15, 0.5)), 2)


def test_collapse_keeps_beta0_and_beta1():
    """A strong collapse keeps the flag complex's homotopy type: the kept sub-complex has the
    full complex's beta_0 and beta_1 (its 2-skeleton's beta_2 may differ)."""
    for m in _collapse_cases():
        skip = O.collapse_skip(m)
        assert _low_betti(m, O.twist_pairs(m, skip=skip), skip) == \
            _low_betti(m, O.twist_pairs(m), np.zeros(m.n, np.uint8))


def test_collapse_betti_check_catches_over_removal():
    """Planted mistake: one more kept edge (the last kept one) removed with its cofaces changes
    beta_0 or beta_1 in some of the same cases."""
    caught = 0
    for m in _collapse_cases():
        skip = O.collapse_skip(m).copy()
        kept_edges = np.where((m.dims == 1) & (skip == 0))[0]
        if len(kept_edges) < 2:
            continue
        skip[kept_edges[-1]] = 1
        for c in np.where(m.dims == 2)[0]:  # its cofaces too
            if skip[m.rows[m.colptr[c]:m.colptr[c + 1]]].any():
                skip[c] = 1
        caught += _low_betti(m, O.twist_pairs(m, skip=skip), skip) != \
            _low_betti(m, O.twist_pairs(m), np.zeros(m.n, np.uint8))
    assert caught >= 10


def test_planted_mistakes_are_caught():
    """A wrong pairing must fail the independent checks used above."""
    m = C.closure_matrix(C.TORUS_7, 2)
    good = O.rank_pairs(m)
    bad = good.copy()
    bad[0, 1] += 1  # wrong death index
    assert not np.array_equal(bad, O.twist_pairs(m))
    # dropping a boundary entry changes the pairing or the Betti numbers
    rows = m.rows.copy()
    rows[-1] = rows[-2]
    cols = [sorted(set(rows[m.colptr[c]:m.colptr[c + 1]].tolist())) for c in range(m.n)]
    m2 = O.Matrix.from_columns(cols, m.dims)
    assert C.betti_from_pairs(m2, O.twist_pairs(m2)) != [1, 2, 1]
    # a collapse that keeps every edge differs from GUDHI's kept set on a dense graph
    u, v, kept = F.kept_edge_cases()[1]
    assert not np.array_equal(np.ones_like(kept), kept)


@pytest.mark.parametrize("seed", range(10))
def test_cohomology_twist_matches_homology_and_rank_formula(seed):
    """Persistence duality: the twist reduction of the anti-transpose gives the same pairs as
    the primal twist and as the rank formula (tiny inputs), with and without a sub-complex."""
    X = C.uniform_points(12 + seed % 6, 3, 300 + seed)
    m = C.vr_matrix(X, 0.55, 3)
    if m.n > 128:
        m = C.vr_matrix(X, 0.4, 2)
    p = O.twist_pairs(m, dual=True)
    assert np.array_equal(p, O.twist_pairs(m))
    if m.n <= 96:
        assert np.array_equal(p, O.rank_pairs(m))
    if (m.dims == 1).sum() > 2:
        skip = O.collapse_skip(m)
        assert np.array_equal(O.twist_pairs(m, skip=skip, dual=True), O.twist_pairs(m, skip=skip))


@pytest.mark.parametrize("name,tri", [("torus", C.TORUS_7), ("rp2", C.RP2_6), ("ball", C.BALL_3)])
def test_cohomology_twist_closed_forms(name, tri):
    m = C.closure_matrix(tri, 5)
    assert np.array_equal(O.twist_pairs(m, dual=True), O.rank_pairs(m))


def test_cohomology_twist_on_fixtures():
    for name in F.DATA_FIXTURES:
        m = O.load_phat(F.fixture_path(name))
        assert np.array_equal(O.twist_pairs(m, dual=True), F.expected_pairs(name)), name


def test_cohomology_planted_mistake_is_caught():
    """A dual that maps pairs back without the reversal (b', d') -> (b', d') disagrees."""
    m = C.closure_matrix(C.TORUS_7, 1)
    p = O.twist_pairs(m, dual=True)
    n = m.n
    wrong = np.stack([n - 1 - p[:, 0], n - 1 - p[:, 1]], 1)
    assert not np.array_equal(wrong[np.argsort(wrong[:, 0])], O.rank_pairs(m))


def _edges_of(m):
    e = np.where(np.asarray(m.dims) == 1)[0]
    cp, rows = np.asarray(m.colptr), np.asarray(m.rows)
    return rows[cp[e]].astype(np.int32), rows[cp[e] + 1].astype(np.int32)


def test_collapse_on_the_large_rips_edge_list():
    """The full edge list of the 4096-rotation Rips complex (r = 1.0, 163,560 edges): the
    library's host collapse (parallel main loop, cascades decided from the last witness where
    they can be) keeps exactly the oracle's edges, and the reference's GUDHI collapser's when
    oracle/_ref is built."""
    import specseqplusplus_amd as S
    from specseqplusplus_amd import configs as CF
    u, v = _edges_of(S.vr_build(CF.c2_points(), 1.0, 2))
    assert len(u) == 163560
    kept = O.collapse_edges(u, v)
    assert np.array_equal(S.collapse_edges(u, v), kept)
    assert np.array_equal(S.collapse_edges(u, v, threads=1), kept)
    if O.ref_available():
        assert np.array_equal(O.ref_collapse_edges(u, v), kept)


@pytest.mark.parametrize("seed", range(12))
def test_library_collapse_equals_oracle(seed):
    import specseqplusplus_amd as S
    rng = np.random.default_rng(600 + seed)
    X = rng.random((60 + 15 * seed, 2 + seed % 3)).astype(np.float32)
    u, v = _edges_of(S.vr_build(X, 0.2 + 0.02 * seed, 1))
    assert np.array_equal(S.collapse_edges(u, v), O.collapse_edges(u, v))


def test_dependency_depth_bounds():
    """The oracle's dependency depth (longest chain of additions when each waits only for the
    pivot it adds) lies between the longest single chain and the number of additions, and is
    0 when nothing is added."""
    import specseqplusplus_amd as S
    for seed in range(6):
        X = np.random.default_rng(70 + seed).random((40 + 20 * seed, 3)).astype(np.float32)
        m = S.vr_build(X, 0.35, 2)
        for dual in (False, True):
            _, st = O.twist_pairs(O.Matrix(m.colptr, m.rows, m.dims), stats=True, dual=dual)
            assert st["max_chain"] <= st["depth"] <= st["additions"]
    tri = O.Matrix(np.array([0, 0, 0, 0, 2, 4, 6, 9]), np.array([0, 1, 0, 2, 1, 2, 3, 4, 5], np.int32),
                   np.array([0, 0, 0, 1, 1, 1, 2], np.int32))
    _, st = O.twist_pairs(tri, stats=True)
    assert st["additions"] == 0 and st["depth"] == 0
