"""Adversarial column additions through the C ABI: matrices whose additions are a merge of
long equal runs that cross every merge segment, one-sided merges, full cancellation and
interleaved lists, each sized for the wavefront LDS merge (|a|+|b| <= 1024), the tiled merge and
the long-column bit set.  The pairs must equal the oracle's; the reduced columns are the
symmetric differences std::set_symmetric_difference would give."""
import numpy as np
import pytest

from oracle import oracle as O

import specseqplusplus_amd as S

pytestmark = pytest.mark.gpu


def crafted(scale, seed):
    """Dimension-0 rows [0, R), then dimension-1 columns that all share the low row R-1, so
    that every later column adds the first one (and what it became)."""
    R = 40 * scale + 8
    top = R - 1
    rng = np.random.default_rng(seed)
    base = list(range(0, 30 * scale))
    cols = [
        base + [top],                                        # pivot of row top
        base + [top],                                        # cancels completely
        list(range(10 * scale, 20 * scale)) + [top],         # equal run in the middle
        [35 * scale, top],                                   # one-sided: nearly all from the pivot
        base + [35 * scale + 1, top],                        # leaves one row
        list(range(0, 30 * scale, 2)) + [top],               # interleaved halves
        sorted(set(rng.choice(30 * scale, 10 * scale, replace=False).tolist())) + [top],
        list(range(1, 30 * scale)) + [36 * scale, top],      # off by one at the start
        list(range(0, 30 * scale - 1)) + [37 * scale, top],  # off by one at the end
    ]
    dims = [0] * R + [1] * len(cols)
    allcols = [[] for _ in range(R)] + [sorted(set(c)) for c in cols]
    return O.Matrix.from_columns(allcols, dims)


@pytest.mark.parametrize("scale", [4, 12, 40, 120])
@pytest.mark.parametrize("dual", [0, 1])
def test_crafted_additions_match_oracle(scale, dual):
    m = crafted(scale, scale)
    h = S.Handle(0)
    try:
        for kw in (dict(), dict(queue_size=1), dict(level1=1, level2=2)):
            pairs, ess = h.reduce(S.Matrix(m.colptr, m.rows, m.dims), S.params(dualize=dual, **kw))
            ref = O.twist_pairs(m)
            assert np.array_equal(pairs, ref), (scale, dual, kw)
            assert 2 * len(pairs) + len(ess) == m.n
    finally:
        h.close()


@pytest.mark.parametrize("env", [{"SSQ_TEST_SHORT_MAX": "16"}, {"SSQ_TEST_SHORT_MAX": "16", "SSQ_TEST_WINDOW_BITS": "64"},
                                 {"SSQ_TEST_SHORT_MAX": "16", "SSQ_TEST_WINDOW_BITS": "64", "SSQ_TEST_GLOBAL_BITS": "64"}])
def test_crafted_additions_long_paths(monkeypatch, env):
    """The same additions forced through the LDS bit set, the device-memory bit set and the
    tiled merges."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    h = S.Handle(0)
    try:
        for scale in (4, 40):
            m = crafted(scale, 100 + scale)
            pairs, _ = h.reduce(S.Matrix(m.colptr, m.rows, m.dims), S.params())
            assert np.array_equal(pairs, O.twist_pairs(m)), scale
    finally:
        h.close()


def test_crafted_reduced_columns_are_symmetric_differences():
    """What the pairs above pin: each later column reduces by the first (so its pair, if any,
    is its largest row of the symmetric difference)."""
    m = crafted(4, 4)
    R = int((m.dims == 0).sum())
    first = set(m.rows[m.colptr[R]:m.colptr[R + 1]].tolist())
    pairs = dict(map(tuple, O.twist_pairs(m)))
    deaths = {d: b for b, d in pairs.items()}
    for c in range(R + 1, m.n):
        col = set(m.rows[m.colptr[c]:m.colptr[c + 1]].tolist())
        diff = sorted(col ^ first)
        if diff and diff[-1] not in pairs:
            assert deaths.get(c) == diff[-1]
