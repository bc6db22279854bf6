"""The reference's 31 shipped boundary matrices on the GPU engine: for every stored matrix, the
pairs of the plain reduction, of the cohomology reduction and of the reduction with high-dim
clearing hash to the oracle's golden SHA-256 (tests/golden/reference_data.npz)."""
import numpy as np
import pytest

import fixtures as F

import specseqplusplus_amd as S

pytestmark = pytest.mark.gpu
STORED = [e for e in F.reference_data_index() if "key" in e]
MODES = [dict(), dict(dualize=1), dict(model=1)]


@pytest.fixture(scope="module")
def h():
    hd = S.Handle(0)
    yield hd
    hd.close()


@pytest.mark.parametrize("e", STORED, ids=lambda e: e["file"].split("/")[-1])
def test_reference_matrix_pairs(h, e):
    cp, rows, dims = F.reference_matrix(e["key"])
    m = S.Matrix(cp, rows, dims)
    for kw in MODES:
        pairs, ess = h.reduce(m, S.params(**kw))
        assert len(pairs) == e["pairs"], (e["file"], kw)
        assert F.pairs_sha256(pairs) == e["pairs_sha256"], (e["file"], kw)
        assert 2 * len(pairs) + len(ess) == e["n"]
