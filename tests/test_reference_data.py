"""The reference's 31 shipped boundary matrices (data/Alpha, data/VR, data/keeping_it_sparse_*),
stored in tests/golden/reference_data.npz with the SHA-256 of the oracle's pairs of each.

Here (CPU): the stored encoding decodes to the three matrices also committed as text; the six
Alpha *_maxclear files are the originals byte for byte; the oracle reproduces the golden hashes
(the five 166,750-cell shuffled matrices, about 12 s each, are left to the GPU tests, which
compare against the same hashes); and on the 20,875-cell matrices the twist reduction agrees
with the standard algorithm on dense bit columns (no clearing, no twist) and with the
cohomology reduction."""
import numpy as np
import pytest

import fixtures as F
from oracle import oracle as O

INDEX = F.reference_data_index()
STORED = [e for e in INDEX if "key" in e]


def test_index_covers_every_shipped_file():
    assert len(INDEX) == 31 and len(STORED) == 25
    same = [e for e in INDEX if "same_as" in e]
    assert len(same) == 6
    for e in same:  # Alpha(indoor01,a)_maxclear.txt is Alpha(indoor01,a).txt
        assert e["file"].replace("_maxclear", "") == e["same_as"]


def test_decoding_matches_the_text_fixtures():
    by_file = {e["file"].split("/")[-1]: e for e in STORED}
    names = {"Alpha_indoor01_0.010.txt": "Alpha(indoor01,0.010).txt",
             "shuffled_50_1_maxclear.txt": "shuffled_50_1_maxclear.txt",
             "vr_random_50_16_maxclear.txt": "vr_random_50_16_maxclear.txt"}
    for fx, orig in names.items():
        m = O.load_phat(F.fixture_path(fx))
        cp, rows, dims = F.reference_matrix(by_file[orig]["key"])
        assert np.array_equal(cp, m.colptr) and np.array_equal(rows, m.rows) and np.array_equal(dims, m.dims)
        assert np.array_equal(O.twist_pairs(m), F.expected_pairs(fx))


@pytest.mark.parametrize("e", [e for e in STORED if "shuffled_100" not in e["file"]], ids=lambda e: e["file"])
def test_oracle_reproduces_golden_hash(e):
    cp, rows, dims = F.reference_matrix(e["key"])
    assert (e["n"], e["nnz"]) == (len(dims), len(rows))
    m = O.Matrix(cp, rows, dims)
    pairs = O.twist_pairs(m)
    assert len(pairs) == e["pairs"] and F.pairs_sha256(pairs) == e["pairs_sha256"]
    if e["n"] <= 25000:
        assert np.array_equal(O.dense_pairs(m), pairs)
    if e["n"] <= 250000:
        assert np.array_equal(O.twist_pairs(m, dual=True), pairs)
