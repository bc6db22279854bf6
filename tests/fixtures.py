"""Helpers that read the committed golden fixtures (no access to the reference tree)."""
import gzip
import os
import tempfile

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DATA_FIXTURES = ["Alpha_indoor01_0.010.txt", "shuffled_50_1_maxclear.txt", "vr_random_50_16_maxclear.txt"]


def fixture_path(name):
    """Decompress a golden matrix to a temp file and return its path."""
    src = os.path.join(GOLDEN, name + ".gz")
    fd, path = tempfile.mkstemp(suffix="_" + name)
    with os.fdopen(fd, "wb") as fo, gzip.open(src, "rb") as fi:
        fo.write(fi.read())
    return path


def expected_pairs(name):
    with gzip.open(os.path.join(GOLDEN, name + ".pairs.gz"), "rt") as f:
        k = int(f.readline())
        arr = np.loadtxt(f, dtype=np.int64, ndmin=2) if k else np.zeros((0, 2), np.int64)
    assert arr.shape == (k, 2)
    return arr


def kept_edge_cases():
    z = np.load(os.path.join(GOLDEN, "gudhi_kept_edges.npz"))
    off = z["offsets"]
    return [(z["u"][off[i]:off[i + 1]], z["v"][off[i]:off[i + 1]], z["kept"][off[i]:off[i + 1]])
            for i in range(len(off) - 1)]


# ---- the reference's shipped boundary matrices (tests/golden/make_reference_data.py) ----------
REFERENCE_DATA = os.path.join(GOLDEN, "reference_data.npz")


def reference_data_index():
    """[{file, key, n, nnz, pairs, pairs_sha256} | {file, same_as}] for the 31 shipped files."""
    import json
    with np.load(REFERENCE_DATA) as z:
        return json.loads(bytes(z["index"]).decode())


def reference_matrix(key):
    """(colptr int64, rows int32, dims int32) of a stored matrix."""
    import lzma
    with np.load(REFERENCE_DATA) as z:
        raw = lzma.decompress(bytes(z[key]))
    idx = {e["key"]: e for e in reference_data_index() if "key" in e}[key]
    n, nnz = idx["n"], idx["nnz"]
    lens = np.frombuffer(raw, np.uint8, n, 0).astype(np.int64)
    dims = np.frombuffer(raw, np.uint8, n, n).astype(np.int32)
    d = np.frombuffer(raw, np.uint32, nnz, 2 * n).astype(np.int64)
    colptr = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=colptr[1:])
    assert colptr[-1] == nnz
    col = np.repeat(np.arange(n, dtype=np.int64), lens)
    first = np.zeros(nnz, bool)
    first[colptr[:-1][lens > 0]] = True
    # rows: the first entry of a column is (column - d), each next one the previous plus d
    vals = np.where(first, col - d, d)
    seg = np.cumsum(first) - 1  # column run of each entry
    starts = np.flatnonzero(first)
    rows = np.cumsum(vals) - np.concatenate([[0], np.cumsum(vals)[starts[1:] - 1]])[seg] if nnz else vals
    return colptr, rows.astype(np.int32), dims


def pairs_sha256(pairs):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(pairs, dtype=np.int64).tobytes()).hexdigest()
