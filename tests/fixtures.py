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
