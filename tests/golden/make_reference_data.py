"""Writes tests/golden/reference_data.npz from the reference's shipped boundary matrices
(data/Alpha, data/VR, data/keeping_it_sparse_datasets_with_highclear_data: 31 files, 25
distinct; the six Alpha *_maxclear files are byte-identical to their originals and stored once).

Each matrix is stored compactly: column lengths and dimensions (uint8), and the rows coded per
column as (column index - first row), then successive differences (uint32), the three arrays
concatenated and LZMA-compressed.  Beside it: n, nnz and the SHA-256 of the oracle's pairs
(twist reduction, primal, int64 (birth, death) rows in birth order), the golden value the GPU
tests compare against.  Also written: indoor01_28k_points.npz, the point cloud of the
data/Alpha matrices (point-net/indoor01_28k.txt), for the alpha generator's tests.

    python tests/golden/make_reference_data.py [/root/reference]
"""
import glob
import hashlib
import json
import lzma
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402


def encode(m):
    cp = np.asarray(m.colptr, np.int64)
    rows = np.asarray(m.rows, np.int64)
    lens = np.diff(cp)
    assert lens.max(initial=0) < 256 and np.asarray(m.dims).max(initial=0) < 256
    col = np.repeat(np.arange(m.n, dtype=np.int64), lens)
    first = np.zeros(len(rows), bool)
    first[cp[:-1][lens > 0]] = True
    d = np.empty_like(rows)
    d[first] = col[first] - rows[first]
    rest = ~first
    d[rest] = rows[rest] - rows[np.flatnonzero(rest) - 1]
    assert (d >= 0).all()
    raw = lens.astype(np.uint8).tobytes() + np.asarray(m.dims).astype(np.uint8).tobytes() + \
        d.astype(np.uint32).tobytes()
    return lzma.compress(raw, preset=9 | lzma.PRESET_EXTREME)


def pairs_sha(pairs):
    return hashlib.sha256(np.ascontiguousarray(pairs, dtype=np.int64).tobytes()).hexdigest()


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    files = sorted(glob.glob(os.path.join(ref, "data", "*", "*.txt")))
    seen, index, blobs = {}, [], {}
    for f in files:
        raw = open(f, "rb").read()
        name = os.path.relpath(f, os.path.join(ref, "data"))
        key = hashlib.sha256(raw).hexdigest()
        if key in seen:
            index.append({"file": name, "same_as": seen[key]})
            continue
        seen[key] = name
        m = O.load_phat(f)
        pairs = O.twist_pairs(m)
        blob = encode(m)
        k = f"m{len(blobs)}"
        blobs[k] = np.frombuffer(blob, np.uint8)
        index.append({"file": name, "key": k, "n": int(m.n), "nnz": int(len(m.rows)), "pairs": int(len(pairs)),
                      "pairs_sha256": pairs_sha(pairs)})
        print(name, m.n, len(m.rows), len(pairs), len(blob))
    np.savez(os.path.join(HERE, "reference_data.npz"), index=np.frombuffer(json.dumps(index).encode(), np.uint8),
             **blobs)
    # the point cloud the data/Alpha matrices were built from (point-net/indoor01_28k.txt)
    pts = np.loadtxt(os.path.join(ref, "point-net", "indoor01_28k.txt"), dtype=np.float64, ndmin=2)
    np.savez_compressed(os.path.join(HERE, "indoor01_28k_points.npz"), points=pts)


if __name__ == "__main__":
    main()
