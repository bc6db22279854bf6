"""Writes the committed golden fixtures (run in the build container, where the reference
tree and oracle/_ref exist).  Inputs are data files the reference ships; expected outputs
come from the oracle (pairs) and from the reference's own GUDHI collapser (kept edges).

  python tests/golden/make_golden.py
"""
import gzip
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
REF = "/root/reference"

from oracle import oracle as O  # noqa: E402
import complexes as C  # noqa: E402

MATRICES = [
    "data/Alpha/Alpha(indoor01,0.010).txt",
    "data/keeping_it_sparse_datasets_with_highclear_data/shuffled_50_1_maxclear.txt",
    "data/keeping_it_sparse_datasets_with_highclear_data/vr_random_50_16_maxclear.txt",
]


def fixture_name(rel):
    return os.path.basename(rel).replace("(", "_").replace(")", "").replace(",", "_")


def write_pairs_gz(path, pairs):
    with gzip.open(path, "wt") as f:
        f.write(f"{len(pairs)}\n")
        for b, d in pairs:
            f.write(f"{b} {d}\n")


def main():
    for f in ["single_triangle.dat", "single_triangle.bin", "single_triangle_persistence_pairs.dat",
              "single_triangle_persistence_pairs.bin"]:
        shutil.copy(os.path.join(REF, "phat-highclear/examples", f), os.path.join(HERE, f))
    for rel in MATRICES:
        name = fixture_name(rel)
        src = os.path.join(REF, rel)
        with open(src, "rb") as fi, gzip.open(os.path.join(HERE, name + ".gz"), "wb") as fo:
            fo.write(fi.read())
        m = O.load_phat(src)
        write_pairs_gz(os.path.join(HERE, name + ".pairs.gz"), O.twist_pairs(m))
        print(name, m.n, flush=True)
    # kept-edge sets of the reference's GUDHI collapser on small VR edge lists
    us, vs, ks, off = [], [], [], [0]
    for seed in range(12):
        n = [30, 60, 90][seed % 3]
        X = C.uniform_points(n, 3, 100 + seed)
        m = C.vr_matrix(X, [0.35, 0.45, 0.55][seed % 3], 1)
        e = np.where(m.dims == 1)[0]
        u = m.rows[m.colptr[e]]
        v = m.rows[m.colptr[e] + 1]
        us.append(u)
        vs.append(v)
        ks.append(O.ref_collapse_edges(u, v))
        off.append(off[-1] + len(e))
    np.savez_compressed(os.path.join(HERE, "gudhi_kept_edges.npz"), u=np.concatenate(us), v=np.concatenate(vs),
                        kept=np.concatenate(ks), offsets=np.array(off, dtype=np.int64))
    print("kept-edge fixtures", len(off) - 1)


if __name__ == "__main__":
    main()
