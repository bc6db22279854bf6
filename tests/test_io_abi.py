"""Host side of libssq without a GPU: loading, writing, error codes, exported symbols,
the Vietoris-Rips builder and the CLI's argument handling."""
import os
import re
import subprocess

import numpy as np
import pytest

import complexes as C
import fixtures as F
from oracle import oracle as O

import specseqplusplus_amd as S
from specseqplusplus_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "specseqplusplus_amd", "bin", "specseq++")


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "ssq.h")).read()
    declared = set(re.findall(r"\b(ssq_[a-z0-9_]+)\s*\(", hdr))
    assert len(declared) >= 18
    lib = _lib.lib()  # loads libssq.so: no compute call is made
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared == set(_lib.SIGNATURES)


@pytest.mark.parametrize("name", F.DATA_FIXTURES)
def test_loader_matches_oracle_loader(name):
    path = F.fixture_path(name)
    a = S.load_phat(path)
    b = O.load_phat(path)
    assert np.array_equal(a.colptr, b.colptr) and np.array_equal(a.rows, b.rows) and np.array_equal(a.dims, b.dims)


def test_ascii_and_binary_agree_and_writer_matches_kat(tmp_path):
    a = S.load_phat(f"{F.GOLDEN}/single_triangle.dat")
    b = S.load_phat(f"{F.GOLDEN}/single_triangle.bin", binary=True)
    assert np.array_equal(a.rows, b.rows) and np.array_equal(a.colptr, b.colptr)
    pairs = np.array([[1, 3], [2, 4], [5, 6]])
    for binary, ext in [(False, "dat"), (True, "bin")]:
        out = tmp_path / f"p.{ext}"
        S.save_pairs(str(out), pairs, binary)
        assert out.read_bytes() == open(f"{F.GOLDEN}/single_triangle_persistence_pairs.{ext}", "rb").read()


@pytest.mark.parametrize("text,status", [
    ("0\n1 0 5\n", _lib.SSQ_EFORMAT),         # index out of range
    ("0\n1 0 x\n", _lib.SSQ_EFORMAT),         # bad token
    ("0\n1 0 -1\n", _lib.SSQ_EFORMAT),        # negative index
])
def test_malformed_ascii_returns_status(tmp_path, text, status):
    p = tmp_path / "bad.txt"
    p.write_text(text)
    with pytest.raises(_lib.SSQError) as e:
        S.load_phat(str(p))
    assert e.value.status == status


def test_missing_file_and_truncated_binary(tmp_path):
    with pytest.raises(_lib.SSQError) as e:
        S.load_phat(str(tmp_path / "nope.txt"))
    assert e.value.status == _lib.SSQ_EIO
    p = tmp_path / "trunc.bin"
    p.write_bytes(np.array([3, 0, 0], dtype=np.int64).tobytes())
    with pytest.raises(_lib.SSQError) as e:
        S.load_phat(str(p), binary=True)
    assert e.value.status == _lib.SSQ_EFORMAT


def test_comments_blank_lines_and_duplicates(tmp_path):
    p = tmp_path / "m.txt"
    p.write_text("# header\n0\n\n0   \n0\n1 1 0 0\n# x\n1 0 2\n1 2 1\n2 3 4 5 5\n")
    m = S.load_phat(str(p))
    assert m.n == 7 and m.rows.tolist() == [0, 1, 0, 2, 1, 2, 3, 4, 5]


@pytest.mark.parametrize("seed,n,r,maxdim", [(0, 30, 0.5, 2), (1, 25, 0.6, 3), (2, 40, 0.35, 2), (3, 12, 2.0, 3)])
def test_vr_builder_matches_enumeration(seed, n, r, maxdim):
    X = np.asarray(C.uniform_points(n, 3, seed), dtype=np.float32)
    a = S.vr_build(X, r, maxdim)
    b = C.vr_matrix(X.astype(np.float64), r, maxdim)
    assert np.array_equal(a.dims, b.dims) and np.array_equal(a.colptr, b.colptr) and np.array_equal(a.rows, b.rows)


def test_cli_help_and_argument_errors(tmp_path):
    r = subprocess.run([CLI, "-h"], capture_output=True, text=True)
    assert r.returncode == 0 and "Usage" in r.stdout
    r = subprocess.run([CLI, "a", "b"], capture_output=True, text=True)
    assert r.returncode == 1
    args = [CLI, "in", "out", "text", "ascii", "0", "0", "0", "102400", "1.05", "65536", "1000", "2000"]
    assert subprocess.run(args, capture_output=True).returncode == 1
    args[3], args[4] = "ascii", "xml"
    assert subprocess.run(args, capture_output=True).returncode == 1
    args[4] = "ascii"
    args[1] = str(tmp_path / "missing.txt")
    assert subprocess.run(args, capture_output=True).returncode == 2


def test_host_collapse_matches_fixtures_and_oracle():
    """The library's parallel host collapse keeps exactly GUDHI's edges: the committed
    kept-edge fixtures, the oracle's sequential restatement, and (where it is built here)
    the reference collapser itself, on graphs dense enough for long cascades."""
    import fixtures as F
    for u, v, kept in F.kept_edge_cases():
        for threads in (1, 4):
            assert np.array_equal(S.collapse_edges(u, v, threads), kept)
    from oracle import oracle as O
    import complexes as C
    for seed in range(6):
        X = C.uniform_points(60 + 25 * seed, 3, 700 + seed)
        m = C.vr_matrix(X, 0.35 + 0.02 * seed, 1)
        e = np.where(m.dims == 1)[0]
        u, v = m.rows[m.colptr[e]], m.rows[m.colptr[e] + 1]
        got = S.collapse_edges(u, v, 3)
        assert np.array_equal(got, O.collapse_edges(u, v)), seed
        if O.ref_available():
            assert np.array_equal(got, O.ref_collapse_edges(u, v)), seed
