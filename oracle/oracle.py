"""ctypes wrapper around oracle/liboracle.so (and oracle/_ref/libgudhi_ref.so if built).

Checker only: see ssq_oracle.h for what each function restates and how it is pinned.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None

_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_dp = ctypes.POINTER(ctypes.c_double)


def build(force=False):
    """Compile liboracle.so (and _ref/ when the reference tree is present)."""
    so = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.check_call(["make", "-s", "-C", _HERE, "all"])


def lib():
    global _lib
    if _lib is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        L = ctypes.CDLL(so)
        L.or_load_phat.argtypes = [ctypes.c_char_p, ctypes.c_int, _i64p,
                                   ctypes.POINTER(_i64p), ctypes.POINTER(_i32p), ctypes.POINTER(_i32p)]
        L.or_free.argtypes = [ctypes.c_void_p]
        L.or_save_pairs.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int64, _i64p]
        for f in ("or_dense_standard", "or_rank_pairs"):
            getattr(L, f).argtypes = [ctypes.c_int64, _i64p, _i32p, _u8p, _i64p]
            getattr(L, f).restype = ctypes.c_int64
        for f in ("or_twist_reduce", "or_twist_reduce_dual"):
            getattr(L, f).argtypes = [ctypes.c_int64, _i64p, _i32p, _i32p, _u8p, _i64p, _dp]
            getattr(L, f).restype = ctypes.c_int64
        L.or_collapse_edges.argtypes = [ctypes.c_int64, _i32p, _i32p, _u8p]
        L.or_apply_collapse.argtypes = [ctypes.c_int64, _i64p, _i32p, _i32p, _u8p]
        L.or_apply_collapse.restype = ctypes.c_int64
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(os.path.join(_HERE, "_ref", "libgudhi_ref.so"))


def ref():
    global _ref
    if _ref is None:
        R = ctypes.CDLL(os.path.join(_HERE, "_ref", "libgudhi_ref.so"))
        R.ref_collapse_edges.argtypes = [ctypes.c_int64, _i32p, _i32p, _u8p]
        _ref = R
    return _ref


def _p(a, t):
    return a.ctypes.data_as(t)


class Matrix:
    """CSC boundary matrix: colptr int64[n+1], rows int32[nnz], dims int32[n]."""

    def __init__(self, colptr, rows, dims):
        self.colptr = np.ascontiguousarray(colptr, dtype=np.int64)
        self.rows = np.ascontiguousarray(rows, dtype=np.int32)
        self.dims = np.ascontiguousarray(dims, dtype=np.int32)

    @property
    def n(self):
        return len(self.dims)

    @classmethod
    def from_columns(cls, cols, dims):
        colptr = np.zeros(len(cols) + 1, dtype=np.int64)
        for i, c in enumerate(cols):
            colptr[i + 1] = colptr[i] + len(c)
        rows = np.array([r for c in cols for r in sorted(set(c))], dtype=np.int32) if cols else np.zeros(0, np.int32)
        colptr = np.zeros(len(cols) + 1, dtype=np.int64)
        for i, c in enumerate(cols):
            colptr[i + 1] = colptr[i] + len(set(c))
        return cls(colptr, rows, np.asarray(dims, dtype=np.int32))


def load_phat(path, binary=False):
    L = lib()
    n = ctypes.c_int64()
    cp, rw, dm = _i64p(), _i32p(), _i32p()
    rc = L.or_load_phat(path.encode(), int(binary), ctypes.byref(n), ctypes.byref(cp),
                        ctypes.byref(rw), ctypes.byref(dm))
    if rc:
        raise IOError(f"or_load_phat({path}) failed rc={rc}")
    N = n.value
    colptr = np.ctypeslib.as_array(cp, shape=(N + 1,)).copy()
    nnz = int(colptr[-1])
    rows = np.ctypeslib.as_array(rw, shape=(max(nnz, 1),))[:nnz].copy()
    dims = np.ctypeslib.as_array(dm, shape=(max(N, 1),))[:N].copy()
    for p in (cp, rw, dm):
        L.or_free(ctypes.cast(p, ctypes.c_void_p))
    return Matrix(colptr, rows, dims)


def save_pairs(path, pairs, binary=False):
    pairs = np.ascontiguousarray(pairs, dtype=np.int64).reshape(-1, 2)
    rc = lib().or_save_pairs(path.encode(), int(binary), len(pairs), _p(pairs, _i64p))
    if rc:
        raise IOError(path)


def _skip(m, skip):
    if skip is None:
        return None, None
    s = np.ascontiguousarray(skip, dtype=np.uint8)
    return s, _p(s, _u8p)


def twist_pairs(m, skip=None, stats=False, dual=False):
    """Persistence pairs (np x 2 int64, sorted by birth) by twist reduction; dual=True reduces
    the anti-transpose (cohomology) and maps the pairs back."""
    out = np.zeros((max(m.n, 1), 2), dtype=np.int64)
    st = np.zeros(4, dtype=np.float64)
    s, sp = _skip(m, skip)
    f = lib().or_twist_reduce_dual if dual else lib().or_twist_reduce
    k = f(m.n, _p(m.colptr, _i64p), _p(m.rows, _i32p), _p(m.dims, _i32p), sp, _p(out, _i64p), _p(st, _dp))
    pairs = out[:k].copy()
    if stats:
        return pairs, {"additions": float(st[0]), "bytes": float(st[1]), "max_chain": float(st[2]),
                       "depth": float(st[3])}
    return pairs


def dense_pairs(m, skip=None):
    out = np.zeros((max(m.n, 1), 2), dtype=np.int64)
    s, sp = _skip(m, skip)
    k = lib().or_dense_standard(m.n, _p(m.colptr, _i64p), _p(m.rows, _i32p), sp, _p(out, _i64p))
    if k < 0:
        raise ValueError("dense_pairs: matrix too large")
    return out[:k].copy()


def rank_pairs(m, skip=None):
    out = np.zeros((max(m.n, 1), 2), dtype=np.int64)
    s, sp = _skip(m, skip)
    k = lib().or_rank_pairs(m.n, _p(m.colptr, _i64p), _p(m.rows, _i32p), sp, _p(out, _i64p))
    if k < 0:
        raise ValueError("rank_pairs: matrix too large")
    return out[:k].copy()


def collapse_edges(u, v):
    u = np.ascontiguousarray(u, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.int32)
    kept = np.zeros(len(u), dtype=np.uint8)
    lib().or_collapse_edges(len(u), _p(u, _i32p), _p(v, _i32p), _p(kept, _u8p))
    return kept


def ref_collapse_edges(u, v):
    u = np.ascontiguousarray(u, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.int32)
    kept = np.zeros(len(u), dtype=np.uint8)
    ref().ref_collapse_edges(len(u), _p(u, _i32p), _p(v, _i32p), _p(kept, _u8p))
    return kept


def collapse_skip(m):
    """Model-2 sub-complex as a skip mask (non-kept edges and their cofaces)."""
    skip = np.zeros(m.n, dtype=np.uint8)
    k = lib().or_apply_collapse(m.n, _p(m.colptr, _i64p), _p(m.rows, _i32p), _p(m.dims, _i32p),
                                _p(skip, _u8p))
    if k < 0:
        raise ValueError("edge column without exactly two vertices")
    return skip
