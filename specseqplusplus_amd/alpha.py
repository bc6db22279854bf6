"""3-D alpha complex boundary matrices (ssq_alpha_build, host C++ in libssq).

The Delaunay triangulation uses exact predicates and symbolic perturbation, so degenerate
inputs (collinear scan lines, cospherical grids) triangulate the same way whatever the input
order.  Filtration values follow phat-highclear/addons/alpha_3.cpp:260-352:
  * a tetrahedron takes its squared circumradius;
  * a triangle takes the value of the tetrahedron it is attached to (that tetrahedron's fourth
    vertex strictly inside the triangle's smallest circumsphere), else its own squared
    circumradius;
  * an edge takes the smallest value of the triangles it is attached to (the third vertex
    strictly inside the edge's diametral sphere), else a quarter of its squared length;
  * vertices take 0.
Cells with value <= cutoff are kept, ordered by (value, dimension, vertex ids) as alpha_3.cpp
writes them, or dimension first (dim_major=True) as the reference's data/Alpha files are
stored.  Duplicate points count once; vertex ids follow each distinct point's first occurrence.
"""
import ctypes

import numpy as np

from . import _lib as L
from .api import Matrix


def alpha_complex(points, cutoff, dim_major=False, values=False):
    """Boundary matrix of the alpha complex of `points` (n x 3) with squared radius <= cutoff
    (cutoff < 0: the whole filtration).  values=True also returns each cell's value."""
    P = np.ascontiguousarray(points, dtype=np.float64)
    if P.ndim != 2 or P.shape[1] != 3:
        raise ValueError("points must be n x 3")
    m = L.ssq_matrix()
    vp = ctypes.POINTER(ctypes.c_double)()
    L.check(L.lib().ssq_alpha_build(P.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(P), float(cutoff),
                                    int(bool(dim_major)), ctypes.byref(m), ctypes.byref(vp) if values else None),
            "ssq_alpha_build")
    try:
        M = Matrix.from_c(m)
        vals = np.ctypeslib.as_array(vp, shape=(max(M.n, 1),))[:M.n].copy() if values else None
    finally:
        L.lib().ssq_matrix_free(ctypes.byref(m))
        if values:
            L.lib().ssq_free(vp)
    return (M, vals) if values else M


def delaunay_tets(points):
    """The Delaunay tetrahedra (ssq_delaunay_3): k x 4 ascending vertex ids, ids as in
    alpha_complex."""
    P = np.ascontiguousarray(points, dtype=np.float64)
    n = ctypes.c_int64()
    cap = 8 * len(P) + 64
    while True:
        out = np.empty((cap, 4), np.int32)
        rc = L.lib().ssq_delaunay_3(P.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(P),
                                    out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), cap, ctypes.byref(n))
        if rc == L.SSQ_ENOMEM and n.value > cap:
            cap = n.value
            continue
        L.check(rc, "ssq_delaunay_3")
        return out[:n.value]


def c3_points(n=100000, seed=0):
    return np.random.default_rng(seed).random((n, 3)) * 100.0
