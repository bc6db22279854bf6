"""3-D alpha complex boundary matrices (input builder for the alpha configuration).

Delaunay tetrahedra come from scipy's Qhull.  Filtration values are squared alpha radii,
assigned top-down as the alpha filtration does (phat-highclear/addons/alpha_3.cpp builds the
same filtration with CGAL):
  * a tetrahedron gets its squared circumradius;
  * going down one dimension, every face f of a cell s (cofaces visited in lexicographic
    order of their vertex tuples) is updated: if f already has a value it becomes
    min(value(f), value(s)); otherwise, if the vertex of s opposite f lies strictly inside
    f's smallest circumsphere (f is not Gabriel), f takes value(s);
  * a face that still has no value gets its own smallest squared circumradius (vertices 0);
  * values are made non-decreasing from faces to cofaces.
Cells with value <= cutoff are kept, ordered by (value as float64, dimension, lexicographic
vertex ids).
"""
import numpy as np

from .api import Matrix


def _pack(cols, bits=21):
    key = np.zeros(len(cols[0]), dtype=np.int64)
    for c in cols:
        key = (key << bits) | c.astype(np.int64)
    return key


def _circ_tet(P, T):
    a, b, c, d = (P[T[:, k]] for k in range(4))
    ba, ca, da = b - a, c - a, d - a
    lb, lc, ld = (ba * ba).sum(1), (ca * ca).sum(1), (da * da).sum(1)
    cr_cd = np.cross(ca, da)
    cr_db = np.cross(da, ba)
    cr_bc = np.cross(ba, ca)
    den = 2.0 * (ba * cr_cd).sum(1)
    u = (lb[:, None] * cr_cd + lc[:, None] * cr_db + ld[:, None] * cr_bc) / den[:, None]
    return (u * u).sum(1)


def _circ_tri(P, F):
    a, b, c = (P[F[:, k]] for k in range(3))
    ab, ac = b - a, c - a
    n = np.cross(ab, ac)
    nn = (n * n).sum(1)
    u = (np.cross(n, ab) * (ac * ac).sum(1)[:, None] + np.cross(ac, n) * (ab * ab).sum(1)[:, None]) / (2.0 * nn[:, None])
    center = a + u
    return (u * u).sum(1), center


def _circ_edge(P, E):
    a, b = P[E[:, 0]], P[E[:, 1]]
    return ((b - a) ** 2).sum(1) / 4.0, (a + b) / 2.0


def _propagate(val_face, face_of_inc, coface_rank, coface_val, attached):
    """GUDHI's per-face rule over its cofaces visited in coface_rank order."""
    order = np.lexsort((coface_rank, face_of_inc))
    f = face_of_inc[order]
    v = coface_val[order]
    att = attached[order]
    starts = np.flatnonzero(np.r_[True, f[1:] != f[:-1]])
    seg = np.repeat(np.arange(len(starts)), np.diff(np.r_[starts, len(f)]))
    pos = np.arange(len(f))
    first_att = np.full(len(starts), np.iinfo(np.int64).max)
    idx = np.flatnonzero(att)
    np.minimum.at(first_att, seg[idx], pos[idx])
    live = pos >= first_att[seg]
    vm = np.where(live, v, np.inf)
    mins = np.minimum.reduceat(vm, starts)
    out = val_face.copy()
    fs = f[starts]
    has = np.isfinite(mins)
    out[fs[has]] = mins[has]
    return out


def alpha_complex(points, cutoff):
    from scipy.spatial import Delaunay

    P = np.asarray(points, dtype=np.float64)
    n = len(P)
    if n >= (1 << 21):
        raise ValueError("at most 2^21 points")
    T = np.sort(Delaunay(P).simplices.astype(np.int64), axis=1)
    T = T[np.lexsort(T.T[::-1])]
    nt = len(T)
    # faces
    drop = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]
    tri_inc = np.concatenate([T[:, list(k)] for k in drop])      # (4 nt, 3)
    tri_opp = np.concatenate([T[:, i] for i in range(4)])         # apex opposite each face
    tri_cof = np.tile(np.arange(nt), 4)
    tk = _pack([tri_inc[:, 0], tri_inc[:, 1], tri_inc[:, 2]])
    tkeys, tri_of_inc = np.unique(tk, return_inverse=True)
    F = np.stack([(tkeys >> 42) & 0x1FFFFF, (tkeys >> 21) & 0x1FFFFF, tkeys & 0x1FFFFF], 1)
    nf = len(F)
    e_drop = [(1, 2), (0, 2), (0, 1)]
    edge_inc = np.concatenate([F[:, list(k)] for k in e_drop])
    edge_opp = np.concatenate([F[:, i] for i in (0, 1, 2)])
    edge_cof = np.tile(np.arange(nf), 3)
    ek = _pack([edge_inc[:, 0], edge_inc[:, 1]])
    ekeys, edge_of_inc = np.unique(ek, return_inverse=True)
    E = np.stack([(ekeys >> 21) & 0x1FFFFF, ekeys & 0x1FFFFF], 1)
    ne = len(E)
    # values, top-down
    vT = _circ_tet(P, T)
    r2F, cF = _circ_tri(P, F)
    att = ((P[tri_opp] - cF[tri_of_inc]) ** 2).sum(1) < r2F[tri_of_inc]
    vF = _propagate(np.full(nf, np.nan), tri_of_inc, tri_cof, vT[tri_cof], att)
    vF = np.where(np.isnan(vF), r2F, vF)
    r2E, cE = _circ_edge(P, E)
    att = ((P[edge_opp] - cE[edge_of_inc]) ** 2).sum(1) < r2E[edge_of_inc]
    vE = _propagate(np.full(ne, np.nan), edge_of_inc, edge_cof, vF[edge_cof], att)
    vE = np.where(np.isnan(vE), r2E, vE)
    vV = np.zeros(n)
    # non-decreasing from faces to cofaces
    vE = np.maximum(vE, 0.0)
    fmax = np.zeros(nf)
    np.maximum.at(fmax, edge_cof, vE[edge_of_inc])
    vF = np.maximum(vF, fmax)
    tmax = np.zeros(nt)
    np.maximum.at(tmax, tri_cof, vF[tri_of_inc])
    vT = np.maximum(vT, tmax)
    # cut off and order by (value, dim, lex)
    cells = []
    for dim, V, vals in [(0, np.arange(n)[:, None], vV), (1, E, vE), (2, F, vF), (3, T, vT)]:
        keep = vals <= cutoff
        cells.append((dim, V[keep], vals[keep], np.flatnonzero(keep)))
    allv = np.concatenate([c[2] for c in cells])
    alld = np.concatenate([np.full(len(c[2]), c[0]) for c in cells])
    lex = [np.concatenate([c[1][:, k] if c[1].shape[1] > k else np.zeros(len(c[1]), np.int64) for c in cells])
           for k in range(4)]
    order = np.lexsort((lex[3], lex[2], lex[1], lex[0], alld, allv))
    N = len(order)
    pos = np.empty(N, np.int64)
    pos[order] = np.arange(N)
    base = np.cumsum([0] + [len(c[2]) for c in cells])
    # global position of each kept cell of each dimension (by original local index)
    gpos = []
    for k, c in enumerate(cells):
        full = np.full({0: n, 1: ne, 2: nf, 3: nt}[c[0]], -1, np.int64)
        full[c[3]] = pos[base[k]:base[k + 1]]
        gpos.append(full)
    dims = alld[order].astype(np.int32)
    counts = np.array([0, 2, 3, 4])[dims]
    colptr = np.zeros(N + 1, np.int64)
    colptr[1:] = np.cumsum(counts)
    rows = np.empty(colptr[-1], np.int64)
    # edges: vertices
    for dim, faces_of in [(1, lambda idx: [E[idx, 0], E[idx, 1]]),
                          (2, lambda idx: [gpos[1][edge_of_inc[idx + k * nf]] for k in range(3)]),
                          (3, lambda idx: [gpos[2][tri_of_inc[idx + k * nt]] for k in range(4)])]:
        kept_local = cells[dim][3]
        g = gpos[dim][kept_local]
        fl = faces_of(kept_local)
        if dim == 1:
            fl = [gpos[0][x] for x in fl]
        B = np.sort(np.stack(fl, 1), axis=1)
        if (B < 0).any():
            raise RuntimeError("face of a kept cell was cut off")
        for k in range(B.shape[1]):
            rows[colptr[g] + k] = B[:, k]
    return Matrix(colptr, rows.astype(np.int32), dims)


def c3_points(n=100000, seed=0):
    return np.random.default_rng(seed).random((n, 3)) * 100.0
