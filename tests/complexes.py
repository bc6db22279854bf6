"""Small filtered complexes for tests (pure numpy, sizes the oracle finishes in seconds).

Order is (filtration, dim, lexicographic vertex ids); filtration of a simplex is its
longest edge (float32), as the Vietoris-Rips configurations specify.
"""
import itertools

import numpy as np

from oracle.oracle import Matrix


def splitmix64(seed):
    """Portable SplitMix64 stream (same constants as the product generator)."""
    M = (1 << 64) - 1
    x = seed & M
    while True:
        x = (x + 0x9E3779B97F4A7C15) & M
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        yield z ^ (z >> 31)


def uniform_points(n, d, seed):
    g = splitmix64(seed)
    return np.array([[(next(g) >> 11) * (1.0 / (1 << 53)) for _ in range(d)] for _ in range(n)])


def vr_simplices(points, r, maxdim):
    """List of (filtration(float32), dim, vertex tuple) for the VR complex, sorted."""
    X = np.asarray(points, dtype=np.float64)
    n = len(X)
    acc = np.zeros((n, n))
    for k in range(X.shape[1]):  # float64 sum in coordinate order, then one float32 rounding
        diff = X[:, None, k] - X[None, :, k]
        acc += diff * diff
    D = np.sqrt(acc).astype(np.float32)
    adj = [set() for _ in range(n)]
    for i in range(n):
        for j in range(i + 1, n):
            if D[i, j] <= np.float32(r):
                adj[i].add(j)
                adj[j].add(i)
    out = [(np.float32(0.0), 0, (i,)) for i in range(n)]
    # enumerate cliques up to maxdim+1 vertices in increasing vertex order
    def extend(simplex, cands):
        if len(simplex) >= 2:
            f = max(D[a, b] for a, b in itertools.combinations(simplex, 2))
            out.append((np.float32(f), len(simplex) - 1, tuple(simplex)))
        if len(simplex) == maxdim + 1:
            return
        for v in sorted(cands):
            if v > simplex[-1]:
                extend(simplex + [v], cands & adj[v])
    for i in range(n):
        extend([i], adj[i])
    out.sort(key=lambda s: (s[0], s[1], s[2]))
    return out


def to_matrix(simplices):
    index = {s[2]: k for k, s in enumerate(simplices)}
    cols, dims = [], []
    for f, d, verts in simplices:
        if d == 0:
            cols.append([])
        else:
            cols.append(sorted(index[tuple(v for v in verts if v != x)] for x in verts))
        dims.append(d)
    return Matrix.from_columns(cols, dims)


def vr_matrix(points, r, maxdim):
    return to_matrix(vr_simplices(points, r, maxdim))


def closure_matrix(top_simplices, shuffle_seed=None):
    """Boundary matrix of the closure of the given simplices, ordered by (dim, lex) or,
    with a seed, by a random filtration order compatible with faces-before-cofaces."""
    faces = set()
    for s in top_simplices:
        s = tuple(sorted(s))
        for k in range(1, len(s) + 1):
            for f in itertools.combinations(s, k):
                faces.add(f)
    faces = sorted(faces, key=lambda f: (len(f), f))
    if shuffle_seed is not None:
        rng = np.random.default_rng(shuffle_seed)
        val = {}
        for f in faces:  # a random filtration value >= every facet's value
            base = max((val[tuple(x for x in f if x != v)] for v in f), default=0.0) if len(f) > 1 else 0.0
            val[f] = base + rng.random()
        faces.sort(key=lambda f: (val[f], len(f), f))
    sims = [(0.0, len(f) - 1, f) for f in faces]
    return to_matrix(sims)


def betti_from_pairs(m, pairs):
    n = m.n
    paired = np.zeros(n, dtype=bool)
    if len(pairs):
        paired[pairs[:, 0]] = True
        paired[pairs[:, 1]] = True
    maxd = int(m.dims.max()) if n else -1
    return [int(((m.dims == d) & ~paired).sum()) for d in range(maxd + 1)]


# Triangulations with known Z2 Betti numbers
TORUS_7 = [tuple(sorted(((i) % 7, (i + 1) % 7, (i + 3) % 7))) for i in range(7)] + \
    [tuple(sorted((i % 7, (i + 2) % 7, (i + 3) % 7))) for i in range(7)]
RP2_6 = [(0, 1, 2), (0, 2, 3), (0, 3, 4), (0, 4, 5), (0, 5, 1), (1, 2, 4), (2, 3, 5),
         (3, 4, 1), (4, 5, 2), (5, 1, 3)]
SPHERE_TET = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]
BALL_3 = [(0, 1, 2, 3)]
