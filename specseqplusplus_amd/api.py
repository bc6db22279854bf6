"""Python host API over libssq: matrices, reduction handles, batches and the server."""
import ctypes

import numpy as np

from . import _lib as L


class Matrix:
    """CSC boundary matrix: colptr int64[n+1], rows int32[nnz] (sorted per column), dims int32[n]."""

    def __init__(self, colptr, rows, dims):
        self.colptr = np.ascontiguousarray(colptr, dtype=np.int64)
        self.rows = np.ascontiguousarray(rows, dtype=np.int32)
        self.dims = np.ascontiguousarray(dims, dtype=np.int32)

    @property
    def n(self):
        return int(len(self.dims))

    @property
    def nnz(self):
        return int(len(self.rows))

    def c_struct(self):
        m = L.ssq_matrix()
        m.n = self.n
        m.nnz = self.nnz
        m.col_ptr = self.colptr.ctypes.data_as(L.i64p)
        m.rows = self.rows.ctypes.data_as(L.i32p) if self.nnz else ctypes.cast(None, L.i32p)
        m.dims = self.dims.ctypes.data_as(L.i32p)
        return m

    @classmethod
    def from_c(cls, m):
        n, nnz = m.n, m.nnz
        colptr = np.ctypeslib.as_array(m.col_ptr, shape=(n + 1,)).copy()
        rows = np.ctypeslib.as_array(m.rows, shape=(max(nnz, 1),))[:nnz].copy()
        dims = np.ctypeslib.as_array(m.dims, shape=(max(n, 1),))[:n].copy()
        return cls(colptr, rows, dims)


def params(dualize=0, model=0, max_iteration=10, ss_block_size=102400, growth=1.05, queue_size=65536,
           level1=1000, level2=2000, add_waves=0, long_grid=0, short_max=0, rounds_per_sync=0):
    """ssq_params: the reference's eight parameters, then the additions' workgroup shape (0 =
    the engine's default)."""
    p = L.ssq_params()
    p.dualize, p.model, p.max_iteration, p.ss_block_size = dualize, model, max_iteration, ss_block_size
    p.growth, p.queue_size, p.level1, p.level2 = growth, queue_size, level1, level2
    p.add_waves, p.long_grid, p.short_max, p.rounds_per_sync = add_waves, long_grid, short_max, rounds_per_sync
    return p


def _pairs_from_c(pr):
    np_ = pr.np
    pairs = np.ctypeslib.as_array(pr.pairs, shape=(max(2 * np_, 1),))[:2 * np_].copy().reshape(-1, 2)
    ess = np.ctypeslib.as_array(pr.essential, shape=(max(pr.ne, 1),))[:pr.ne].copy()
    L.lib().ssq_pairs_free(ctypes.byref(pr))
    return pairs, ess


def load_phat(path, binary=False):
    m = L.ssq_matrix()
    L.check(L.lib().ssq_load_phat(path.encode(), int(binary), ctypes.byref(m)), path)
    out = Matrix.from_c(m)
    L.lib().ssq_matrix_free(ctypes.byref(m))
    return out


def save_pairs(path, pairs, binary=False):
    pairs = np.ascontiguousarray(pairs, dtype=np.int64).reshape(-1)
    pr = L.ssq_pairs()
    pr.np = len(pairs) // 2
    pr.pairs = pairs.ctypes.data_as(L.i64p)
    pr.ne = 0
    L.check(L.lib().ssq_save_pairs(path.encode(), int(binary), ctypes.byref(pr)), path)


def vr_build(points, r, maxdim):
    pts = np.ascontiguousarray(points, dtype=np.float32)
    m = L.ssq_matrix()
    L.check(L.lib().ssq_vr_build(pts.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), pts.shape[0], pts.shape[1],
                                 float(r), int(maxdim), ctypes.byref(m)))
    out = Matrix.from_c(m)
    L.lib().ssq_matrix_free(ctypes.byref(m))
    return out


def collapse_edges(u, v, threads=0):
    """Kept-edge mask of the flag-complex edge collapse (edges in filtration order)."""
    u = np.ascontiguousarray(u, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.int32)
    kept = np.zeros(len(u), dtype=np.uint8)
    L.check(L.lib().ssq_collapse_edges(len(u), u.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                       v.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                       kept.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), int(threads)))
    return kept


class Handle:
    """One engine (HIP stream + device state) on one GPU."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        L.check(L.lib().ssq_create(int(device), ctypes.byref(self._h)), "ssq_create")

    def close(self):
        if self._h:
            L.lib().ssq_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def error(self):
        return L.lib().ssq_last_error(self._h).decode()

    def reduce(self, m, p=None, stats=False):
        p = p or params()
        pr = L.ssq_pairs()
        st = L.ssq_stats()
        ms = m.c_struct()
        L.check(L.lib().ssq_reduce(self._h, ctypes.byref(ms), ctypes.byref(p), ctypes.byref(pr), ctypes.byref(st)),
                self.error())
        pairs, ess = _pairs_from_c(pr)
        return (pairs, ess, st.as_dict()) if stats else (pairs, ess)

    def init(self, m, p=None):
        """Phase one (ssq_init): upload and prepare m; returns the stats (t_init_ms)."""
        p = p or params()
        st = L.ssq_stats()
        ms = m.c_struct()
        L.check(L.lib().ssq_init(self._h, ctypes.byref(ms), ctypes.byref(p), ctypes.byref(st)), self.error())
        return st.as_dict()

    def compute(self, stats=False):
        """Phase two (ssq_compute): the page loop on the matrix init() left; (pairs, essential)."""
        pr = L.ssq_pairs()
        st = L.ssq_stats()
        L.check(L.lib().ssq_compute(self._h, ctypes.byref(pr), ctypes.byref(st)), self.error())
        pairs, ess = _pairs_from_c(pr)
        return (pairs, ess, st.as_dict()) if stats else (pairs, ess)

    def reduce_device(self, n, nnz, d_colptr, d_rows, d_dims, d_pairs, p=None):
        """Device-resident reduction: arguments are device pointers (ints)."""
        p = p or params()
        np_ = ctypes.c_int64()
        st = L.ssq_stats()
        L.check(L.lib().ssq_reduce_device(self._h, n, nnz, d_colptr, d_rows, d_dims, ctypes.byref(p), d_pairs,
                                          ctypes.byref(np_), ctypes.byref(st)), self.error())
        return np_.value, st.as_dict()

    def vr_build_gpu(self, points, r, maxdim=2):
        """Vietoris-Rips matrix built on this handle's GPU (same cells and order as vr_build)."""
        pts = np.ascontiguousarray(points, dtype=np.float32)
        m = L.ssq_matrix()
        L.check(L.lib().ssq_vr_build_gpu(self._h, pts.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), pts.shape[0],
                                         pts.shape[1], float(r), int(maxdim), ctypes.byref(m)), self.error())
        out = Matrix.from_c(m)
        L.lib().ssq_matrix_free(ctypes.byref(m))
        return out

    # ---- device-resident buffers (HBM), owned by the caller through this handle ----
    def batch_reduce_points(self, clouds, maxdim=2, p=None, stats=False):
        """Pairs of many point clouds' Rips complexes, built and reduced as one batch on the
        GPU.  clouds: list of (points (npts x pdim), r), all with the same pdim."""
        p = p or params()
        arrs = [np.ascontiguousarray(X, dtype=np.float32) for X, _ in clouds]
        pdim = arrs[0].shape[1] if arrs else 1
        if any(a.ndim != 2 or a.shape[1] != pdim for a in arrs):
            raise ValueError("all clouds need the same dimension")
        n = len(arrs)
        fp = ctypes.POINTER(ctypes.c_float)
        ptrs = (fp * n)(*[a.ctypes.data_as(fp) for a in arrs])
        npts = (ctypes.c_int64 * n)(*[a.shape[0] for a in arrs])
        rad = (ctypes.c_float * n)(*[float(r) for _, r in clouds])
        outs = (L.ssq_pairs * n)()
        st = L.ssq_stats()
        L.check(L.lib().ssq_batch_reduce_points(self._h, n, ptrs, npts, pdim, rad, int(maxdim), ctypes.byref(p), outs,
                                                ctypes.byref(st)), self.error())
        res = [_pairs_from_c(outs[i]) for i in range(n)]
        return (res, st.as_dict()) if stats else res

    def rips_count(self, d_points, pbase, radius, pdim=3):
        """Exact edge and triangle counts of each cloud (points resident: d_points a
        DeviceArray of float32, cloud c = rows pbase[c]:pbase[c+1])."""
        pb = np.ascontiguousarray(pbase, dtype=np.int64)
        rad = np.ascontiguousarray(radius, dtype=np.float32)
        c = len(rad)
        e = np.zeros(c, np.int64)
        t = np.zeros(c, np.int64)
        L.check(L.lib().ssq_rips_count(self._h, c, d_points.ptr, pb.ctypes.data_as(L.i64p), int(pdim),
                                       rad.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), e.ctypes.data_as(L.i64p),
                                       t.ctypes.data_as(L.i64p)), self.error())
        return e, t

    def batch_reduce_points_device(self, d_points, pbase, radius, d_pairs, pairs_cap, p=None, pdim=3, stats=False):
        """Batched Rips cohomology with points and pairs in HBM: returns the per-cloud pair
        offsets (cloud c's pairs are d_pairs[2*off[c] : 2*off[c+1]])."""
        p = p or params(dualize=1, ss_block_size=1 << 30)
        pb = np.ascontiguousarray(pbase, dtype=np.int64)
        rad = np.ascontiguousarray(radius, dtype=np.float32)
        c = len(rad)
        off = np.zeros(c + 1, np.int64)
        st = L.ssq_stats()
        L.check(L.lib().ssq_batch_reduce_points_device(self._h, c, d_points.ptr, pb.ctypes.data_as(L.i64p), int(pdim),
                                                       rad.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 2,
                                                       ctypes.byref(p), d_pairs.ptr, int(pairs_cap),
                                                       off.ctypes.data_as(L.i64p), ctypes.byref(st)), self.error())
        return (off, st.as_dict()) if stats else off

    def device_array(self, arr):
        """Copy a numpy array into a new device buffer; returns a DeviceArray."""
        arr = np.ascontiguousarray(arr)
        d = DeviceArray(self, arr.nbytes, arr.dtype, arr.shape)
        L.check(L.lib().ssq_copy_to_device(self._h, d.ptr, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes),
                self.error())
        return d

    def device_empty(self, shape, dtype):
        dt = np.dtype(dtype)
        return DeviceArray(self, int(np.prod(shape)) * dt.itemsize, dt, shape)

    def synchronize(self):
        L.check(L.lib().ssq_device_synchronize(self._h), self.error())

    def upload(self, m):
        """Matrix resident in HBM: (n, nnz, colptr, rows, dims) device arrays."""
        return DeviceMatrix(m.n, m.nnz, self.device_array(m.colptr), self.device_array(m.rows if m.nnz else
                            np.zeros(1, np.int32)), self.device_array(m.dims if m.n else np.zeros(1, np.int32)))

    def reduce_resident(self, dm, d_pairs, p=None):
        """ssq_reduce_device on an uploaded matrix; returns (n_pairs, stats)."""
        return self.reduce_device(dm.n, dm.nnz, dm.colptr.ptr, dm.rows.ptr, dm.dims.ptr, d_pairs.ptr, p)

    def to_device(self, m):
        """ssq_matrix_to_device: m copied into device memory of its own, checked on the device.
        Freed by free_device_matrix (or when the handle closes its process)."""
        dm = L.ssq_device_matrix()
        ms = m.c_struct()
        L.check(L.lib().ssq_matrix_to_device(self._h, ctypes.byref(ms), ctypes.byref(dm)), self.error())
        return dm

    def free_device_matrix(self, dm):
        L.check(L.lib().ssq_device_matrix_free(self._h, ctypes.byref(dm)), self.error())

    def batch_reduce_resident(self, dms, p=None, stats=False):
        """ssq_batch_reduce_resident over matrices from to_device: [(pairs, essential), ...]."""
        p = p or params()
        arr = (L.ssq_device_matrix * len(dms))(*dms)
        outs = (L.ssq_pairs * max(len(dms), 1))()
        st = L.ssq_stats()
        L.check(L.lib().ssq_batch_reduce_resident(self._h, arr, len(dms), ctypes.byref(p), outs, ctypes.byref(st)),
                self.error())
        res = [_pairs_from_c(outs[i]) for i in range(len(dms))]
        return (res, st.as_dict()) if stats else res

    def batch_reduce(self, mats, p=None, stats=False):
        p = p or params()
        structs = [m.c_struct() for m in mats]
        arr = (ctypes.POINTER(L.ssq_matrix) * len(mats))(*[ctypes.pointer(s) for s in structs])
        outs = (L.ssq_pairs * len(mats))()
        st = L.ssq_stats()
        L.check(L.lib().ssq_batch_reduce(self._h, arr, len(mats), ctypes.byref(p), outs, ctypes.byref(st)),
                self.error())
        res = [_pairs_from_c(outs[i]) for i in range(len(mats))]
        return (res, st.as_dict()) if stats else res


class DeviceArray:
    """A device (HBM) buffer allocated through libssq; freed when collected."""

    def __init__(self, handle, nbytes, dtype, shape):
        self.handle, self.nbytes, self.dtype, self.shape = handle, int(nbytes), np.dtype(dtype), tuple(np.atleast_1d(shape))
        p = ctypes.c_void_p()
        L.check(L.lib().ssq_device_alloc(handle._h, max(self.nbytes, 1), ctypes.byref(p)), "ssq_device_alloc")
        self.ptr = p.value

    def numpy(self, count=None):
        n = int(np.prod(self.shape)) if count is None else int(count)
        out = np.empty(n, dtype=self.dtype)
        if n:
            L.check(L.lib().ssq_copy_to_host(self.handle._h, out.ctypes.data_as(ctypes.c_void_p), self.ptr,
                                            n * self.dtype.itemsize), "ssq_copy_to_host")
        return out

    def free(self):
        if self.ptr and self.handle._h:
            L.lib().ssq_device_free(self.handle._h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceMatrix:
    def __init__(self, n, nnz, colptr, rows, dims):
        self.n, self.nnz, self.colptr, self.rows, self.dims = n, nnz, colptr, rows, dims


_default = None


def reduce(m, p=None, stats=False):
    global _default
    if _default is None:
        _default = Handle(0)
    return _default.reduce(m, p, stats)
