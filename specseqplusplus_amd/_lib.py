"""ctypes bindings of libssq (include/ssq.h).  The library is required: importing a
function that needs it raises if lib/libssq.so has not been built."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libssq.so")

SSQ_OK, SSQ_EIO, SSQ_EFORMAT, SSQ_ENOMEM, SSQ_EDEVICE, SSQ_EINVAL, SSQ_EBUSY = range(7)
STATUS_NAMES = {0: "SSQ_OK", 1: "SSQ_EIO", 2: "SSQ_EFORMAT", 3: "SSQ_ENOMEM", 4: "SSQ_EDEVICE",
                5: "SSQ_EINVAL", 6: "SSQ_EBUSY"}

i64p = ctypes.POINTER(ctypes.c_int64)
i32p = ctypes.POINTER(ctypes.c_int32)


class ssq_matrix(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("nnz", ctypes.c_int64), ("col_ptr", i64p), ("rows", i32p),
                ("dims", i32p)]


class ssq_params(ctypes.Structure):
    _fields_ = [("dualize", ctypes.c_int32), ("model", ctypes.c_int32), ("max_iteration", ctypes.c_int32),
                ("ss_block_size", ctypes.c_int32), ("growth", ctypes.c_float), ("queue_size", ctypes.c_int32),
                ("level1", ctypes.c_int32), ("level2", ctypes.c_int32), ("add_waves", ctypes.c_int32),
                ("long_grid", ctypes.c_int32), ("short_max", ctypes.c_int32), ("rounds_per_sync", ctypes.c_int32)]


class ssq_pairs(ctypes.Structure):
    _fields_ = [("np", ctypes.c_int64), ("pairs", i64p), ("ne", ctypes.c_int64), ("essential", i64p)]


class ssq_stats(ctypes.Structure):
    _fields_ = [("t_init_ms", ctypes.c_double), ("t_reduce_ms", ctypes.c_double), ("t_total_ms", ctypes.c_double),
                ("t_collapse_ms", ctypes.c_double), ("n_pending", ctypes.c_int64), ("n_apparent", ctypes.c_int64),
                ("n_cleared_hd", ctypes.c_int64), ("n_kept_edges", ctypes.c_int64), ("n_pages", ctypes.c_int64),
                ("n_rounds", ctypes.c_int64), ("n_additions", ctypes.c_int64), ("add_bytes", ctypes.c_double),
                ("store_bytes", ctypes.c_int64), ("t_add_kernel_ms", ctypes.c_double),
                ("n_add_launches", ctypes.c_int64), ("n_tiled_additions", ctypes.c_int64),
                ("n_dense_additions", ctypes.c_int64), ("dense_bytes", ctypes.c_double),
                ("t_add_long_kernel_ms", ctypes.c_double),
                ("max_chain", ctypes.c_int64), ("max_addition_len", ctypes.c_int64),
                ("n_huge_additions", ctypes.c_int64), ("huge_bytes", ctypes.c_double),
                ("t_add_huge_kernel_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ssq_device_matrix(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("nnz", ctypes.c_int64), ("col_ptr", ctypes.c_void_p),
                ("rows", ctypes.c_void_p), ("dims", ctypes.c_void_p)]


class ssq_server_config(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("streams", ctypes.c_int), ("max_batch", ctypes.c_int),
                ("max_batch_cells", ctypes.c_int64), ("mem_budget_bytes", ctypes.c_int64), ("params", ssq_params),
                ("points_params", ssq_params)]


class ssq_server_stats(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_int64 * 3), ("requests", ctypes.c_int64 * 3), ("cells", ctypes.c_int64 * 3),
                ("service_ms", ctypes.c_double * 3), ("max_service_ms", ctypes.c_double * 3)]


# every symbol declared in include/ssq.h: name -> (restype, argtypes)
SIGNATURES = {
    "ssq_load_phat": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ssq_matrix)]),
    "ssq_save_pairs": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ssq_pairs)]),
    "ssq_matrix_free": (None, [ctypes.POINTER(ssq_matrix)]),
    "ssq_pairs_free": (None, [ctypes.POINTER(ssq_pairs)]),
    "ssq_default_params": (None, [ctypes.POINTER(ssq_params)]),
    "ssq_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "ssq_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "ssq_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "ssq_reduce": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ssq_matrix), ctypes.POINTER(ssq_params),
                                  ctypes.POINTER(ssq_pairs), ctypes.POINTER(ssq_stats)]),
    "ssq_init": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ssq_matrix), ctypes.POINTER(ssq_params),
                                ctypes.POINTER(ssq_stats)]),
    "ssq_compute": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ssq_pairs), ctypes.POINTER(ssq_stats)]),
    "ssq_reduce_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ssq_params),
                                         ctypes.c_void_p, i64p, ctypes.POINTER(ssq_stats)]),
    "ssq_batch_reduce": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.POINTER(ssq_matrix)), ctypes.c_int,
                                        ctypes.POINTER(ssq_params), ctypes.POINTER(ssq_pairs),
                                        ctypes.POINTER(ssq_stats)]),
    "ssq_matrix_to_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ssq_matrix),
                                            ctypes.POINTER(ssq_device_matrix)]),
    "ssq_device_matrix_free": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ssq_device_matrix)]),
    "ssq_batch_reduce_resident": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ssq_device_matrix), ctypes.c_int,
                                                 ctypes.POINTER(ssq_params), ctypes.POINTER(ssq_pairs),
                                                 ctypes.POINTER(ssq_stats)]),
    "ssq_server_register_matrix": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ssq_matrix), i64p]),
    "ssq_server_submit_registered": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, i64p]),
    "ssq_server_unregister_matrix": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "ssq_server_create": (ctypes.c_int, [ctypes.POINTER(ssq_server_config), ctypes.POINTER(ctypes.c_void_p)]),
    "ssq_server_submit_matrix": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ssq_matrix), i64p]),
    "ssq_server_submit_points": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_int64,
                                                ctypes.c_int, ctypes.c_float, ctypes.c_int, i64p]),
    "ssq_server_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ssq_pairs),
                                       ctypes.POINTER(ctypes.c_double)]),
    "ssq_server_wait_all": (ctypes.c_int, [ctypes.c_void_p]),
    "ssq_server_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ssq_server_stats)]),
    "ssq_server_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "ssq_device_alloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    "ssq_device_free": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "ssq_copy_to_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    "ssq_copy_to_host": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    "ssq_device_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "ssq_vr_build": (ctypes.c_int, [ctypes.POINTER(ctypes.c_float), ctypes.c_int64, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_int, ctypes.POINTER(ssq_matrix)]),
    "ssq_vr_build_gpu": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_int64, ctypes.c_int,
                                        ctypes.c_float, ctypes.c_int, ctypes.POINTER(ssq_matrix)]),
    "ssq_alpha_build": (ctypes.c_int, [ctypes.POINTER(ctypes.c_double), ctypes.c_int64, ctypes.c_double, ctypes.c_int,
                                       ctypes.POINTER(ssq_matrix), ctypes.POINTER(ctypes.POINTER(ctypes.c_double))]),
    "ssq_free": (None, [ctypes.c_void_p]),
    "ssq_delaunay_3": (ctypes.c_int, [ctypes.POINTER(ctypes.c_double), ctypes.c_int64, ctypes.POINTER(ctypes.c_int32),
                                      ctypes.c_int64, i64p]),
    "ssq_collapse_edges": (ctypes.c_int, [ctypes.c_int64, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                          ctypes.POINTER(ctypes.c_uint8), ctypes.c_int]),
    "ssq_batch_reduce_points": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int,
                                               ctypes.POINTER(ctypes.POINTER(ctypes.c_float)),
                                               ctypes.POINTER(ctypes.c_int64), ctypes.c_int,
                                               ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                               ctypes.POINTER(ssq_params), ctypes.POINTER(ssq_pairs),
                                               ctypes.POINTER(ssq_stats)]),
    "ssq_rips_count": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, i64p, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_float), i64p, i64p]),
    "ssq_batch_reduce_points_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, i64p,
                                                      ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                                      ctypes.POINTER(ssq_params), ctypes.c_void_p, ctypes.c_int64,
                                                      i64p, ctypes.POINTER(ssq_stats)]),
}

_lib = None


def lib():
    """Load libssq.so (raises if it has not been built: there is no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libssq is not built ({LIB_PATH} missing); run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class SSQError(RuntimeError):
    def __init__(self, status, msg=""):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def check(status, msg=""):
    if status != SSQ_OK:
        raise SSQError(status, msg)
