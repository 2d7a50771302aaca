"""ctypes loader of librmetis_hip.so (the in-tree HIP build).

The library is the product: there is no CPU fallback.  Importing works
without a GPU (the C ABI then reports METIS_ERROR for compute calls), but a
missing library raises immediately.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librmetis_hip.so")

i32p = C.POINTER(C.c_int32)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
i64p = C.POINTER(C.c_int64)

MAX_LEVELS = 64


class HipStats(C.Structure):
    _fields_ = [
        ("nlevels", C.c_int32), ("coarsest_n", C.c_int32), ("initial_ms", C.c_float),
        ("level_n", C.c_int32 * MAX_LEVELS), ("level_m", C.c_int32 * MAX_LEVELS),
        ("coarsen_ms", C.c_float * MAX_LEVELS), ("refine_ms", C.c_float * MAX_LEVELS),
        ("passes", C.c_int32 * MAX_LEVELS), ("attempts_per_call", C.c_int32),
    ]


# symbol -> (restype, argtypes); every symbol declared in include/rmetis.h
SIGNATURES = {
    "RMETIS_SetDefaultOptions": (C.c_int32, [i32p]),
    "RMETIS_PartGraphKway": (C.c_int32, [i32p, i32p, i32p, i32p, i32p, i32p, i32p, i32p, f32p, f32p, i32p,
                                         i32p, i32p]),
    "RMETIS_PartGraphRecursive": (C.c_int32, [i32p, i32p, i32p, i32p, i32p, i32p, i32p, i32p, f32p, f32p, i32p,
                                              i32p, i32p]),
    "RMETIS_NodeND": (C.c_int32, [i32p, i32p, i32p, i32p, i32p, i32p, i32p]),
    "rmetis_hip_recursive_device": (C.c_int32, [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                                C.c_void_p, C.c_int32, f32p, f32p, i32p, i32p, C.c_void_p]),
    "rmetis_hip_nodend_device": (C.c_int32, [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, i32p, C.c_void_p,
                                             C.c_void_p]),
    "rmetis_read_graph": (C.c_int32, [C.c_char_p, i32p, i32p, C.POINTER(i32p), C.POINTER(i32p), C.POINTER(i32p),
                                      C.POINTER(i32p), C.POINTER(i32p)]),
    "rmetis_write_graph": (C.c_int32, [C.c_char_p, C.c_int32, C.c_int32, i32p, i32p, i32p, i32p, i32p]),
    "rmetis_free": (None, [C.c_void_p]),
    "rmetis_hip_fm2_refine": (C.c_int32, [C.c_int32, C.c_int32, i32p, i32p, i32p, i32p, i32p, f64p, f64p,
                                          C.c_int32, i32p]),
    "rmetis_hip_kway_device": (C.c_int32, [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_int32, f32p, f32p, i32p, i32p, C.c_void_p,
                                           C.POINTER(HipStats)]),
    "rmetis_hip_kway_device_noalloc": (C.c_int32, [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                                   C.c_void_p, C.c_int32, f32p, f32p, i32p, i32p, C.c_void_p]),
    "rmetis_hip_kway_batch_device": (C.c_int32, [C.c_int32, i32p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                                 C.c_void_p, i32p, f32p, i32p, C.POINTER(C.c_uint64), i32p,
                                                 C.c_void_p, C.POINTER(HipStats)]),
    "rmetis_hip_kway_batch": (C.c_int32, [C.c_int32, i32p, C.c_int32, i32p, i32p, i32p, i32p, i32p, f32p, i32p,
                                          C.POINTER(C.c_uint64), i32p, i32p]),
    "rmetis_hip_server_create": (C.c_void_p, [C.c_int32, C.c_int64, C.c_int64]),
    "rmetis_hip_server_submit": (C.c_int64, [C.c_void_p, C.c_int32, C.c_int32, i32p, i32p, i32p, i32p, C.c_int32,
                                             f32p, i32p, C.c_uint64]),
    "rmetis_hip_server_submit_device": (C.c_int64, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                                    C.c_void_p, C.c_void_p, C.c_int32, f32p, i32p, C.c_uint64,
                                                    C.c_void_p]),
    "rmetis_hip_server_result": (C.c_int32, [C.c_void_p, C.c_int64, C.c_int32, i32p, i32p, f64p, i32p]),
    "rmetis_hip_server_stats": (C.c_int32, [C.c_void_p, C.POINTER(C.c_int64)]),
    "rmetis_hip_server_destroy": (None, [C.c_void_p]),
    "rmetis_hip_edge_cut": (C.c_int32, [C.c_int32, i32p, i32p, i32p, i32p, i32p]),
    "rmetis_hip_partition_weights": (C.c_int32, [C.c_int32, C.c_int32, i32p, i32p, C.c_int32, i32p]),
    "rmetis_hip_gain_table": (C.c_int32, [C.c_int32, i32p, i32p, i32p, i32p, C.c_int32, i32p, i32p, i32p,
                                          i32p]),
    "rmetis_hip_greedy_pass": (C.c_int32, [C.c_int32, C.c_int32, i32p, i32p, i32p, i32p, i32p, C.c_int32,
                                           f64p, f64p, C.c_int32, i32p, i32p, i32p]),
    "rmetis_hip_greedy_refine": (C.c_int32, [C.c_int32, C.c_int32, i32p, i32p, i32p, i32p, i32p, C.c_int32,
                                             f64p, f64p, C.c_int32, i32p]),
    "rmetis_hip_fm_refine": (C.c_int32, [C.c_int32, C.c_int32, i32p, i32p, i32p, i32p, i32p, C.c_int32,
                                         f64p, f64p, C.c_int32, i32p]),
    "rmetis_hip_match": (C.c_int32, [C.c_int32, i32p, i32p, i32p, C.c_int32, C.c_uint64, i32p]),
    "rmetis_hip_match_capped": (C.c_int32, [C.c_int32, C.c_int32, i32p, i32p, i32p, i32p, C.c_int32, C.c_uint64,
                                            i64p, i32p]),
    "rmetis_hip_build_coarse_graph": (C.c_int32, [C.c_int32, C.c_int32, i32p, i32p, i32p, i32p, i32p, i32p,
                                                  i32p, i32p, i32p, i32p, i32p, i32p, i32p, i32p, i32p]),
    "rmetis_hip_validate_csr": (C.c_int32, [C.c_int32, C.c_int32, i32p, C.c_int64, i32p, C.c_int64, i32p,
                                            C.c_int64, i32p, C.c_int64]),
    "rmetis_hip_parse_options": (C.c_int32, [i32p, C.POINTER(C.c_int64)]),
    "rmetis_hip_grow_partition": (C.c_int32, [C.c_int32, C.c_int32, i32p, i32p, i32p, C.c_int32, f64p,
                                              C.c_uint64, i32p]),
    "rmetis_hip_level_thresholds": (C.c_int32, [C.c_int32, C.c_int32, f32p, i32p, f32p, C.c_int32,
                                                C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "rmetis_hip_debug_flags": (None, [C.c_int32]),
    "rmetis_hip_get_tuning": (C.c_int32, [i32p, C.c_int32]),
    "rmetis_hip_set_tuning": (C.c_int32, [i32p, C.c_int32]),
    "rmetis_hip_tuning_source": (C.c_char_p, []),
    "rmetis_hip_device_alloc": (C.c_void_p, [C.c_int64]),
    "rmetis_hip_device_free": (None, [C.c_void_p]),
    "rmetis_hip_memcpy": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32]),
    "rmetis_hip_device_synchronize": (C.c_int32, []),
    "rmetis_hip_profile_enable": (None, [C.c_int32]),
    "rmetis_hip_profile_filter": (None, [C.c_int32, C.c_int64]),
    "rmetis_hip_profile_read": (C.c_int32, [C.c_int32, C.c_int64, C.POINTER(C.c_int64),
                                            C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "rmetis_hip_last_error": (C.c_char_p, []),
    "rmetis_hip_device_count": (C.c_int32, []),
}

_lib = None


def lib():
    """Load the built library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m rmetis_amd.build` "
                              "(the k-way path has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class DeviceArray:
    """A device allocation made through the library's own HIP runtime
    (keeps a single runtime in the process)."""

    def __init__(self, host=None, count=None, dtype=None):
        import numpy as np
        if host is not None:
            host = np.ascontiguousarray(host)
            self.dtype, self.count = host.dtype, host.size
        else:
            self.dtype, self.count = np.dtype(dtype), int(count)
        nbytes = max(self.count * self.dtype.itemsize, 1)
        self.ptr = lib().rmetis_hip_device_alloc(nbytes)
        if not self.ptr:
            raise MemoryError("device allocation failed")
        if host is not None and host.size:
            rc = lib().rmetis_hip_memcpy(self.ptr, host.ctypes.data, host.nbytes, 1)
            if rc != 1:
                raise RuntimeError("host->device copy failed")

    def numpy(self):
        import numpy as np
        out = np.empty(self.count, self.dtype)
        if self.count:
            rc = lib().rmetis_hip_memcpy(out.ctypes.data, self.ptr, out.nbytes, 2)
            if rc != 1:
                raise RuntimeError("device->host copy failed")
        return out

    def free(self):
        if self.ptr:
            lib().rmetis_hip_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def last_error():
    msg = lib().rmetis_hip_last_error()
    return msg.decode() if msg else ""
