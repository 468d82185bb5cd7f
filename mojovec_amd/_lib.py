"""ctypes binding of libmojovec_gpu.so (declarations in include/mojovec_gpu.h).

The library is the only compute path: importing it fails loudly when the
shared object has not been built, and every call that reaches the GPU raises
when no HIP device is present. There is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmojovec_gpu.so")

MVG_OK = 0
MVG_ERR_INVALID = 1
MVG_ERR_DEVICE = 2
MVG_ERR_ALLOC = 3
MVG_ERR_STATE = 4


class MojoVecError(Exception):
    """Raised with the library's message when a native call fails."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


_p = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_f = ctypes.c_float

# name -> (restype, argtypes); pointers travel as c_void_p.
SIGNATURES = {
    "mvg_last_error": (ctypes.c_char_p, []),
    "mvg_version": (ctypes.c_char_p, []),
    "mvg_device_info": (_i, [ctypes.c_char_p, _i, ctypes.POINTER(_i)]),
    "mvg_flat_create": (_i, [_i, _i, ctypes.POINTER(_p)]),
    "mvg_flat_add": (_i, [_p, _p, _i64]),
    "mvg_flat_search": (_i, [_p, _p, _i64, _i, _p, _i64, _p, _p]),
    "mvg_flat_set_search_mode": (_i, [_p, _i, _i]),
    "mvg_flat_distances": (_i, [_p, _p, _i64, _p, _i64, _i, _p]),
    "mvg_flat_pair_distances": (_i, [_p, _p, _p, _i64, _p]),
    "mvg_fp8_encode": (_i, [_p, _i64, _i, _p, _p, _p]),
    "mvg_flat_ntotal": (_i64, [_p]),
    "mvg_flat_get_vector": (_i, [_p, _i64, _p]),
    "mvg_flat_get_vectors": (_i, [_p, _p, _i64, _p]),
    "mvg_flat_destroy": (None, [_p]),
    "mvg_sq8_create": (_i, [_i, _i, ctypes.POINTER(_p)]),
    "mvg_sq8_add": (_i, [_p, _p, _i64]),
    "mvg_sq8_search": (_i, [_p, _p, _i64, _i, _i, _p, _i64, _p, _p]),
    "mvg_sq8_distances": (_i, [_p, _p, _i64, _p, _i64, _p]),
    "mvg_sq8_pair_distances": (_i, [_p, _p, _p, _i64, _p]),
    "mvg_sq8_params": (_i, [_p, ctypes.POINTER(_f), ctypes.POINTER(_f), ctypes.POINTER(_f),
                            ctypes.POINTER(_i64)]),
    "mvg_sq8_get_codes": (_i, [_p, _p, _p]),
    "mvg_sq8_get_vector": (_i, [_p, _i64, _p]),
    "mvg_sq8_get_vectors": (_i, [_p, _p, _i64, _p]),
    "mvg_sq8_set_search_mode": (_i, [_p, _i]),
    "mvg_sq8_destroy": (None, [_p]),
    "mvg_kmeans_train": (_i, [_p, _i64, _i, _i, _i, _p, _i64, _u64, _p, _p]),
    "mvg_pq_train": (_i, [_p, _i64, _i, _i, _i, _u64, _p]),
    "mvg_pq_encode": (_i, [_p, _i64, _i, _i, _p, _p]),
    "mvg_pq_lut": (_i, [_p, _i64, _i, _i, _p, _i, _p]),
    "mvg_pq_decode": (_i, [_p, _i64, _i, _i, _p, _p]),
    "mvg_ivfpq_create": (_i, [_i, _i, _i, _i, ctypes.POINTER(_p)]),
    "mvg_ivfpq_train": (_i, [_p, _p, _i64, _i, _i, _u64]),
    "mvg_ivfpq_set_codebooks": (_i, [_p, _p, _p]),
    "mvg_ivfpq_get_codebooks": (_i, [_p, _p, _p]),
    "mvg_ivfpq_add": (_i, [_p, _p, _i64]),
    "mvg_ivfpq_add_with_ids": (_i, [_p, _p, _i64, _p]),
    "mvg_ivfpq_set_keep_vectors": (_i, [_p, _i]),
    "mvg_ivfpq_set_nprobe": (_i, [_p, _i]),
    "mvg_ivfpq_search": (_i, [_p, _p, _i64, _i, _p, _i, _p, _i64, _p, _p]),
    "mvg_ivfpq_search_device": (_i, [_p, _p, _i64, _i, _p, _i, _i, _p, _i64, _p, _p, _p]),
    "mvg_ivfpq_search_device_k": (_i, [_p, _p, _i64, _p, _i, _p, _i, _i, _p, _i64, _p, _p, _p]),
    "mvg_ivfpq_search_device_noalloc": (_i, [_p, _p, _i64, _i, _p, _i, _i, _p, _i64, _p, _p, _p]),
    "mvg_ivfpq_set_search_mode": (_i, [_p, _i, _i]),
    "mvg_ivfpq_set_prefilter": (_i, [_p, _i]),
    "mvg_ivfpq_set_table_bits": (_i, [_p, _i]),
    "mvg_ivfpq_set_work_queue": (_i, [_p, _i]),
    "mvg_ivfpq_set_deep_select": (_i, [_p, _i, _i]),
    "mvg_ivfpq_set_tile_min_batch": (_i, [_p, _i64]),
    "mvg_ivfpq_set_profiling": (_i, [_p, _i]),
    "mvg_ivfpq_read_profile": (_i, [_p, _p, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "mvg_ivfpq_info": (_i, [_p, ctypes.POINTER(_i64), ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "mvg_ivfpq_shape": (_i, [_p, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i),
                             ctypes.POINTER(_i)]),
    "mvg_ivfpq_list_sizes": (_i, [_p, _p]),
    "mvg_ivfpq_get_list": (_i, [_p, _i, _p, _p]),
    "mvg_ivfpq_destroy": (None, [_p]),
    "mvg_ivfpq_capture": (_i, [_p, _p, _i64, _i, _p, _i, _i, _p, _p, ctypes.POINTER(_p)]),
    "mvg_ivfpq_capture_k": (_i, [_p, _p, _i64, _p, _i, _p, _i, _i, _p, _p, ctypes.POINTER(_p)]),
    "mvg_graph_launch": (_i, [_p, _p]),
    "mvg_ivfpq_autotune": (_i, [_p, _p, _i64, _i, _p, _i, _i, _p, _p, ctypes.POINTER(_i),
                                ctypes.POINTER(_f)]),
    "mvg_graph_destroy": (None, [_p]),
    "mvg_server_create": (_i, [_p, _i, _i, _i64, ctypes.POINTER(_p)]),
    "mvg_server_create_sq8": (_i, [_p, _i, _i, _i64, ctypes.POINTER(_p)]),
    "mvg_server_create_multi": (_i, [_p, _p, _i, _i, _i64, ctypes.POINTER(_p)]),
    "mvg_server_submit": (_i, [_p, _p, _i64, _i, _i, _i, _p, _p, ctypes.POINTER(_i64)]),
    "mvg_server_submit_to": (_i, [_p, _i, _p, _i64, _i, _i, _i, _p, _p, ctypes.POINTER(_i64)]),
    "mvg_server_wait": (_i, [_p, _i64, ctypes.POINTER(ctypes.c_double)]),
    "mvg_server_stats": (_i, [_p, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "mvg_server_destroy": (None, [_p]),
    "mvg_coll_create": (_i, [_i, _i, _i, _i, ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(_p)]),
    "mvg_coll_train": (_i, [_p, _p, _i64]),
    "mvg_coll_add": (_i, [_p, _p, _i64, _p, _i64]),
    "mvg_coll_query": (_i, [_p, _p, _i64, _i, _i, _p, _p, ctypes.POINTER(_i64)]),
    "mvg_coll_set_nprobe": (_i, [_p, _i]),
    "mvg_coll_set_train_params": (_i, [_p, _i, _i, _u64]),
    "mvg_coll_info": (_i, [_p, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i64),
                           ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i),
                           ctypes.POINTER(_i)]),
    "mvg_coll_name": (_i, [_p, ctypes.c_char_p, _i]),
    "mvg_coll_index": (_p, [_p]),
    "mvg_coll_destroy": (None, [_p]),
    "mvg_coll_save": (_i, [_p, ctypes.c_char_p]),
    "mvg_coll_load": (_i, [ctypes.c_char_p, ctypes.POINTER(_p)]),
    "mvg_ivfpq_save": (_i, [_p, ctypes.c_char_p]),
    "mvg_ivfpq_load": (_i, [ctypes.c_char_p, ctypes.POINTER(_p)]),
    "mvg_ivfflat_create": (_i, [_i, _i, _i, ctypes.POINTER(_p)]),
    "mvg_ivfflat_create_with_quantizer": (_i, [_i, _i, _i, _i, ctypes.POINTER(_p)]),
    "mvg_ivfflat_train": (_i, [_p, _p, _i64, _i, _u64]),
    "mvg_ivfflat_set_centroids": (_i, [_p, _p]),
    "mvg_ivfflat_get_centroids": (_i, [_p, _p]),
    "mvg_ivfflat_add": (_i, [_p, _p, _i64, _p]),
    "mvg_ivfflat_set_nprobe": (_i, [_p, _i]),
    "mvg_ivfflat_search": (_i, [_p, _p, _i64, _i, _p, _i64, _p, _p]),
    "mvg_ivfflat_search_device": (_i, [_p, _p, _i64, _i, _p, _i64, _p, _p, _p]),
    "mvg_ivfflat_info": (_i, [_p, ctypes.POINTER(_i64), ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "mvg_ivfflat_shape": (_i, [_p, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i),
                               ctypes.POINTER(_i)]),
    "mvg_ivfflat_list_sizes": (_i, [_p, _p]),
    "mvg_ivfflat_get_list": (_i, [_p, _i, _p, _p]),
    "mvg_ivfflat_destroy": (None, [_p]),
    "mvg_ivfflat_save": (_i, [_p, ctypes.c_char_p]),
    "mvg_ivfflat_load": (_i, [ctypes.c_char_p, ctypes.POINTER(_p)]),
    "mvg_snapshot_checksum": (_i, [_p, _i64, ctypes.POINTER(_u64)]),
}

_lib = None


def load():
    """Returns the loaded library; raises ImportError if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python mojovec_amd/build.py`"
            )
        # One HIP runtime per process. PyTorch ships its own libamdhip64; when
        # it is loaded first the library binds to it (same soname), whereas
        # loading the library first starts a second runtime next to PyTorch's,
        # and whichever initialises second finds no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc: int) -> None:
    if rc != MVG_OK:
        msg = load().mvg_last_error().decode("utf-8", "replace")
        raise MojoVecError(rc, msg)
