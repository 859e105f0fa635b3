"""Loads the HIP library and declares the C ABI (include/onion_gen_hip.h).

There is no fallback: if ``libonion_gen_hip.so`` is missing, importing the
engine raises, so nothing can silently run anything but the HIP path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("OG_LIB_PATH") or os.path.join(_HERE, "libonion_gen_hip.so")

# Every entry point include/onion_gen_hip.h declares, with (restype, argtypes).
_c = ctypes
_u8p = _c.POINTER(_c.c_uint8)
SIGNATURES = {
    "og_hip_device_count": (_c.c_int, []),
    "og_hip_device_info": (_c.c_int, [_c.c_int, _c.c_char_p, _c.c_size_t,
                                      _c.POINTER(_c.c_uint64), _c.POINTER(_c.c_int)]),
    "og_hit_room": (_c.c_uint32, [_c.c_uint64, _c.c_uint32]),
    "og_hip_engine_create": (_c.c_int, [_c.c_int, _c.c_char_p, _c.c_uint32, _c.c_void_p,
                                        _c.c_size_t, _c.c_uint32, _c.POINTER(_c.c_void_p),
                                        _c.c_char_p, _c.c_size_t]),
    "og_hip_engine_create_ex": (_c.c_int, [_c.c_int, _c.c_char_p, _c.c_void_p, _c.c_void_p,
                                           _c.POINTER(_c.c_void_p), _c.c_char_p, _c.c_size_t]),
    "og_hip_engine_threads": (_c.c_uint32, [_c.c_void_p]),
    "og_hip_engine_batch": (_c.c_uint64, [_c.c_void_p]),
    "og_hip_engine_advance": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_uint32,
                                         _c.POINTER(_c.c_uint32), _c.c_char_p, _c.c_size_t]),
    "og_hip_engine_last_ms": (_c.c_double, [_c.c_void_p]),
    "og_hip_engine_destroy": (None, [_c.c_void_p]),
    "og_hip_field_op": (_c.c_int, [_c.c_int, _c.c_int, _c.c_char_p, _c.c_char_p, _c.c_char_p,
                                   _c.c_size_t, _c.c_char_p, _c.c_size_t]),
    "og_hip_radix_bench": (_c.c_int, [_c.c_int, _c.c_int, _c.c_uint32, _c.c_uint32,
                                      _c.POINTER(_c.c_double), _c.c_char_p, _c.c_char_p,
                                      _c.c_size_t]),
    "og_hip_issue_bench": (_c.c_int, [_c.c_int, _c.c_int, _c.c_uint32, _c.c_uint32,
                                      _c.POINTER(_c.c_double), _c.POINTER(_c.c_double),
                                      _c.c_char_p, _c.c_size_t]),
    "og_hip_search_noalloc": (_c.c_int, [_c.c_int, _c.c_char_p, _c.c_uint64, _c.c_uint32,
                                         _c.c_uint32, _c.c_void_p, _c.c_void_p, _c.c_uint32,
                                         _c.c_void_p, _c.c_void_p, _c.c_char_p, _c.c_size_t]),
    "og_hip_search_noalloc_bytes": (_c.c_size_t, [_c.c_uint32, _c.c_uint32]),
    "og_hip_search_noalloc_max": (_c.c_size_t, []),
    "og_field_op_host": (_c.c_int, [_c.c_int, _c.c_char_p, _c.c_char_p, _c.c_char_p,
                                    _c.c_size_t]),
    "og_filters_parse": (_c.c_void_p, [_c.c_char_p, _c.c_uint32, _c.c_char_p, _c.c_size_t]),
    "og_filters_free": (None, [_c.c_void_p]),
    "og_filters_count": (_c.c_size_t, [_c.c_void_p]),
    "og_filters_index_words": (_c.c_uint32, [_c.c_void_p, _c.c_void_p, _c.c_size_t]),
    "og_filters_match_key": (_c.c_int, [_c.c_void_p, _c.c_char_p]),
    "og_filters_match_address": (_c.c_int, [_c.c_void_p, _c.c_char_p]),
    "og_filters_describe": (_c.c_size_t, [_c.c_void_p, _c.c_char_p, _c.c_size_t]),
    "og_filters_address_window_host": (_c.c_int, [_c.c_void_p, _c.c_char_p,
                                                  _c.POINTER(_c.c_int), _c.POINTER(_c.c_int)]),
    "og_filters_exact_table_host": (_c.c_int, [_c.c_void_p, _c.c_char_p, _c.POINTER(_c.c_int)]),
    "og_filters_scan_pair_host": (_c.c_int, [_c.c_void_p, _c.c_char_p, _c.c_char_p,
                                             _c.POINTER(_c.c_int), _c.POINTER(_c.c_int)]),
    "og_filters_scan_host": (_c.c_int, [_c.c_void_p, _c.c_char_p, _c.c_int,
                                        _c.POINTER(_c.c_int), _c.POINTER(_c.c_int),
                                        _c.POINTER(_c.c_int)]),
    "og_key_secret_scalar": (None, [_c.c_char_p, _c.c_char_p]),
    "og_key_expanded_secret": (None, [_c.c_char_p, _c.c_char_p]),
    "og_key_start_point": (None, [_c.c_char_p, _c.c_char_p]),
    "og_key_public_from_scalar": (None, [_c.c_char_p, _c.c_char_p]),
    "og_key_expanded_secret_at_offset": (_c.c_int, [_c.c_char_p, _c.c_uint64, _c.c_char_p]),
    "og_key_address_bytes": (None, [_c.c_char_p, _c.c_char_p]),
    "og_key_address": (None, [_c.c_char_p, _c.c_char_p]),
    "og_run_block_seed": (None, [_c.c_char_p, _c.c_uint64, _c.c_char_p]),
    "og_hip_finish_hits": (_c.c_int, [_c.c_int, _c.c_char_p, _c.c_void_p, _c.c_size_t,
                                      _c.c_void_p, _c.c_char_p, _c.c_size_t]),
    "og_write_key": (_c.c_int, [_c.c_char_p, _c.c_char_p, _c.c_char_p, _c.c_char_p,
                                _c.c_size_t]),
}



class TuningShape(_c.Structure):
    _fields_ = [("first", _c.c_uint32), ("second", _c.c_uint32), ("rate", _c.c_double),
                ("at", _c.c_int64), ("rounds", _c.c_uint32), ("radix", _c.c_uint32),
                ("slots", _c.c_uint32)]


SIGNATURES.update({
    "og_hip_engine_room": (_c.c_uint32, [_c.c_void_p]),
    "og_hip_engine_restart": (_c.c_int, [_c.c_void_p, _c.c_char_p, _c.c_char_p, _c.c_size_t]),
    "og_hip_engine_device": (_c.c_int, [_c.c_void_p]),
    "og_hip_run_lease": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_void_p, _c.c_uint32,
                                    _c.c_void_p, _c.c_size_t, _c.POINTER(_c.c_size_t),
                                    _c.POINTER(_c.c_uint64), _c.POINTER(_c.c_uint64),
                                    _c.c_char_p, _c.c_size_t]),
    "og_hip_engine_tune": (_c.c_int, [_c.c_int, _c.c_char_p, _c.c_void_p, _c.c_uint64,
                                      _c.POINTER(TuningShape), _c.c_char_p, _c.c_size_t]),
    "og_hip_engine_bytes": (_c.c_uint64, [_c.c_void_p]),
    "og_tuning_device_key": (_c.c_size_t, [_c.c_char_p, _c.c_char_p, _c.c_uint32, _c.c_char_p,
                                           _c.c_size_t]),
    "og_tuning_set_version": (_c.c_int, [_c.c_char_p]),
    "og_tuning_choose": (_c.c_size_t, [_c.POINTER(_c.c_double), _c.POINTER(_c.c_uint64),
                                       _c.c_size_t, _c.c_uint64, _c.c_double]),
    "og_hip_engine_shape": (None, [_c.c_void_p, _c.POINTER(_c.c_uint32),
                                   _c.POINTER(_c.c_uint32), _c.POINTER(_c.c_uint32)]),
    "og_hip_engine_sign_stage": (_c.c_int, [_c.c_void_p]),
    "og_hip_engine_kernels": (_c.c_int, [_c.c_void_p]),
    "og_tuning_path": (_c.c_int, [_c.c_char_p, _c.c_size_t]),
    "og_tuning_get": (_c.c_int, [_c.c_char_p, _c.POINTER(TuningShape)]),
    "og_tuning_put": (_c.c_int, [_c.c_char_p, _c.POINTER(TuningShape)]),
})

SIGNATURES.update({
    "og_batch_create": (_c.c_int, [_c.c_int, _c.c_void_p, _c.POINTER(_c.c_void_p), _c.c_char_p,
                                   _c.c_size_t]),
    "og_batch_tune": (_c.c_int, [_c.c_int, _c.c_void_p, _c.c_size_t, _c.c_uint64, _c.c_double,
                                 _c.c_void_p, _c.POINTER(_c.c_double), _c.c_char_p,
                                 _c.c_size_t]),
    "og_batch_footprint": (_c.c_size_t, [_c.c_void_p, _c.c_void_p]),
    "og_batch_admit": (_c.c_int, [_c.c_void_p, _c.c_char_p, _c.c_void_p, _c.c_uint64,
                                  _c.POINTER(_c.c_int), _c.c_char_p, _c.c_size_t]),
    "og_batch_set_want": (_c.c_int, [_c.c_void_p, _c.c_int, _c.c_uint64]),
    "og_batch_search_found": (_c.c_uint64, [_c.c_void_p, _c.c_int]),
    "og_batch_search_width": (_c.c_uint32, [_c.c_void_p, _c.c_int]),
    "og_batch_examined": (_c.c_uint64, [_c.c_void_p]),
    "og_batch_retire": (_c.c_int, [_c.c_void_p, _c.c_int]),
    "og_batch_step": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_uint32, _c.POINTER(_c.c_uint32),
                                 _c.c_char_p, _c.c_size_t]),
    "og_batch_search_state": (_c.c_int, [_c.c_void_p, _c.c_int, _c.POINTER(_c.c_uint64),
                                         _c.POINTER(_c.c_uint64), _c.POINTER(_c.c_uint64),
                                         _c.POINTER(_c.c_int)]),
    "og_batch_resident_bytes": (_c.c_uint64, [_c.c_void_p]),
    "og_batch_per_step": (_c.c_uint64, [_c.c_void_p]),
    "og_batch_last_ms": (_c.c_double, [_c.c_void_p]),
    "og_batch_destroy": (None, [_c.c_void_p]),
})

SIGNATURES.update({
    "og_serve_run": (_c.c_int, [_c.c_int, _c.c_void_p, _c.c_void_p, _c.c_size_t, _c.c_double,
                                _c.c_uint32, _c.c_uint64, _c.c_void_p, _c.c_void_p, _c.c_char_p,
                                _c.c_size_t]),
})

_lib = None


def lib():
    """The loaded library; raises OSError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(
                f"{LIB_PATH} is missing: build it with `make -C onion_gen_amd` "
                "or __graft_entry__.build(); there is no CPU fallback")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


class OgError(RuntimeError):
    """An error reported through the C ABI's (status, message) pair."""


def err_buf():
    return ctypes.create_string_buffer(512)
