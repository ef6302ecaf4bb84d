"""ctypes bindings of include/cospan_hip.h and include/cospan_workloads.h (no fallback: a missing library raises)."""
import ctypes as C
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")

_S = C.c_size_t
_P = C.c_void_p
_I = C.c_int
_U32 = C.c_uint32
_F = C.c_float


class CsError(RuntimeError):
    """A non-zero status from the C boundary, with its message."""

    def __init__(self, status, message):
        super().__init__(f"status {status}: {message}")
        self.status = status
        self.message = message


def _load(name):
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built: run __graft_entry__.build() (make -C cospanmdao_amd/csrc)")
    return C.CDLL(path, mode=C.RTLD_GLOBAL)


# (name, restype, argtypes) of every entry point of include/cospan_hip.h
HIP_API = [
    ("cs_hip_last_error", _I, [C.c_char_p, _S]),
    ("cs_hip_build_id", C.c_char_p, []),
    ("cs_hip_init", _I, [_I]),
    ("cs_hip_device_count", _I, [C.POINTER(_I)]),
    ("cs_hip_device_info", _I, [_I, C.c_char_p, _S, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
    ("cs_hip_device_synchronize", _I, []),
    ("cs_hip_stream_create", _I, [C.POINTER(_P)]),
    ("cs_hip_stream_destroy", _I, [_P]),
    ("cs_hip_stream_synchronize", _I, [_P]),
    ("cs_hip_event_create", _I, [C.POINTER(_P), _I]),
    ("cs_hip_event_destroy", _I, [_P]),
    ("cs_hip_event_record", _I, [_P, _P]),
    ("cs_hip_stream_wait_event", _I, [_P, _P]),
    ("cs_hip_event_query", _I, [_P, C.POINTER(_I)]),
    ("cs_hip_event_synchronize", _I, [_P]),
    ("cs_hip_event_elapsed_ms", _I, [C.POINTER(_F), _P, _P]),
    ("cs_hip_malloc", _I, [C.POINTER(_P), _S]),
    ("cs_hip_free", _I, [_P]),
    ("cs_hip_host_malloc", _I, [C.POINTER(_P), _S]),
    ("cs_hip_host_free", _I, [_P]),
    ("cs_hip_memcpy_async", _I, [_P, _P, _S, _P]),
    ("cs_hip_fill_u32_async", _I, [_P, _U32, _S, _P]),
    ("cs_hip_fill_u8_async", _I, [_P, C.c_uint8, _S, _P]),
    ("cs_hip_launch", _I, [_P, C.POINTER(_U32), C.POINTER(_U32), _S, C.POINTER(_P), _P]),
    ("cs_hip_occupancy", _I, [_P, _I, _S, C.POINTER(_I), C.POINTER(_I)]),
    ("cs_hip_graph_begin", _I, [_P]),
    ("cs_hip_graph_end", _I, [_P, C.POINTER(_P)]),
    ("cs_hip_graph_launch", _I, [_P, _P]),
    ("cs_hip_graph_destroy", _I, [_P]),
    ("cs_hip_stream_capturing", _I, [_P, C.POINTER(_I)]),
    ("cs_hip_set_spin_limit", _I, [_U32]),
    ("cs_hip_spin_limit", _U32, []),
    ("cs_hip_fault_word", _I, [C.POINTER(_P)]),
    ("cs_hip_take_fault", _I, [C.POINTER(_U32)]),
    ("cs_hip_scan_workspace_bytes", _S, [_S]),
    ("cs_hip_scan_workspace_init", _I, [_P, _S, _P]),
    ("cs_hip_exclusive_scan_u32", _I, [_P, _P, _S, _P, _P]),
    ("cs_hip_exclusive_scan_u32_static", _I, [_P, _P, _S, _P]),
    ("cs_hip_build_column", _I, [_I, _P, _S, _S, _I, _P, _P, _P, _P, _P]),
    ("cs_hip_fill_hash_f32", _I, [_U32, _P, _S, _S, _S, _P]),
    ("cs_hip_quantize_bf16", _I, [_P, _P, _S, _S, _S, _S, _P]),
    ("cs_hip_quantize_e4m3_rows", _I, [_P, _P, _P, _S, _S, _S, _S, _P]),
]

WORKLOADS_API = [
    ("cs_workloads_build_id", C.c_char_p, []),
    ("cs_gemver_create", _I, [C.POINTER(_P), _I, _I, _S, _S, _F, _F, _U32]),
    ("cs_gemver_set_tile", _I, [_P, _I]),
    ("cs_gemver_set_replay", _I, [_P, _I]),
    ("cs_gemver_run", _I, [_P]),
    ("cs_gemver_wait", _I, [_P]),
    ("cs_gemver_read", _I, [_P, C.c_char_p, _P, _S]),
    ("cs_gemver_time", _I, [_P, _I, C.POINTER(_F)]),
    ("cs_gemver_time_launches", _I, [_P, _I, C.POINTER(_F), _I, C.POINTER(_I)]),
    ("cs_gemver_time_stage", _I, [_P, _I, _I, C.POINTER(_F)]),
    ("cs_gemver_stats", _I, [_P, C.POINTER(C.c_uint64)]),
    ("cs_gemver_destroy", _I, [_P]),
    ("cs_ns_create", _I, [C.POINTER(_P), _I, _I, _S, _S, C.c_double, C.c_double, _U32, _I]),
    ("cs_ns_run", _I, [_P]),
    ("cs_ns_wait", _I, [_P]),
    ("cs_ns_read", _I, [_P, C.c_char_p, _P, _S]),
    ("cs_ns_write", _I, [_P, C.c_char_p, _P, _S]),
    ("cs_ns_time", _I, [_P, _I, C.POINTER(_F)]),
    ("cs_ns_time_launches", _I, [_P, _I, C.POINTER(_F), _I, C.POINTER(_I)]),
    ("cs_ns_set_replay", _I, [_P, _I]),
    ("cs_ns_destroy", _I, [_P]),
    ("cs_chain_create", _I, [C.POINTER(_P), _I, _S, _S, _U32]),
    ("cs_chain_run", _I, [_P]),
    ("cs_chain_wait", _I, [_P]),
    ("cs_chain_read", _I, [_P, C.c_char_p, C.c_void_p, _S]),
    ("cs_chain_time", _I, [_P, _I, C.POINTER(_F)]),
    ("cs_chain_time_launches", _I, [_P, _I, C.POINTER(_F), _I, C.POINTER(_I)]),
    ("cs_chain_stats", _I, [_P, C.POINTER(C.c_uint64)]),
    ("cs_chain_destroy", _I, [_P]),
    ("cs_nsc_create", _I, [C.POINTER(_P), _I, _S, _S, _I, _I, _I]),
    ("cs_nsc_init", _I, [_P, _F]),
    ("cs_nsc_write", _I, [_P, C.c_void_p, _S]),
    ("cs_nsc_run", _I, [_P]),
    ("cs_nsc_wait", _I, [_P]),
    ("cs_nsc_read", _I, [_P, C.c_char_p, C.c_void_p, _S]),
    ("cs_nsc_time", _I, [_P, _I, C.POINTER(_F)]),
    ("cs_nsc_time_launches", _I, [_P, _I, C.POINTER(_F), _I, C.POINTER(_I)]),
    ("cs_nsc_stats", _I, [_P, C.POINTER(C.c_uint64)]),
    ("cs_nsc_destroy", _I, [_P]),
    ("cs_batch_create", _I, [C.POINTER(_P), _U32, _U32, _U32, _F, _F, _U32]),
    ("cs_batch_sizes", _I, [_P, C.POINTER(_U32), _S]),
    ("cs_batch_set_params", _I, [_P, _U32, _U32, _U32]),
    ("cs_batch_get_params", _I, [_P, C.POINTER(_U32)]),
    ("cs_batch_set_mode", _I, [_P, C.c_int]),
    ("cs_batch_time_stage", _I, [_P, _I, _I, C.POINTER(_F)]),
    ("cs_batch_autotune", _I, [_P, _I, C.c_char_p, C.POINTER(_U32)]),
    ("cs_batch_run", _I, [_P]),
    ("cs_batch_wait", _I, [_P]),
    ("cs_batch_read", _I, [_P, C.c_char_p, _P, _S]),
    ("cs_batch_time", _I, [_P, _I, C.POINTER(_F)]),
    ("cs_batch_time_each", _I, [_P, _I, C.POINTER(_F)]),
    ("cs_batch_stats", _I, [_P, C.POINTER(C.c_uint64)]),
    ("cs_batch_destroy", _I, [_P]),
]


def _bind(lib, api):
    for name, res, args in api:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


class _Lazy:
    """Loads a library on first use, so that importing the package never touches the GPU."""

    def __init__(self, name, api, deps=()):
        self._name, self._api, self._deps, self._lib = name, api, deps, None

    def _get(self):
        if self._lib is None:
            for d in self._deps:
                d._get()
            self._lib = _bind(_load(self._name), self._api)
        return self._lib

    def __getattr__(self, item):
        return getattr(self._get(), item)

    @property
    def path(self):
        return os.path.join(LIB_DIR, self._name)


hip = _Lazy("libcospan_hip.so", HIP_API)
workloads = _Lazy("libcospan_workloads.so", WORKLOADS_API, deps=(hip,))


def check(status):
    """Raises CsError for a non-zero status of the C boundary."""
    if status != 0:
        buf = C.create_string_buffer(1024)
        hip.cs_hip_last_error(buf, len(buf))
        raise CsError(status, buf.value.decode(errors="replace"))
    return status
