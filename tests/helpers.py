"""Device memory for the GPU tests through the C boundary itself (cs_hip_malloc / cs_hip_memcpy_async)."""
import ctypes as C

import numpy as np

from cospanmdao_amd import check, hip


class DeviceArray:
    """A device allocation holding a copy of a numpy array (or `nbytes` uninitialised bytes)."""

    def __init__(self, host=None, nbytes=None):
        self.ptr = C.c_void_p()
        self.nbytes = host.nbytes if host is not None else nbytes
        check(hip.cs_hip_malloc(C.byref(self.ptr), max(self.nbytes, 16)))
        if host is not None and self.nbytes:
            host = np.ascontiguousarray(host)
            check(hip.cs_hip_memcpy_async(self.ptr, host.ctypes.data, self.nbytes, None))
            check(hip.cs_hip_stream_synchronize(None))

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        if out.nbytes:
            check(hip.cs_hip_stream_synchronize(None))
            check(hip.cs_hip_memcpy_async(out.ctypes.data, self.ptr, out.nbytes, None))
            check(hip.cs_hip_stream_synchronize(None))
        return out

    def fill_u32(self, value):
        check(hip.cs_hip_fill_u32_async(self.ptr, value, self.nbytes // 4, None))

    def free(self):
        if self.ptr:
            check(hip.cs_hip_free(self.ptr))
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def init_device():
    check(hip.cs_hip_init(0))


def record_errors(key, record):
    """Merges `record` (per-output errors of one full-size parity case) under `key` into $CS_PARITY_OUT/parity_errors.json
    when CS_PARITY_OUT is set; the committed summary profiles/parity_errors_full_size.md is made from that file."""
    import json
    import os

    out = os.environ.get("CS_PARITY_OUT")
    if not out:
        return
    os.makedirs(out, exist_ok=True)
    path = os.path.join(out, "parity_errors.json")
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data[key] = record
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
