"""The C boundary: every entry point declared in include/*.h is exported by the built libraries, and the Python bindings
cover them. No compute call is made (these run without a GPU)."""
import os
import re
import subprocess

import pytest

from cospanmdao_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = {"cospan_hip.h": "libcospan_hip.so", "cospan_workloads.h": "libcospan_workloads.so"}


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|size_t|void|uint32_t|const\s+char\s*\*)\s*(cs_\w+)\s*\(", text, flags=re.M)))


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(_lib.LIB_DIR, lib)], capture_output=True, text=True, check=True)
    return {line.split()[-1] for line in out.stdout.splitlines() if " T " in line}


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_every_declared_entry_point_is_exported(header):
    names = declared(header)
    assert names, header
    missing = [n for n in names if n not in exported(HEADERS[header])]
    assert not missing, f"{HEADERS[header]} lacks {missing}"


def test_bindings_cover_the_headers():
    bound = {n for n, _, _ in _lib.HIP_API} | {n for n, _, _ in _lib.WORKLOADS_API}
    for header in HEADERS:
        missing = [n for n in declared(header) if n not in bound]
        assert not missing, f"no Python binding for {missing}"


def test_libraries_load_without_a_gpu():
    # dlopen + symbol resolution only
    from cospanmdao_amd import hip, workloads

    assert hip.cs_hip_scan_workspace_bytes(0) >= 16
    assert hip.cs_hip_scan_workspace_bytes(4096) == 16 + 8 * 2
    assert workloads.cs_gemver_create is not None


def test_device_count_without_a_gpu_is_not_an_error():
    import ctypes as C

    from cospanmdao_amd import check, hip

    n = C.c_int(-1)
    check(hip.cs_hip_device_count(C.byref(n)))
    assert n.value >= 0


def test_libraries_are_built_from_this_tree():
    """Each library carries the hash of the sources it was built from; a library built from other sources (a stale build)
    fails here before any of its kernels could run."""
    from cospanmdao_amd import hip, workloads
    from cospanmdao_amd.source_hash import source_hash

    want = source_hash(ROOT)
    assert hip.cs_hip_build_id().decode() == want, "libcospan_hip.so is stale: run __graft_entry__.build()"
    assert workloads.cs_workloads_build_id().decode() == want, "libcospan_workloads.so is stale: run __graft_entry__.build()"
