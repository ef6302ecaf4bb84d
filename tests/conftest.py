import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP libraries")


_BUILD_CHECKED = []


@pytest.fixture(autouse=True)
def _gpu_libraries_match_tree(request):
    """Before the first GPU test: each HIP library must be built from the sources in this tree (its build id, the hash of
    those sources). A stale library stops the session before any of its kernels runs."""
    if _BUILD_CHECKED or request.node.get_closest_marker("gpu") is None:
        return
    from cospanmdao_amd import hip, workloads
    from cospanmdao_amd.source_hash import source_hash

    want = source_hash(ROOT)
    got = {"libcospan_hip.so": hip.cs_hip_build_id().decode(), "libcospan_workloads.so": workloads.cs_workloads_build_id().decode()}
    stale = {k: v for k, v in got.items() if v != want}
    if stale:
        pytest.exit(f"stale HIP libraries {stale} (the sources hash to {want}): run __graft_entry__.build()", returncode=3)
    _BUILD_CHECKED.append(True)
