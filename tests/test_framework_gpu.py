"""The fused-group machinery through cs::Model on the GPU: tests/framework/fused_groups.cpp (segments, range modes, local
memory, work-group shape errors, bare components, group waits, global barriers, cross-group reads over several forwards),
built by __graft_entry__.build() / make -C tests/framework."""
import os
import subprocess

import pytest

from cospanmdao_amd.source_hash import ROOT, source_hash

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def run_program(name):
    """Runs tests/framework/_build/<name> after checking that it was built from the sources in this tree."""
    exe = os.path.join(HERE, "framework", "_build", name)
    assert os.path.exists(exe), f"tests/framework/_build/{name} is not built (make -C tests/framework)"
    print("running", exe, flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout[-4000:], r.stderr[-2000:], flush=True)
    want = source_hash(ROOT, [os.path.join(HERE, "framework", name + ".cpp")])
    assert r.stdout.startswith(f"build {want}\n"), f"{name} is stale (built from other sources than {want}): make -C tests/framework"
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout[-3000:]


def test_fused_groups_program():
    run_program("fused_groups")


def test_ragged_program():
    """Ragged ports through cs::Model on the device: tests/framework/ragged.cpp (counts_from / offsets_from hooks, declared
    counts, dense sources read as rows, a ragged port in a fused group, the error paths), every value checked."""
    run_program("ragged")


def test_policies_program():
    """Run policies, rearm and depends_on through cs::Model on the device: tests/framework/policies.cpp (the reference's
    test_r04.cpp policy and dearm cases and test_system.cpp's depends_on side channel, over captured and replayed runs)."""
    run_program("policies")


def test_ragged_gemver_program():
    """Batched GEMVERs of different sizes written only with ragged ports and counts_from hooks (tests/framework/
    ragged_gemver.cpp): bare and fused passes, sizes and instance counts changing between forwards, every instance against
    FP64 (rel-L2 <= 1e-5)."""
    run_program("ragged_gemver")


def test_range_split_program():
    """The range_split strategies and launch shapes (tests/framework/range_split.cpp, after the reference's
    tests/test_range_split.cpp:569-667): more than 65535 groups per stage under each strategy, a stage taller than a
    hardware grid's 65535 in y, bare group-range and index-range components of 67600 groups, rank-3 and rank-1 stages, a
    bare component's native launch, the physical group coordinates of each strategy, padded refusing rank 3."""
    run_program("range_split")


def test_getting_started_demo():
    """examples/getting_started.cpp (the reference's getting-started/main.cpp with the three changes of INTEGRATION.md),
    built with INTEGRATION.md's build line, prints y = 2 x + 1 for x = 1, 2, 3, 4 on the GPU."""
    exe = os.path.join(HERE, "framework", "_build", "getting_started")
    assert os.path.exists(exe), "tests/framework/_build/getting_started is not built (make -C tests/framework)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.split() == ["3", "5", "7", "9"], r.stdout + r.stderr
