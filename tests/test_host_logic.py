"""The models' host logic on the CPU: the real Model / System / Executor code, linked against a CPU implementation of the
C boundary (tests/host/mock_hip.cpp: adversarial stream order, garbage-filled fresh allocations, workspace checks, graph
capture and replay) with CPU emulations of the kernels. Needs hipcc for the HIP-language headers, not a GPU."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HOST = os.path.join(HERE, "host")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build():
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc is not installed")
    r = subprocess.run(["make", "-C", HOST, f"HIPCC={HIPCC}"], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]


def run(name):
    r = subprocess.run([os.path.join(HOST, "_build", name)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout[-4000:] + r.stderr[-2000:]
    return r.stdout


def test_batched_gemver_host_logic():
    """Plan, device scan, unit map and the kernels of both forms wired through the System: shapes, pointers, grids against
    the unit counts, zeroed work-queue state at every start, x bit for bit and w against FP64, over eager runs, a capture,
    replays and parameter changes."""
    build()
    out = run("batch_host")
    assert "capture + replays" in out
    assert "a work-queue wait that gave up is reported by Model::wait" in out


def test_chain_host_logic():
    """The GEMVER -> BiCGK -> GEMVER chain in its six styles: every input has a source (connections into promoted and tied
    aliases reach every member), every shape resolves, and each style issues the launches of its plan, eager and replayed."""
    build()
    out = run("chain_host")
    for style, launches in (("unfused", 11), ("parallel", 11), ("fused", 8), ("onepass", 9), ("naive", 8), ("crafted", 8)):
        assert f"{style}: {launches} launches per chain" in out


def test_ragged_host_logic():
    """Ragged ports on the CPU (tests/framework/ragged.cpp built against the mock boundary): column builds from counts and
    offsets, block offsets in every view, hook placement, declared counts merged at connect, and every error message."""
    build()
    out = run("ragged_host")
    assert "offsets_from" in out and "fused group" in out


def test_nsc_host_logic():
    """The compressible NS residual's three variants: tied atomic accumulators (System::tie_outputs and cs::ties_desc in a
    fused group) are one variable, the reductions read the tile's partials, shapes resolve, launches per evaluation are
    6 / 1 / 4, eager and replayed; a plain tie of outputs without allow_plain is refused."""
    build()
    out = run("nsc_host")
    for name, launches in (("unfused", 6), ("fused", 1), ("launches", 4)):
        assert f"{name}: {launches} launches per evaluation" in out


def test_policies_host_logic():
    """Run policies (run_n, do_once, transient), rearm, released memory and refused views, and depends_on through a side
    channel (tests/framework/policies.cpp against the mock boundary): without the edge the mock's stream order makes the
    reader run first and read the stale value; with it, every eager, captured and replayed run reads the writer's value."""
    build()
    out = run("policies_host")
    assert "depends_on: B reads what A wrote" in out and "transient" in out


def test_shared_write_back_protocol():
    """The shared L2 write-back that publishes a work-group's stores in ticket mode and in the persistent batched queue
    (kernel.hpp write_back_step, the function the device runs), simulated over 200 seeded adversarial interleavings of
    store arrivals, polls, claims and write-back starts and completions on 1 to 4 XCDs: every published store is in
    memory, one write-back is in flight per XCD, nothing is stuck; a planted mistake in the decision is caught."""
    build()
    out = run("writeback_host")
    assert "every published store in memory" in out
    assert "planted mistake caught in 200 of 200" in out


def test_getting_started_demo_on_the_cpu_boundary():
    """examples/getting_started.cpp, the reference's one-component demo, included as it is and run against the CPU
    boundary with its kernel emulated by the user's own compute_forward: it prints y = 2 x + 1 for x = 1, 2, 3, 4."""
    build()
    r = subprocess.run([os.path.join(HOST, "_build", "getting_started_host")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.split() == ["3", "5", "7", "9"], r.stdout + r.stderr
