"""bench.py pieces that run without a GPU: the one-GPU contract, the HBM
traffic read from the committed rocprofv3 table, and the CPU baseline on the
reference's own routines."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
import oracle  # noqa: E402


def test_multi_gpu_request_exits_with_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert r.stdout == ""
    assert "one GPU" in r.stderr


def test_traffic_from_profile_table():
    """The roofline's traffic comes from the committed FETCH_SIZE/WRITE_SIZE
    table, one row per step kernel."""
    for kernel in ("k_force", "k_main"):
        t = bench.pmc_traffic(kernel)
        assert isinstance(t, int) and t > 10**6
    assert bench.pmc_traffic("no_such_kernel") is None


@pytest.mark.skipif(not oracle.RefLib.available(), reason="oracle/_ref (the reference's routines) not built")
def test_reference_cpu_baseline_is_quiet(capfd):
    """The reference writes a warning line per force clamp; the baseline keeps
    them off standard output, which carries the bench's one JSON line."""
    systems, coords, _ = bench.build_workload(nchains=6, nrep=2)
    r = bench.cpu_baseline_reference(systems, coords, budget_s=1.0)
    out, _ = capfd.readouterr()
    assert out == ""
    assert r["kind"] == "reference" and r["cores"] == 1 and r["unit"] == "bead-steps/s"
    assert r["value"] > 0
