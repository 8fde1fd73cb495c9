"""The host side of bench.py: a request for more than one GPU is refused before anything touches
a device, and the CPU-baseline legs (the oracle on a bounded sample of the workload, pinned
to one core) return the fields the bench line carries."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert r.stdout == ""
    assert "one GPU only" in r.stderr


def test_cpu_baseline_fields():
    import bench
    from avalancheflume3d_amd import workloads
    inst = workloads.small()
    cb = bench.cpu_window(inst, 300, 5)
    assert set(cb) == {"value", "unit", "cores", "kind", "sample", "cpu"}
    assert cb["value"] > 0 and cb["unit"] == "clump-steps/s"
    assert cb["cores"] == 1 and cb["kind"] == "port" and cb["cpu"]
    assert "the first 300 of the workload's 2000 clumps, a 5-step window" in cb["sample"]
    assert "pinned to 1 core" in cb["sample"]


def test_cpu_batch_baseline_fields():
    import bench
    cb = bench.cpu_batch_baseline(0.2, 2)
    assert cb["cores"] == 2 and cb["kind"] == "port" and cb["value"] > 0
    assert "the first 2 of the 64 batch instances" in cb["sample"]
