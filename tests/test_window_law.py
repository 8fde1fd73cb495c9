"""Host check of the hand-written contact-law adjoint (contact_force_bwd in
avalancheflume3d_amd/csrc/dem_window.h) against dual-number tangents of the templated
forward contact_force, over random contacts covering every branch of the law: stick,
stick after a reset, slide and slide after a reset, for the particle rule and the wall
rule (force_part.cu:188-242).  Compiled for the host with hipcc (no GPU needed)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not found")
def test_contact_law_adjoint_matches_dual_tangent(tmp_path):
    exe = str(tmp_path / "window_law_check")
    src = os.path.join(ROOT, "tests", "native", "window_law_check.cpp")
    inc = os.path.join(ROOT, "avalancheflume3d_amd", "csrc")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-I", inc, src, "-o", exe],
                   check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    line = r.stdout.strip().split()
    # "branches stick A stick-reset B slide C slide-reset D worst E"
    vals = {line[i]: int(line[i + 1]) for i in range(1, 9, 2)}
    assert min(vals.values()) > 100, vals
