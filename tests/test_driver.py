"""The C++ host driver `avalanche_amd` (the reference's main on the C ABI).

CPU: argument and DataInput.dat error paths (they return before any GPU call).
GPU: the first steps of the reference run (hopper filling) against the oracle on the same
seeded setup; the lossless binary checkpoint (a run split by a restart equals the straight
run bit for bit); the reference's text checkpoint is read back.
"""
import os
import subprocess

import numpy as np
import pytest

from avalancheflume3d_amd import abi, dem
from oracle import oracle
from tests import scenarios

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "avalancheflume3d_amd", "avalanche_amd")
MAGIC_BYTES = 8
HDR = 4 * 8

needs_driver = pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")


def run(cwd, args, text, timeout=300):
    return subprocess.run([DRIVER, *args], input=text, cwd=cwd, capture_output=True, text=True,
                          timeout=timeout)


def read_bin(path, n):
    raw = open(path, "rb").read()
    off = MAGIC_BYTES
    hdr = np.frombuffer(raw, np.int64, 4, off)
    off += HDR + 256
    pp = np.frombuffer(raw, abi.particle_dtype, n, off)
    off += n * abi.particle_dtype.itemsize
    aa = np.frombuffer(raw, np.float64, 3 * n, off).reshape(n, 3)
    ga = np.frombuffer(raw, np.float64, 3 * n, off + 24 * n).reshape(n, 3)
    return hdr, pp, aa, ga


@needs_driver
def test_driver_usage_and_input_errors(tmp_path):
    r = run(tmp_path, ["--bogus"], "")
    assert r.returncode == 2 and "usage" in r.stderr
    r = run(tmp_path, [], "not a DataInput file\n")
    assert r.returncode == 1
    r = run(tmp_path, [], scenarios.datainput(300, mu_pw=1.5))  # error_check.cu
    assert r.returncode == 1 and "mu_pw" in r.stderr


@needs_driver
@pytest.mark.gpu
def test_driver_first_steps_match_oracle(gpu, tmp_path):
    n, steps = 300, 6
    text = scenarios.datainput(n, print_flag=1)
    r = run(tmp_path, ["--max-steps", str(steps), "--final-bin", "final.bin"], text)
    assert r.returncode == 0, r.stderr
    for f in ("log_amd.dat", "last_config.dat", "last_config.bin", "mean_velocity_amd.csv",
              "xcm_amd.csv", "snapshots_amd.xyz", "velocities_amd.xyz", "final.bin"):
        assert (tmp_path / f).exists(), f
    hdr, pp, aa, ga = read_bin(tmp_path / "final.bin", n)
    assert hdr[0] == 0 and hdr[1] == steps  # filling phase, cc
    # the same seeded setup through the ABI, stepped by the oracle
    sys, geo, fun, ntags = scenarios.reference_system(n)
    q = dem.new_particles(n)
    dem.generate_particles(q, sys, 2.6)
    dem.calculate_particle_parameters(q, sys, geo)
    dem.place_particles(q, sys, geo)
    dem.set_initial_velocity(q, sys)
    o = oracle.OracleSim(sys, geo, fun, ntags, abi.PHASE_FILL, q)
    o.initial()
    o.step(steps)
    ref = o.get()["pp"]
    for fld in ("rr", "vv", "ww", "Rmatx", "Rmaty", "Rmatz"):
        err = np.abs(pp[fld] - ref[fld]).max() / max(np.abs(ref[fld]).max(), 1e-300)
        assert err <= 1e-9, (fld, err)
    # the reference's text checkpoint: 4 header lines, then the clump index and 18 columns
    # (rr, R rows, vv, ww) per clump (avalanche_hst.cu:225-245)
    rows = [l.split() for l in open(tmp_path / "last_config.dat").read().splitlines()[4:] if l.strip()]
    assert len(rows) == n
    assert all(len(r_) == 19 and int(r_[0]) == i for i, r_ in enumerate(rows))


@needs_driver
@pytest.mark.gpu
def test_driver_binary_restart_is_lossless(gpu, tmp_path):
    n = 300
    text = scenarios.datainput(n)
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    assert run(a, ["--max-steps", "14", "--final-bin", "end.bin"], text).returncode == 0
    assert run(b, ["--max-steps", "5", "--final-bin", "mid.bin"], text).returncode == 0
    r = run(b, ["--restart-bin", "mid.bin", "--max-steps", "9", "--final-bin", "end.bin"], text)
    assert r.returncode == 0, r.stderr
    ha, pa, aa_, ga_ = read_bin(a / "end.bin", n)
    hb, pb, ab_, gb_ = read_bin(b / "end.bin", n)
    assert ha[1] == hb[1] == 14
    assert pa.tobytes() == pb.tobytes()
    assert aa_.tobytes() == ab_.tobytes() and ga_.tobytes() == gb_.tobytes()
    # the default candidate lists and per-step binning (--skin 0) write the same state
    c = tmp_path / "c"
    c.mkdir()
    assert run(c, ["--skin", "0", "--max-steps", "14", "--final-bin", "end.bin"], text).returncode == 0
    hc, pc, ac_, gc_ = read_bin(c / "end.bin", n)
    assert pc.tobytes() == pa.tobytes() and ac_.tobytes() == aa_.tobytes()
    # the text checkpoint (6 decimals) is read back by --restart
    r = run(b, ["--restart", "--max-steps", "2"], text)
    assert r.returncode == 0, r.stderr
