"""The C++ host driver `avalanche_amd` (the reference's main on the C ABI).

CPU: argument and DataInput.dat error paths (they return before any GPU call).
GPU: the first steps of the reference run (hopper filling) against the oracle on the same
seeded setup; the lossless binary checkpoint (a run split by a restart equals the straight
run bit for bit); the reference's text checkpoint is read back.
"""
import os
import re
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


# ---- the two phases, the stale-mean-speed stops and the output formats -----------------
#
# avalanche_hst.cu runs the filling loop while time < 0 and stops it early when the mean
# speed last computed (at the output steps, cc % 20000 == 0) is below v0 / 10 (:597-598);
# the gate then opens: side.x = xfinal, cc = 0, time = 0 (:606-613); the avalanche runs while
# time < 2 and stops when that stale mean speed is below v0 / 10 after time 0.2 (:854-855).
# The writers are :519-575 (filling) and :763-829 (avalanche).

F6 = r"-?\d+\.\d{6}"


def write_last_config(path, pp, aval_start, time, cc, jj):
    """last_config.dat as avalanche_hst.cu:521-545 writes it (the restart's input)."""
    with open(path, "w") as f:
        f.write("%d\n%f\n%d\n%d\n" % (aval_start, time, cc, jj))
        for i, p in enumerate(pp):
            vals = [*p["rr"], *p["Rmatx"], *p["Rmaty"], *p["Rmatz"], *p["vv"], *p["ww"]]
            f.write("%d " % i + " ".join("%f" % v for v in vals) + "\n")


def read_last_config(path, n, sys, geo):
    """The restart's reading (avalanche_hst.cu:222-249, 265-293) onto freshly generated
    clumps: centres, rotation rows, velocities from the file; sphere centres recomputed."""
    lines = open(path).read().splitlines()
    aval, t, cc, jj = int(lines[0]), float(lines[1]), int(lines[2]), int(lines[3])
    pp = dem.new_particles(n)
    dem.generate_particles(pp, sys, 2.6)
    dem.calculate_particle_parameters(pp, sys, abi.Geometry.from_buffer_copy(geo))
    for ln in lines[4:]:
        f = ln.split()
        i, v = int(f[0]), [float(x) for x in f[1:]]
        pp["rr"][i], pp["Rmatx"][i], pp["Rmaty"][i], pp["Rmatz"][i] = v[0:3], v[3:6], v[6:9], v[9:12]
        pp["vv"][i], pp["ww"][i] = v[12:15], v[15:18]
    dem.sphere_centres(pp, geo)
    return pp, aval, t, cc, jj


def mean_speeds(pp):
    """avalanche_hst.cu:547-575: mean |v| and |w| summed in clump order."""
    sv = sw = 0.0
    for v, w in zip(pp["vv"], pp["ww"]):
        vm2 = v[0] * v[0]
        vm2 += v[1] * v[1]
        vm2 += v[2] * v[2]
        sv += float(np.sqrt(vm2))
        wm2 = w[0] * w[0]
        wm2 += w[1] * w[1]
        wm2 += w[2] * w[2]
        sw += float(np.sqrt(wm2))
    return sv / len(pp), sw / len(pp)


def close6(line_vals, want, what):
    for a, b in zip(line_vals, want):
        assert abs(float(a) - b) <= 1.5e-6 + 1e-9 * abs(b), (what, a, b)


def start_state(n, sys, geo):
    """The reference's seeded clumps, placed in the hopper and at rest."""
    pp = dem.new_particles(n)
    dem.generate_particles(pp, sys, 2.6)
    dem.calculate_particle_parameters(pp, sys, geo)
    dem.place_particles(pp, sys, geo)
    dem.set_initial_velocity(pp, sys)
    pp["vv"] = 0.0
    pp["ww"] = 0.0
    return pp


@needs_driver
@pytest.mark.gpu
def test_driver_phases_stale_speed_stops_and_formats(gpu, tmp_path):
    n = 300
    text = scenarios.datainput(n, print_flag=1)
    sys, geo, fun, ntags = scenarios.reference_system(n)
    xfinal = (geo.lfunnel + sys.Lflume) * sys.cost  # avalanche_hst.cu:116
    # A: filling restarted at t = -0.0005 on an output step (cc = 20000) with the clumps at
    # rest: the first step's mean speed (~ g dt) is below v0 / 10, so the filling stops after
    # it and the gate opens; 299 avalanche steps follow (--max-steps 300)
    a = tmp_path / "a"
    a.mkdir()
    write_last_config(a / "last_config.dat", start_state(n, sys, geo), 0, -0.0005, 20000, 5)
    pp0, _, t0, cc0, jj0 = read_last_config(a / "last_config.dat", n, sys, geo)
    r = run(a, ["--restart", "--max-steps", "300", "--final-bin", "end.bin"], text)
    assert r.returncode == 0, r.stderr
    assert "Opening hopper... Avalanche starts!" in r.stdout
    hdr, pa, aa_, ga_ = read_bin(a / "end.bin", n)
    sys_end = abi.System.from_buffer_copy(open(a / "end.bin", "rb").read()[MAGIC_BYTES + HDR:][:256])
    assert hdr[0] == 1 and hdr[1] == 299  # avalanche phase, cc counted from the opening
    t = 0.0
    for _ in range(299):
        t += sys.dt
    assert sys_end.time == t and sys_end.side.x == xfinal
    # the same protocol through the C ABI: bit for bit
    s2 = abi.System.from_buffer_copy(sys)
    s2.time = t0
    inst = dem.Instance(sys=s2, geo=geo, fun=fun, particles=pp0.copy(), ntags=ntags,
                        phase=abi.PHASE_FILL)
    ctx = dem.Context([inst])
    ctx.set_phase(0, abi.PHASE_FILL, t0, sys.side.x, cc0)
    ctx.regrid()
    ctx.initial_forces()
    ctx.step(1)
    ctx.set_phase(0, abi.PHASE_AVALANCHE, 0.0, xfinal, 0)
    ctx.step(1)
    ctx.step(298)
    pc, ac, gc = ctx.download()
    ctx.close()
    assert pc.tobytes() == pa.tobytes() and ac.tobytes() == aa_.tobytes()
    # the oracle over the two output steps (the first step of each phase)
    o = oracle.OracleSim(s2, geo, fun, ntags, abi.PHASE_FILL, pp0)
    o.initial()
    o.step(1)
    fill_pp = o.get()["pp"]
    v_fill, w_fill = mean_speeds(fill_pp)
    assert v_fill < 0.1 * sys.v0  # the stop the driver took
    o.open(xfinal)
    o.step(1)
    av_pp = o.get()["pp"]
    v_av, w_av = mean_speeds(av_pp)
    mv = [l.split(", ") for l in open(a / "mean_velocity_amd.csv").read().splitlines()]
    assert all(re.fullmatch(rf"{F6}, {F6}, {F6}", ", ".join(x)) for x in mv)
    assert len(mv) == 2
    close6(mv[0], [t0, v_fill, w_fill], "filling mean speeds")  # time - dt of step 1 = t0
    close6(mv[1], [0.0, v_av, w_av], "avalanche mean speeds")
    xc = [l.split(", ") for l in open(a / "xcm_amd.csv").read().splitlines()]
    assert len(xc) == 1 and re.fullmatch(rf"{F6}, {F6}, {F6}", ", ".join(xc[0]))
    tot = 0.0
    for m in pp0["mass"]:
        tot += float(m)
    xm = zm = 0.0
    for m, rr in zip(av_pp["mass"], av_pp["rr"]):
        xm += m * rr[0]
        zm += m * rr[2]
    close6(xc[0], [0.0, xm / tot, zm / tot], "centre of mass")
    # text checkpoint of the avalanche's output step
    lc = open(a / "last_config.dat").read().splitlines()
    assert lc[0] == "1" and re.fullmatch(F6, lc[1]) and lc[2] == "1" and lc[3] == str(jj0 + 2)
    assert float(lc[1]) == pytest.approx(sys.dt, abs=1e-6)
    for i, ln in enumerate(lc[4:]):
        f = ln.split(" ")
        assert len(f) == 19 and f[0] == str(i) and all(re.fullmatch(F6, x) for x in f[1:])
        p = av_pp[i]
        close6(f[1:], [*p["rr"], *p["Rmatx"], *p["Rmaty"], *p["Rmatz"], *p["vv"], *p["ww"]],
               "checkpoint row")
    # snapshots: two frames of 4 n spheres + 2 ghosts; velocities: n clumps + 2 ghosts
    xyz = open(a / "snapshots_amd.xyz").read().split("\n")
    vel = open(a / "velocities_amd.xyz").read().split("\n")
    k = 0
    for frame, state in ((0, fill_pp), (1, av_pp)):
        assert xyz[k] == str(4 * n + 2) and xyz[k + 1] == ""
        k += 2
        for i in range(n):
            for s, key in enumerate(("rr_a", "rr_b", "rr_c", "rr_d")):
                f = xyz[k].split(" ")
                assert re.fullmatch(rf"{F6} {F6} {F6} {F6} \d+", xyz[k]) and f[4] == str(i)
                close6(f[:4], [state["radius"][i], *state[key][i]], "sphere line")
                k += 1
        # ghost clumps: xfinal, or side.x (= xfinal once open), :563 vs :808
        ghost = "0.000001" if frame == 0 else "0.0001"
        assert xyz[k] == "%s %f %f %f %d" % (ghost, xfinal, geo.yfunnel, sys.side.z, n), xyz[k]
        assert xyz[k + 1] == "0.000001 0.000000 0.000000 0.000000 %d" % n
        k += 2
    k = 0
    for state in (fill_pp, av_pp):
        assert vel[k] == str(n + 2) and vel[k + 1] == ""
        k += 2
        for i in range(n):
            f = vel[k].split(" ")
            assert re.fullmatch(rf"{F6} {F6} {F6} {F6} {F6} \d+", vel[k]) and f[5] == str(i)
            close6(f[:5], [*state["rr"][i], state["vv"][i][0], state["ww"][i][1]], "velocity line")
            k += 1
        assert vel[k] == "%f %f %f 0.000000 0.000000 %d" % (xfinal, geo.yfunnel, sys.side.z, n)
        assert vel[k + 1] == "0.000000 0.000000 0.000000 0.000000 0.000000 %d" % n
        k += 2
    # B: avalanche restarted at t = 0.19995 on an output step with the clumps at rest: the
    # mean speed of its first step stays the stale value, below v0 / 10, so the run ends on
    # the first step whose time exceeds 0.2 (not at t = 2)
    b = tmp_path / "b"
    b.mkdir()
    write_last_config(b / "last_config.dat", start_state(n, sys, geo), 1, 0.19995, 40000, 7)
    r = run(b, ["--restart", "--final-bin", "end.bin"], text)
    assert r.returncode == 0, r.stderr
    hdr, _, _, _ = read_bin(b / "end.bin", n)
    sys_b = abi.System.from_buffer_copy(open(b / "end.bin", "rb").read()[MAGIC_BYTES + HDR:][:256])
    t, k = 0.19995, 0
    while True:
        t += sys.dt
        k += 1
        if t > 0.2:
            break
    assert hdr[0] == 1 and hdr[1] == 40000 + k and sys_b.time == t
    assert len(open(b / "mean_velocity_amd.csv").read().splitlines()) == 1
