"""Product host setup and the oracle against fixtures written from the reference's own host
ops (tests/golden/make_golden.py).  Runs anywhere: the fixtures are committed data."""
import math
import os

import numpy as np
import pytest

from avalancheflume3d_amd import abi, dem
from oracle import oracle
from tests import scenarios

G = scenarios.GOLDEN


def load(name):
    return np.load(os.path.join(G, name))


def bits_equal(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_rannew64_sequences():
    z = load("rannew64.npz")
    for s in (26752658, 12345, 987654321):
        rng, idum = dem.Rng(), s
        got = []
        for _ in range(2000):
            v, idum = rng(idum)
            got.append(v)
        assert bits_equal(np.array(got), z[f"seed_{s}"]), s
    rng, idum, got = dem.Rng(), 777, []
    for _ in range(300):
        v, idum = rng(idum)
        got.append(v)
    idum = -4242
    for _ in range(300):
        v, idum = rng(idum)
        got.append(v)
    assert bits_equal(np.array(got), z["reseed_777_then_-4242"])


def test_initial_velocity():
    z = load("initial_velocity.npz")
    sys, geo, fun, nt = scenarios.reference_system(300)
    pp = dem.new_particles(300)
    dem.set_initial_velocity(pp, sys)
    assert bits_equal(pp["vv"], z["vv"]) and bits_equal(pp["ww"], z["ww"])
    # upper hemisphere, |v| = v0, w = v / (4 rmax) (set_initial_velocity.cu:20-35)
    assert (pp["vv"][:, 2] >= 0).all()
    np.testing.assert_allclose(np.linalg.norm(pp["vv"], axis=1), sys.v0, rtol=1e-14)
    np.testing.assert_allclose(pp["ww"], pp["vv"] / (4 * sys.rmax), rtol=1e-13)


def test_generation_and_placement():
    ref = load("setup.npz")["pp"].view(abi.particle_dtype)
    sys, geo, fun, nt = scenarios.reference_system(1000)
    pp = dem.new_particles(1000)
    dem.generate_particles(pp, sys)
    dem.calculate_particle_parameters(pp, sys, geo)
    dem.place_particles(pp, sys, geo)
    for f in ("radius", "density", "mass", "moment_of_inertia", "rr", "rr_a", "rr_b", "rr_c",
              "rr_d", "Rmatx", "Rmaty", "Rmatz"):
        assert bits_equal(pp[f], ref[f]), f


@pytest.mark.parametrize("name", ["steps_fill", "steps_avalanche"])
def test_oracle_trajectory_fixtures(name):
    z = load(f"{name}.npz")
    ntags, phase, nsteps, chunks = z["meta"].tolist()
    sys = abi.System.from_buffer_copy(z["sys"].tobytes())
    fun = abi.Funnel.from_buffer_copy(z["fun"].tobytes())
    geo = abi.Geometry.from_buffer_copy(z["geo"].tobytes())
    pp = z["input"].view(abi.particle_dtype)
    o = oracle.OracleSim(sys, geo, fun, ntags, phase, pp)
    o.initial()
    g = o.get()
    assert bits_equal(g["ff"], z["ff_0"]) and bits_equal(g["aa_i"], z["aa_0"])
    assert np.abs(z["ff_0"]).max() > 0
    for k in range(1, chunks + 1):
        o.step(nsteps // chunks)
        g = o.get()
        ref = z[f"pp_{k}"].view(abi.particle_dtype)
        for f in abi.PARTICLE_FIELDS:
            if phase == abi.PHASE_FILL and f.startswith("rrot"):
                continue  # never computed while filling (rotate_vectors not called)
            assert bits_equal(g["pp"][f], ref[f]), (k, f)
        assert bits_equal(g["ff"], z[f"ff_{k}"]) and bits_equal(g["tq"], z[f"tq_{k}"])
        assert g["sys"].time == z[f"time_{k}"][0]


def test_derived_constants():
    """avalanche_hst.cu:84-103,187-195 against closed forms and the constants the reference
    logs for its DataInput.dat (kappa 1.01305e9, gamma 1.02165e4, xfinal 172.02)."""
    sys, geo, fun, nt = scenarios.reference_system(20358)
    tc, e = 1.0e-4, 0.6
    assert sys.kappa_pp == pytest.approx((math.pi / tc) ** 2 + (math.log(e) / tc) ** 2, rel=1e-15)
    assert sys.gamma_pp == pytest.approx(-2 * math.log(e) / tc, rel=1e-15)
    assert sys.kappa_pp == pytest.approx(1.01305e9, rel=1e-5)
    assert sys.gamma_pp == pytest.approx(1.02165e4, rel=1e-5)
    assert (geo.lfunnel + sys.Lflume) * sys.cost == pytest.approx(172.02, abs=5e-3)
    assert sys.dy2 == 7.5 and sys.hdt == 0.5e-6 and sys.side.x == 15.0
    assert fun.tanfi == 7.5 / 70.0
    assert sys.fr_fact_pp == 0.4 / 1.0e12
    assert geo.cell_dist == 4.499489743528236  # GYR + 0.05f (float literal)


def test_error_check_codes():
    base, geo, fun, nt = scenarios.reference_system(100)
    cases = {0: {}, 1: {"npart": 0}, 3: {"rmin": -1.0}, 4: {"rmin": 0.3}, 5: {"eps_pw": 1.5},
             6: {"eps_pp": 0.0}, 7: {"mu_pw": -0.1}, 8: {"mu_pp": 2.0}, 9: {"mu_pw": 0.9}}
    for code, kv in cases.items():
        s = abi.System.from_buffer_copy(base)
        for k, v in kv.items():
            setattr(s, k, v)
        got, msg = dem.error_check(s)
        assert got == code, (kv, got, msg)
        if oracle.ref_available():
            import ctypes as C
            buf = C.create_string_buffer(200)
            assert oracle.ref_lib().ref_error_check(abi.ptr(s), buf) == code
    s = abi.System.from_buffer_copy(base)
    s.side.z = 0.0
    assert dem.error_check(s)[0] == 2


def test_parse_datainput_errors():
    with pytest.raises(dem.AvfError) as e:
        dem.parse_datainput("\n".join(scenarios.DATAINPUT.splitlines()[:10]))
    assert e.value.code == 8
    bad = scenarios.DATAINPUT.splitlines()
    bad[6] = "fast   Time step."
    with pytest.raises(dem.AvfError):
        dem.parse_datainput("\n".join(bad))
