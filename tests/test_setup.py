"""Host setup (DataInput.dat parsing, derived constants, error_check) against closed forms
and the reference's own input file.  CPU only."""
import math

import numpy as np

import pytest

from avalancheflume3d_amd import abi, dem
from tests import scenarios


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


def test_generate_particles_bidisperse_pattern():
    """generate_particles.cu:17-29: density 2.6, every 20th clump (index 0, 20, 40, ...) of
    radius rmax, the others rmin."""
    sys, geo, fun, nt = scenarios.reference_system(95, rmin=0.1, rmax=0.17)
    pp = dem.new_particles(95)
    dem.generate_particles(pp, sys)
    assert (pp["density"] == 2.6).all()
    big = np.arange(95) % 20 == 0
    assert (pp["radius"][big] == 0.17).all() and (pp["radius"][~big] == 0.1).all()


def test_place_particles_cubic_lattice():
    """place_particles.cu:8-90: identity orientations; centres on a cubic lattice of spacing
    (2 + ROOT2 + 0.05 rmax) rmax inside the hopper (x in [0, side.x], y in [YHOP_L, YHOP_R],
    z above HOP_HEIGHT), all distinct, none closer than one lattice spacing."""
    n = 400
    sys, geo, fun, nt = scenarios.reference_system(n)
    pp = dem.new_particles(n)
    dem.generate_particles(pp, sys)
    dem.calculate_particle_parameters(pp, sys, geo)
    dem.place_particles(pp, sys, geo)
    r = sys.rmax
    a = (2.0 + abi.ROOT2 + 0.05 * r) * r
    assert np.array_equal(pp["Rmatx"], np.tile([1.0, 0, 0], (n, 1)))
    assert np.array_equal(pp["Rmaty"], np.tile([0, 1.0, 0], (n, 1)))
    rr = pp["rr"]
    assert (rr[:, 0] > 0).all() and (rr[:, 0] < sys.side.x).all()
    assert (rr[:, 1] > geo.yhop_l).all() and (rr[:, 1] < geo.yhop_r).all()
    assert (rr[:, 2] > geo.hop_height).all()
    g = (rr - rr.min(0)) / a
    np.testing.assert_allclose(g, np.round(g), atol=1e-9)
    d = np.linalg.norm(rr[:, None, :] - rr[None, :, :], axis=2) + np.eye(n) * 1e9
    assert d.min() == pytest.approx(a, rel=1e-12)


def test_initial_velocity_upper_hemisphere():
    """set_initial_velocity.cu:17-35: |v| = v0 with angles 2 pi U - pi and pi U, so v_z >= 0
    and every direction of the upper hemisphere occurs; w = v / (4 rmax)."""
    n = 3000
    sys, geo, fun, nt = scenarios.reference_system(n)
    pp = dem.new_particles(n)
    dem.set_initial_velocity(pp, sys)
    v = pp["vv"]
    np.testing.assert_allclose(np.linalg.norm(v, axis=1), sys.v0, rtol=1e-14)
    assert (v[:, 2] >= 0).all()
    np.testing.assert_allclose(pp["ww"], v / (4 * sys.rmax), rtol=1e-13)
    # both signs of v_x and v_y, and the mean of sin(elevation) = 2 / pi for a uniform angle
    assert (v[:, 0] > 0).mean() == pytest.approx(0.5, abs=0.05)
    assert (v[:, 1] > 0).mean() == pytest.approx(0.5, abs=0.05)
    assert (v[:, 2] / sys.v0).mean() == pytest.approx(2 / math.pi, abs=0.03)


def rannew64_published(seed: int, count: int, lags=(249, 102)):
    """rannew64.cu:55-126 as published: a lagged-Fibonacci generator x_n = x_{n-250} -
    x_{n-103} (mod 2^64) seeded by the 64-bit LCG of Knuth (a = 6364136223846793005,
    c = 1442695040888963407; 1000 warm-up draws, then 250 seeds), minus a second 64-bit LCG
    (a = 2862933555777941757, c = 3037000493) started from the last seed, scaled by
    1 / (2^64 - 1).  Python integers, so the modular arithmetic is exact.  lags: the two
    positions of the lagged generator (static in the reference and not reset when a negative
    seed re-initialises the table)."""
    M = 1 << 64
    u = seed % M
    for _ in range(1000):
        u = (u * 6364136223846793005 + 1442695040888963407) % M
    seeds = []
    for _ in range(250):
        u = (u * 6364136223846793005 + 1442695040888963407) % M
        seeds.append(u)
    y = u
    np_, nq = lags
    out = []
    for _ in range(count):
        x = (seeds[np_] - seeds[nq]) % M
        seeds[np_] = x
        np_ = np_ - 1 if np_ > 0 else 249
        nq = nq - 1 if nq > 0 else 249
        y = (2862933555777941757 * y + 3037000493) % M
        out.append(float((x - y) % M) * (1.0 / 18446744073709551615.0))
    return out


def test_rannew64_against_the_published_generator():
    for seed in (26752658, 12345, 987654321):
        rng, idum, got = dem.Rng(), seed, []
        for _ in range(600):
            v, idum = rng(idum)
            got.append(v)
        assert got == rannew64_published(seed, 600)
    # a negative seed re-initialises the table (rannew64.cu:69-94) and is made positive; the
    # lag positions carry on from the draws before
    rng = dem.Rng()
    _, idum = rng(777)
    v, idum = rng(-4242)
    assert idum == 4242 and v == rannew64_published(4242, 1, lags=(248, 101))[0]
    xs = np.array(rannew64_published(99, 20000))
    assert xs.min() >= 0 and xs.max() <= 1
    assert xs.mean() == pytest.approx(0.5, abs=0.01) and xs.var() == pytest.approx(1 / 12, abs=0.003)
