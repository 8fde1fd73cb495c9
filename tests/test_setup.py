"""Host setup (DataInput.dat parsing, derived constants, error_check) against closed forms
and the reference's own input file.  CPU only."""
import math

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
