"""The C-ABI library loads and exports every entry point include/avflume.h declares."""
import ctypes as C
import os
import re

import pytest

from avalancheflume3d_amd import abi, dem
from tests.conftest import gpu_available

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "avflume.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(avf_[a-z_0-9]+)\s*\(",
                                 text, re.M)))


def test_header_and_exports_match():
    names = declared()
    assert len(names) >= 30
    lib = dem.lib()
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(abi.EXPORTED)


def test_record_sizes():
    assert abi.particle_dtype.itemsize == 408
    assert C.sizeof(abi.System) == 256 and C.sizeof(abi.Cell) == 168
    assert C.sizeof(abi.Funnel) == 24
    assert C.sizeof(abi.Geometry) == 16 * 8 + 6 * 4
    assert C.sizeof(abi.Contact) if hasattr(abi, "Contact") else True


@pytest.mark.skipif(gpu_available(), reason="checks the no-device error path")
def test_create_without_device_reports_error():
    inst = dem.Instance(*dem_parts(), particles=dem.new_particles(4))
    with pytest.raises(dem.AvfError) as e:
        dem.Context([inst])
    assert e.value.code == 2 and "hip" in str(e.value).lower()


def test_create_rejects_empty_and_bad_batches():
    """Argument errors come back before any device call: an instance without clumps, an
    empty batch."""
    sys, geo, fun = dem_parts()
    sys.npart = 0
    desc = (abi.InstanceDesc * 1)()
    desc[0].sys, desc[0].geo, desc[0].fun, desc[0].ntags = sys, geo, fun, 9
    h = C.c_void_p()
    lib = abi.load_library()
    assert lib.avf_create(C.byref(h), desc, 1, None) == abi.STATUS_CODES["AVF_ERR_ARG"]
    assert b"npart" in lib.avf_last_error(None)
    assert lib.avf_create(C.byref(h), desc, 0, None) == abi.STATUS_CODES["AVF_ERR_ARG"]
    assert lib.avf_create(None, desc, 1, None) == abi.STATUS_CODES["AVF_ERR_ARG"]


def dem_parts():
    from tests import scenarios
    sys, geo, fun, nt = scenarios.reference_system(4)
    return sys, geo, fun
