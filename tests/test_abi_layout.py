"""The records of the C ABI have the same size and field offsets in the Python declarations
(avalancheflume3d_amd/abi.py: ctypes structures and numpy dtypes) as in include/avflume.h.
A probe compiled against the header with the host C compiler prints sizeof / offsetof of
every field the Python side declares; a field missing from the header fails the compile."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

from avalancheflume3d_amd import abi

INCLUDE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")

STRUCTS = [("avf_system", abi.System), ("avf_funnel", abi.Funnel), ("avf_cell", abi.Cell),
           ("avf_geometry", abi.Geometry), ("avf_instance_desc", abi.InstanceDesc),
           ("avf_options", abi.Options), ("avf_diag", abi.Diag),
           ("avf_observables", abi.Observables), ("avf_sphere_grid", abi.SphereGrid),
           ("avf_profile_spec", abi.ProfileSpec), ("avf_window_spec", abi.WindowSpec),
           ("avf_window_diag", abi.WindowDiag), ("avf_rng", abi.Rng)]
DTYPES = [("avf_particle", abi.particle_dtype), ("avf_contact", abi.contact_dtype)]


def expected():
    want = {}
    for cname, cls in STRUCTS:
        want[f"{cname} size"] = C.sizeof(cls)
        for f, _ in cls._fields_:
            want[f"{cname}.{f}"] = getattr(cls, f).offset
    for cname, dt in DTYPES:
        want[f"{cname} size"] = dt.itemsize
        for f in dt.names:
            want[f"{cname}.{f}"] = dt.fields[f][1]
    return want


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs a host C compiler")
def test_record_layouts_match_header(tmp_path):
    want = expected()
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "avflume.h"',
             "int main(void) {"]
    for key in want:
        if key.endswith(" size"):
            lines.append(f'  printf("{key} %zu\\n", sizeof({key[:-5]}));')
        else:
            rec, field = key.split(".")
            lines.append(f'  printf("{key} %zu\\n", offsetof({rec}, {field}));')
    lines.append("  return 0;\n}")
    src, exe = tmp_path / "probe.c", tmp_path / "probe"
    src.write_text("\n".join(lines) + "\n")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-I", INCLUDE, str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    got = {}
    for line in out.splitlines():
        key, val = line.rsplit(" ", 1)
        got[key] = int(val)
    assert set(got) == set(want)
    bad = {k: (got[k], want[k]) for k in want if got[k] != want[k]}
    assert not bad, f"header vs Python (header, python): {bad}"
