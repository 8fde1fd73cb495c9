"""ctypes bindings of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/liboracle.so (this repository's C restatement of the reference's host ops).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module; the
product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from avalancheflume3d_amd import abi
from avalancheflume3d_amd.abi import ptr

HERE = os.path.dirname(os.path.abspath(__file__))
# AVF_ORACLE_LIB: another build of the same restatement (e.g. a deliberately altered copy
# built to check that the tests detect the alteration).
ORACLE_LIB = os.environ.get("AVF_ORACLE_LIB", os.path.join(HERE, "liboracle.so"))

_orc = None


def build() -> None:
    """Compiles the oracle."""
    subprocess.run(["make", "-C", HERE, "oracle"], check=True, capture_output=True)


def oracle_lib() -> C.CDLL:
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"{ORACLE_LIB} missing: run `make -C oracle oracle`")
        lib = C.CDLL(ORACLE_LIB)
        vp, i64, i32 = C.c_void_p, C.c_int64, C.c_int32
        lib.orc_sim_new.restype = vp
        lib.orc_sim_new.argtypes = [vp, vp, vp, i64, i32, vp]
        lib.orc_sim_delete.argtypes = [vp]
        lib.orc_sim_steps.argtypes = [vp, i64]
        lib.orc_sim_initial.argtypes = [vp]
        lib.orc_sim_forces.argtypes = [vp]
        lib.orc_sim_regrid.argtypes = [vp]
        lib.orc_sim_open.argtypes = [vp, C.c_double]
        lib.orc_sim_get.argtypes = [vp] * 8
        lib.orc_sim_set.argtypes = [vp] * 4
        lib.orc_sim_record.argtypes = [vp, vp, i64]
        lib.orc_sim_recorded.restype = i64
        lib.orc_sim_recorded.argtypes = [vp]
        lib.orc_sim_error.restype = i32
        lib.orc_sim_error.argtypes = [vp, vp]
        lib.orc_brute_contacts.restype = i64
        lib.orc_brute_contacts.argtypes = [vp, i64, vp, vp, i64]
        lib.orc_cell_setup.argtypes = [vp, vp, vp, vp]
        lib.orc_cell_locate.restype = C.c_int
        lib.orc_cell_locate.argtypes = [vp, vp, vp, vp, vp, vp]
        _orc = lib
    return _orc


def cell_table(cell: abi.Cell, sys: abi.System, pp: np.ndarray):
    """clear_cell / clear_ocup / cell_locate (cells.cu:11-111) of the restatement:
    (cell_vec, nocup, status, offending clump)."""
    cell_vec = np.full(cell.ncell3 * cell.ntags, -1, dtype=np.int64)
    nocup = np.zeros(cell.ncell3, dtype=np.int32)
    s = abi.System.from_buffer_copy(sys)
    s.npart = len(pp)
    bad = C.c_int64(-1)
    rc = oracle_lib().orc_cell_locate(ptr(cell), ptr(s), ptr(np.ascontiguousarray(pp)),
                                      ptr(nocup), ptr(cell_vec), C.byref(bad))
    return cell_vec, nocup, rc, bad.value


class _Sim:
    """Common wrapper: one instance of the host step loop."""

    def __init__(self, sys: abi.System, n: int):
        self.n = n
        self.sys0 = sys

    def get(self) -> dict:
        n = self.n
        out = {"pp": np.zeros(n, dtype=abi.particle_dtype), "aa_i": abi.v3(n),
               "anga_i": abi.v3(n), "ff": abi.v3(n), "tq": abi.v3(n)}
        sys, cell = abi.System(), abi.Cell()
        self._get(ptr(out["pp"]), ptr(out["aa_i"]), ptr(out["anga_i"]), ptr(out["ff"]),
                  ptr(out["tq"]), ptr(sys), ptr(cell))
        out["sys"], out["cell"] = sys, cell
        return out


class OracleSim(_Sim):
    def __init__(self, sys: abi.System, geo: abi.Geometry, fun: abi.Funnel, ntags: int,
                 phase: int, pp: np.ndarray):
        super().__init__(sys, len(pp))
        self.lib = oracle_lib()
        s = abi.System.from_buffer_copy(sys)
        s.npart = len(pp)
        self.h = self.lib.orc_sim_new(ptr(s), ptr(geo), ptr(fun), ntags, phase,
                                      ptr(np.ascontiguousarray(pp)))
        if not self.h:
            raise RuntimeError("orc_sim_new failed")
        self._recbuf = None

    def _get(self, *a):
        self.lib.orc_sim_get(self.h, *a)

    def set(self, pp=None, aa=None, anga=None):
        self.lib.orc_sim_set(self.h, ptr(pp), ptr(aa), ptr(anga))

    def _rc(self, rc):
        if rc:
            bad = C.c_int64()
            self.lib.orc_sim_error(self.h, C.byref(bad))
            raise RuntimeError(f"oracle error {abi.STATUS.get(rc, rc)} at clump {bad.value}")

    def initial(self):
        self._rc(self.lib.orc_sim_initial(self.h))

    def forces(self):
        self._rc(self.lib.orc_sim_forces(self.h))

    def regrid(self):
        self._rc(self.lib.orc_sim_regrid(self.h))

    def step(self, k: int = 1):
        self._rc(self.lib.orc_sim_steps(self.h, k))

    def open(self, xfinal: float):
        self.lib.orc_sim_open(self.h, C.c_double(xfinal))

    def record(self, cap: int):
        self._recbuf = np.zeros(cap, dtype=abi.contact_dtype)
        self.lib.orc_sim_record(self.h, ptr(self._recbuf), cap)

    def recorded(self) -> np.ndarray:
        k = self.lib.orc_sim_recorded(self.h)
        assert k <= len(self._recbuf), "contact recorder overflow"
        return self._recbuf[:k].copy()

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_sim_delete(self.h)
            self.h = None


def brute_contacts(pp: np.ndarray, sys: abi.System) -> np.ndarray:
    """O(N^2) contact enumeration (independent of any cell grid), sorted (l, sm, sl, m)."""
    lib = oracle_lib()
    s = abi.System.from_buffer_copy(sys)
    s.npart = len(pp)
    k = lib.orc_brute_contacts(ptr(pp), len(pp), ptr(s), None, 0)
    buf = np.zeros(max(k, 1), dtype=abi.contact_dtype)
    lib.orc_brute_contacts(ptr(pp), len(pp), ptr(s), ptr(buf), len(buf))
    return buf[:k]


def oracle_cell_setup(pp: np.ndarray, sys: abi.System, geo: abi.Geometry, ntags: int) -> abi.Cell:
    cell = abi.Cell()
    cell.ntags = ntags
    s = abi.System.from_buffer_copy(sys)
    s.npart = len(pp)
    oracle_lib().orc_cell_setup(ptr(cell), ptr(pp), ptr(s), ptr(geo))
    return cell
