"""Python front end of the C ABI (include/avflume.h): setup helpers and the device context.

The reference's operator set (cabecera.hu:115-157) is reached through one context object:
``Context.step`` runs the step loop of avalanche_hst.cu:668-710 (or :438-467 while filling)
on the GPU, ``upload``/``download`` move reference-layout ``particle`` arrays.  Every call
checks the status code and raises ``AvfError`` with the library's message.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import abi
from .abi import ptr


class AvfError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{abi.STATUS.get(code, code)}: {msg}")
        self.code = code


def lib() -> C.CDLL:
    return abi.load_library()


def _check(rc: int, ctx=None) -> None:
    if rc != 0:
        msg = lib().avf_last_error(ctx)
        raise AvfError(rc, msg.decode() if msg else "")


# --------------------------------------------------------------------------- setup ---

def reference_geometry() -> abi.Geometry:
    g = abi.Geometry()
    lib().avf_geometry_reference(ptr(g))
    return g


def default_options() -> abi.Options:
    o = abi.Options()
    lib().avf_options_default(ptr(o))
    return o


class Rng:
    """rannew64 (rannew64.cu:55-126) with explicit state."""

    def __init__(self):
        self.st = abi.Rng()
        lib().avf_rng_init(ptr(self.st))

    def __call__(self, idum: int) -> tuple[float, int]:
        d = C.c_int64(idum)
        v = lib().avf_rannew64(ptr(self.st), C.byref(d))
        return v, d.value


def parse_datainput(text: str) -> tuple[abi.System, int, int]:
    sys = abi.System()
    ntags, pf = C.c_int64(0), C.c_int32(0)
    _check(lib().avf_parse_datainput(text.encode(), ptr(sys), C.byref(ntags), C.byref(pf)))
    return sys, ntags.value, pf.value


def derive_system(sys: abi.System, geo: abi.Geometry) -> abi.Funnel:
    fun = abi.Funnel()
    lib().avf_derive_system(ptr(sys), ptr(fun), ptr(geo))
    return fun


def error_check(sys: abi.System) -> tuple[int, str]:
    buf = C.create_string_buffer(256)
    code = lib().avf_error_check(ptr(sys), buf, 256)
    return code, buf.value.decode()


def new_particles(n: int) -> np.ndarray:
    return np.zeros(n, dtype=abi.particle_dtype)


def generate_particles(pp: np.ndarray, sys: abi.System, density: float = 2.6) -> None:
    lib().avf_generate_particles(ptr(pp), ptr(sys), C.c_double(density))


def calculate_particle_parameters(pp: np.ndarray, sys: abi.System, geo: abi.Geometry) -> None:
    lib().avf_calculate_particle_parameters(ptr(pp), ptr(sys), ptr(geo))


def place_particles(pp: np.ndarray, sys: abi.System, geo: abi.Geometry) -> None:
    _check(lib().avf_place_particles(ptr(pp), ptr(sys), ptr(geo)))


def set_initial_velocity(pp: np.ndarray, sys: abi.System, rng: Rng | None = None,
                         idum0: int = 0) -> Rng:
    rng = rng or Rng()
    lib().avf_set_initial_velocity(ptr(pp), ptr(sys), ptr(rng.st), C.c_int64(idum0))
    return rng


def place_lattice(pp: np.ndarray, geo: abi.Geometry, lo, hi, spacing: float, jitter: float,
                  seed: int) -> None:
    _check(lib().avf_place_lattice(ptr(pp), C.c_int64(len(pp)), ptr(geo), abi.Double3(*lo),
                                   abi.Double3(*hi), C.c_double(spacing), C.c_double(jitter),
                                   C.c_int64(seed)))


def sphere_centres(pp: np.ndarray, geo: abi.Geometry) -> None:
    lib().avf_sphere_centres(ptr(pp), C.c_int64(len(pp)), ptr(geo))


def cell_setup(pp: np.ndarray, sys: abi.System, geo: abi.Geometry, ntags: int) -> abi.Cell:
    cell = abi.Cell()
    cell.ntags = ntags
    lib().avf_cell_setup(ptr(cell), ptr(pp), ptr(sys), ptr(geo))
    return cell


# ------------------------------------------------------------------------- context ---

@dataclass
class Instance:
    """One independent avalanche: its system, geometry, funnel and clump records."""
    sys: abi.System
    geo: abi.Geometry
    fun: abi.Funnel
    particles: np.ndarray
    ntags: int = 9
    phase: int = abi.PHASE_FILL
    aa_i: np.ndarray | None = None
    anga_i: np.ndarray | None = None
    extra: dict = field(default_factory=dict)

    def desc(self) -> abi.InstanceDesc:
        d = abi.InstanceDesc()
        d.sys = self.sys
        d.sys.npart = len(self.particles)
        d.geo = self.geo
        d.fun = self.fun
        d.ntags = self.ntags
        d.phase = self.phase
        return d


class Context:
    """Device context stepping one or more instances in the same kernels."""

    def __init__(self, instances: list[Instance], options: abi.Options | None = None,
                 upload: bool = True):
        self.instances = list(instances)
        descs = (abi.InstanceDesc * len(self.instances))(*[i.desc() for i in self.instances])
        self.options = options or default_options()
        h = C.c_void_p()
        _check(lib().avf_create(C.byref(h), descs, len(self.instances), ptr(self.options)))
        self.h = h
        if upload:
            for k, inst in enumerate(self.instances):
                self.upload(k, inst.particles, inst.aa_i, inst.anga_i)

    def _ck(self, rc: int) -> None:
        _check(rc, self.h)

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().avf_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def n(self, inst: int = 0) -> int:
        return len(self.instances[inst].particles)

    def upload(self, inst: int, pp: np.ndarray, aa=None, anga=None) -> None:
        n = self.n(inst)
        if aa is None:
            aa = abi.v3(n)
        if anga is None:
            anga = abi.v3(n)
        pp = np.ascontiguousarray(pp, dtype=abi.particle_dtype)
        self._ck(lib().avf_upload(self.h, inst, ptr(pp), ptr(np.ascontiguousarray(aa, np.float64)),
                                  ptr(np.ascontiguousarray(anga, np.float64))))

    def download(self, inst: int = 0):
        n = self.n(inst)
        pp = np.zeros(n, dtype=abi.particle_dtype)
        aa, anga = abi.v3(n), abi.v3(n)
        self._ck(lib().avf_download(self.h, inst, ptr(pp), ptr(aa), ptr(anga)))
        return pp, aa, anga

    def set_phase(self, inst: int, phase: int, time: float, side_x: float, cc: int = 0) -> None:
        self._ck(lib().avf_set_phase(self.h, inst, phase, C.c_double(time), C.c_double(side_x),
                                     C.c_int64(cc)))

    def time(self, inst: int = 0) -> tuple[float, int]:
        t, cc = C.c_double(), C.c_int64()
        self._ck(lib().avf_get_time(self.h, inst, C.byref(t), C.byref(cc)))
        return t.value, cc.value

    def regrid(self) -> None:
        self._ck(lib().avf_regrid(self.h))

    def cell(self, inst: int = 0) -> abi.Cell:
        c = abi.Cell()
        self._ck(lib().avf_get_cell(self.h, inst, ptr(c)))
        return c

    def initial_forces(self) -> None:
        self._ck(lib().avf_initial_forces(self.h))

    def step(self, n: int = 1) -> None:
        self._ck(lib().avf_step(self.h, C.c_int64(n)))

    def step_profiled(self, n: int) -> tuple[dict, float]:
        """n eager steps with HIP events between kernels: ({kernel: mean ms}, total ms)."""
        ks = np.zeros(len(abi.KERNELS), dtype=np.float64)
        tot = C.c_double()
        self._ck(lib().avf_step_profiled(self.h, C.c_int64(n), ptr(ks), C.byref(tot)))
        return dict(zip(abi.KERNELS, ks.tolist())), tot.value

    def step_timed(self, n: int, kernel: str, every: int = 1) -> tuple[float, float]:
        """n steps with HIP events around one kernel on every `every`-th step only:
        (its mean ms per timed launch, total ms)."""
        km, tot = C.c_double(), C.c_double()
        self._ck(lib().avf_step_timed(self.h, C.c_int64(n), abi.KERNELS.index(kernel),
                                      C.c_int32(every), C.byref(km), C.byref(tot)))
        return km.value, tot.value

    def sync(self) -> None:
        self._ck(lib().avf_sync(self.h))

    def forces(self, inst: int = 0):
        n = self.n(inst)
        f, t = abi.v3(n), abi.v3(n)
        self._ck(lib().avf_get_forces(self.h, inst, ptr(f), ptr(t)))
        return f, t

    def clump_cells(self, inst: int = 0) -> np.ndarray:
        out = np.zeros(self.n(inst), dtype=np.int64)
        self._ck(lib().avf_get_clump_cells(self.h, inst, ptr(out)))
        return out

    def check_cells(self, inst: int = 0) -> tuple[int, int]:
        bad = C.c_int64(-1)
        rc = lib().avf_check_cells(self.h, inst, C.byref(bad))
        if rc not in (0, 4, 5):
            self._ck(rc)
        return rc, bad.value

    def cell_table(self, inst: int = 0) -> tuple[np.ndarray, np.ndarray]:
        """The reference's (cell_vec, nocup) of the current state (cells.cu:11-111)."""
        cl = self.cell(inst)
        ncell3 = cl.ncell.x * cl.ncell.y * cl.ncell.z
        cell_vec = np.zeros(ncell3 * self.instances[inst].ntags, dtype=np.int64)
        nocup = np.zeros(ncell3, dtype=np.int32)
        bad = C.c_int64(-1)
        self._ck(lib().avf_get_cell_table(self.h, inst, ptr(cell_vec), ptr(nocup),
                                          C.c_int64(ncell3), C.byref(bad)))
        return cell_vec, nocup

    def sphere_keys(self, inst: int = 0):
        """(keys per sphere 4 m + s, sort order of the binned sphere ids, sphere grid)."""
        n = self.n(inst)
        keys = np.zeros(4 * n, dtype=np.uint32)
        order = np.zeros(4 * n, dtype=np.int64)
        nb = C.c_int64()
        grid = abi.SphereGrid()
        self._ck(lib().avf_get_sphere_keys(self.h, inst, ptr(keys), ptr(order), C.byref(nb),
                                           C.byref(grid)))
        return keys, order[:nb.value], grid

    def contacts(self, inst: int = 0, cap: int | None = None) -> np.ndarray:
        cap = cap if cap is not None else 64 * self.n(inst)
        out = np.zeros(cap, dtype=abi.contact_dtype)
        cnt = C.c_int64()
        self._ck(lib().avf_get_contacts(self.h, inst, ptr(out), C.c_int64(cap), C.byref(cnt)))
        if cnt.value > cap:
            return self.contacts(inst, cnt.value)
        return out[:cnt.value]

    def observe(self, inst: int = 0, profile: abi.ProfileSpec | None = None):
        """Reference stats (vmean, wmean, xcm, zcm), runout and the deposit profile."""
        o = abi.Observables()
        prof = np.zeros(profile.nbins, np.float64) if profile is not None else None
        self._ck(lib().avf_observe(self.h, inst, ptr(o), ptr(profile) if profile else None,
                                   ptr(prof)))
        d = {k: getattr(o, k) for k, _ in abi.Observables._fields_}
        return (d, prof) if profile is not None else d

    def window_run(self, steps: int, checkpoint: int = 100, skin: float = 0.0,
                   w_runout: float = 1.0, w_profile: float = 0.0,
                   profile: abi.ProfileSpec | None = None, theta=None, dv=None, target=None,
                   grad: bool = True, tan_theta=None, tan_dv=None, tangent: bool = False,
                   final_state: bool = False) -> dict:
        """Differentiable window from the current state (left unchanged): the loss per
        instance, optional gradients wrt theta (ninst x 6: mu, eps, kappa for pp and pw)
        and the start-velocity perturbation dv (clumps x 3, instance-major reference order),
        and optionally the dual-number directional derivative along (tan_theta, tan_dv)."""
        ninst = len(self.instances)
        ntot = sum(self.n(i) for i in range(ninst))
        spec = abi.WindowSpec()
        spec.steps, spec.checkpoint, spec.skin = steps, checkpoint, skin
        spec.w_runout, spec.w_profile = w_runout, w_profile
        nbins = 0
        if profile is not None:
            spec.profile = profile
            nbins = profile.nbins

        def arr(a, shape):
            if a is None:
                return None
            return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(shape))

        theta_a = arr(theta, (ninst, abi.NTHETA))
        dv_a = arr(dv, (ntot, 3))
        tgt_a = arr(target, (ninst, max(nbins, 1))) if target is not None else None
        tt_a = arr(tan_theta, (ninst, abi.NTHETA))
        tdv_a = arr(tan_dv, (ntot, 3))
        loss = np.zeros(ninst)
        obs = (abi.Observables * ninst)()
        prof = np.zeros((ninst, max(nbins, 1)))
        gth = np.zeros((ninst, abi.NTHETA)) if grad else None
        gdv = np.zeros((ntot, 3)) if grad else None
        dl = np.zeros(ninst) if tangent else None
        fs = np.zeros((ntot, 24)) if final_state else None
        dg = abi.WindowDiag()
        self._ck(lib().avf_window_run(self.h, ptr(spec), ptr(theta_a), ptr(dv_a), ptr(tgt_a),
                                      ptr(loss), ptr(obs), ptr(prof), ptr(gth), ptr(gdv), ptr(tt_a),
                                      ptr(tdv_a), ptr(dl), ptr(fs), ptr(dg)))
        out = {"loss": loss, "profile": prof[:, :nbins],
               "obs": [{k: getattr(o, k) for k, _ in abi.Observables._fields_} for o in obs],
               "diag": {k: getattr(dg, k) for k, _ in abi.WindowDiag._fields_}}
        if grad:
            out["grad_theta"], out["grad_dv"] = gth, gdv
        if tangent:
            out["dloss"] = dl
        if final_state:
            out["final_state"] = fs
        return out

    def diag(self) -> dict:
        d = abi.Diag()
        self._ck(lib().avf_get_diag(self.h, ptr(d)))
        return {k: getattr(d, k) for k, _ in abi.Diag._fields_}
