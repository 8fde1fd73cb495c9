"""ctypes / numpy mirror of include/avflume.h.

The records keep the reference's layouts (particle = 408 B, cabecera.hu:77-98;
system_par = 256 B, cabecera.hu:34-66; cell_par = 168 B, cabecera.hu:101-112;
Funnel = 24 B, cabecera.hu:69-74), so arrays written by this module are byte-compatible
with the reference host driver's arrays.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libavflume.so")

# Literals of cabecera.hu:5-12, as written.
PI = 3.1415926535897932
GYR = 4.4494897427831781
SPH_FACT = 4.188790204786391
ROOT2 = 1.414213562373095
GRAV = 981.0
DRAG = 1.895e-3

MOD_PP, MOD_HOP, MOD_FUNNEL, MOD_FCORNERS, MOD_WALLS, MOD_FLUME, MOD_DEP = (1 << i for i in range(7))
MODULES_FILL = MOD_PP | MOD_HOP
MODULES_AVALANCHE = MOD_HOP | MOD_FUNNEL | MOD_FCORNERS | MOD_WALLS | MOD_PP | MOD_FLUME
PHASE_FILL, PHASE_AVALANCHE = 0, 1

STATUS = {
    0: "AVF_OK", 1: "AVF_ERR_ARG", 2: "AVF_ERR_HIP", 3: "AVF_ERR_NOMEM",
    4: "AVF_ERR_CELL_NEG", 5: "AVF_ERR_CELL_FULL", 6: "AVF_ERR_NEIGH_OVERFLOW",
    7: "AVF_ERR_NOFIT", 8: "AVF_ERR_INPUT",
}
STATUS_CODES = {v: k for k, v in STATUS.items()}


class Double3(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class Long3(C.Structure):
    _fields_ = [("x", C.c_int64), ("y", C.c_int64), ("z", C.c_int64)]


_V3 = ("f8", (3,))
PARTICLE_FIELDS = ["radius", "mass", "density", "moment_of_inertia", "rr_a", "rr_b", "rr_c",
                   "rr_d", "rrot_a", "rrot_b", "rrot_c", "rrot_d", "rr", "rrot", "vv", "ww",
                   "Rmatx", "Rmaty", "Rmatz"]
particle_dtype = np.dtype([(f, "f8") if i < 3 else (f, *_V3) for i, f in enumerate(PARTICLE_FIELDS)])
assert particle_dtype.itemsize == 408

contact_dtype = np.dtype([("l", "i8"), ("m", "i8"), ("sl", "i4"), ("sm", "i4"),
                          ("repulsive", "i4"), ("reserved", "i4")])
assert contact_dtype.itemsize == 32


class System(C.Structure):
    _fields_ = [("npart", C.c_int64)] + [(n, C.c_double) for n in (
        "dt", "hdt", "rmin", "rmax", "v0", "time", "eps_pp", "eps_pw", "kappa_pp", "kappa_pw",
        "gamma_pp", "gamma_pw", "mu_pp", "mu_pw", "t_coll", "tot_mass", "tilt", "xhopper",
        "Lflume", "gyr_max", "fr_fact_pp", "fr_fact_pw", "sint", "cost", "sintg", "costg",
        "xfunnel", "dy2")] + [("side", Double3)]


class Funnel(C.Structure):
    _fields_ = [("sinfi", C.c_double), ("cosfi", C.c_double), ("tanfi", C.c_double)]


class Cell(C.Structure):
    _fields_ = [("ntags", C.c_int64), ("ncell3", C.c_int64), ("ncell", Long3),
                ("ntot", C.c_int64), ("cell_side", Double3), ("cs_inv", Double3),
                ("rr_max", Double3), ("rr_min", Double3), ("length", Double3)]


class Geometry(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "hop_height", "hop_width", "yfunnel", "lfunnel", "yhop_l", "yhop_r", "grav", "drag",
        "sphere_div", "cell_dist", "dep_sin", "dep_cos", "dep_x0", "dep_x1", "dep_z0",
        "regrid_ddt_steps")] + [(n, C.c_int32) for n in (
        "regrid_interval", "inter_flag", "clump_mode", "modules_fill", "modules_avalanche",
        "reserved")]


class InstanceDesc(C.Structure):
    _fields_ = [("sys", System), ("geo", Geometry), ("fun", Funnel), ("ntags", C.c_int64),
                ("phase", C.c_int32), ("reserved", C.c_int32)]


class Options(C.Structure):
    _fields_ = [("reorder_interval", C.c_int32), ("use_graph", C.c_int32),
                ("sphere_cell_factor", C.c_double), ("block_clumps", C.c_int32),
                ("max_neighbours", C.c_int32), ("device", C.c_int32), ("lean_state", C.c_int32),
                ("stream", C.c_void_p), ("verlet_skin", C.c_double),
                ("closed_forms", C.c_int32), ("step_binning", C.c_int32)]


class Diag(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "cell_below", "cell_beyond", "neigh_overflow", "max_neighbours", "max_cell_spheres",
        "steps", "list_builds", "steps_undone", "closed_instances", "sweep_global_bricks",
        "sweep_fp64_pairs")]


class Observables(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "vmean", "wmean", "xcm", "zcm", "runout_mean", "runout_front", "mass_live")] + \
        [("live", C.c_int64)]


class SphereGrid(C.Structure):
    _fields_ = [("lo", C.c_double * 3), ("h", C.c_double), ("inv_h", C.c_double),
                ("n", C.c_int32 * 3), ("reserved", C.c_int32)]


class ProfileSpec(C.Structure):
    _fields_ = [("x0", C.c_double), ("x1", C.c_double), ("width", C.c_double),
                ("sigma", C.c_double), ("nbins", C.c_int32), ("reserved", C.c_int32)]


class WindowSpec(C.Structure):
    _fields_ = [("steps", C.c_int64), ("checkpoint", C.c_int64), ("skin", C.c_double),
                ("w_runout", C.c_double), ("w_profile", C.c_double), ("profile", ProfileSpec)]


class WindowDiag(C.Structure):
    _fields_ = [("max_disp", C.c_double), ("max_list", C.c_int64), ("max_in", C.c_int64),
                ("missed", C.c_int64), ("list_valid", C.c_int32), ("reserved", C.c_int32), ("forward_ms", C.c_double),
                ("tangent_ms", C.c_double), ("backward_ms", C.c_double)]


# theta layout per instance (avf_window_run)
THETA = ["mu_pp", "eps_pp", "kappa_pp", "mu_pw", "eps_pw", "kappa_pw"]
NTHETA = len(THETA)


class Rng(C.Structure):
    _fields_ = [("useed", C.c_uint64 * 250), ("uuy", C.c_uint64), ("np", C.c_int32),
                ("nq", C.c_int32), ("npp", C.c_int32), ("not_init", C.c_int32)]


assert C.sizeof(System) == 256 and C.sizeof(Cell) == 168 and C.sizeof(Funnel) == 24


def reference_geometry() -> Geometry:
    """Reference values of the cabecera.hu:5-19 geometry (mirror of avf_geometry_reference)."""
    g = Geometry()
    g.hop_height, g.hop_width, g.yfunnel, g.lfunnel = 25.0, 15.0, 30.0, 70.0
    g.yhop_l, g.yhop_r, g.grav, g.drag = 7.5, 22.5, GRAV, DRAG
    g.sphere_div = ROOT2
    g.cell_dist = GYR + float(np.float32(0.05))  # `GYR + 0.05f`, avalanche_hst.cu:115
    g.dep_sin, g.dep_cos, g.dep_x0, g.dep_x1, g.dep_z0 = 0.0, 1.0, 0.0, 0.0, 0.0
    g.regrid_ddt_steps = 2000.0
    g.regrid_interval = 2000
    g.inter_flag = 1
    g.clump_mode = 0
    g.modules_fill = MODULES_FILL
    g.modules_avalanche = MODULES_AVALANCHE
    return g


def ptr(a) -> C.c_void_p:
    """Raw pointer of a numpy array or ctypes object (None -> NULL)."""
    if a is None:
        return C.c_void_p(0)
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return C.c_void_p(a.ctypes.data)
    return C.cast(C.pointer(a), C.c_void_p)


def v3(n: int) -> np.ndarray:
    """Zeroed (n, 3) float64 array (an array of double3)."""
    return np.zeros((n, 3), dtype=np.float64)


_LIB = None


def load_library() -> C.CDLL:
    """Loads libavflume.so; raises if it has not been built (no fallback path exists)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m avalancheflume3d_amd.build`")
        lib = C.CDLL(LIB_PATH)
        vp, i32, i64, d = C.c_void_p, C.c_int32, C.c_int64, C.c_double
        sig = {
            "avf_geometry_reference": (None, [vp]),
            "avf_options_default": (None, [vp]),
            "avf_rng_init": (None, [vp]),
            "avf_rannew64": (d, [vp, vp]),
            "avf_parse_datainput": (C.c_int, [C.c_char_p, vp, vp, vp]),
            "avf_derive_system": (None, [vp, vp, vp]),
            "avf_error_check": (C.c_int, [vp, C.c_char_p, C.c_int]),
            "avf_generate_particles": (None, [vp, vp, d]),
            "avf_calculate_particle_parameters": (None, [vp, vp, vp]),
            "avf_place_particles": (C.c_int, [vp, vp, vp]),
            "avf_set_initial_velocity": (None, [vp, vp, vp, i64]),
            "avf_place_lattice": (C.c_int, [vp, i64, vp, Double3, Double3, d, d, i64]),
            "avf_sphere_centres": (None, [vp, i64, vp]),
            "avf_cell_setup": (None, [vp, vp, vp, vp]),
            "avf_create": (C.c_int, [vp, vp, i32, vp]),
            "avf_destroy": (C.c_int, [vp]),
            "avf_last_error": (C.c_char_p, [vp]),
            "avf_upload": (C.c_int, [vp, i32, vp, vp, vp]),
            "avf_download": (C.c_int, [vp, i32, vp, vp, vp]),
            "avf_set_phase": (C.c_int, [vp, i32, i32, d, d, i64]),
            "avf_get_time": (C.c_int, [vp, i32, vp, vp]),
            "avf_regrid": (C.c_int, [vp]),
            "avf_get_cell": (C.c_int, [vp, i32, vp]),
            "avf_initial_forces": (C.c_int, [vp]),
            "avf_step": (C.c_int, [vp, i64]),
            "avf_get_forces": (C.c_int, [vp, i32, vp, vp]),
            "avf_get_clump_cells": (C.c_int, [vp, i32, vp]),
            "avf_check_cells": (C.c_int, [vp, i32, vp]),
            "avf_get_contacts": (C.c_int, [vp, i32, vp, i64, vp]),
            "avf_get_cell_table": (C.c_int, [vp, i32, vp, vp, i64, vp]),
            "avf_get_sphere_keys": (C.c_int, [vp, i32, vp, vp, vp, vp]),
            "avf_get_diag": (C.c_int, [vp, vp]),
            "avf_sync": (C.c_int, [vp]),
            "avf_step_profiled": (C.c_int, [vp, i64, vp, vp]),
            "avf_step_timed": (C.c_int, [vp, i64, i32, i32, vp, vp]),
            "avf_observe": (C.c_int, [vp, i32, vp, vp, vp]),
            "avf_observe_noalloc": (C.c_int, [vp, i32, vp, vp, vp]),
            "avf_window_run": (C.c_int, [vp] * 15),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


# Entry points declared by include/avflume.h (checked by the CPU test suite).
EXPORTED = [
    "avf_geometry_reference", "avf_options_default", "avf_rng_init", "avf_rannew64",
    "avf_parse_datainput", "avf_derive_system", "avf_error_check", "avf_generate_particles",
    "avf_calculate_particle_parameters", "avf_place_particles", "avf_set_initial_velocity",
    "avf_place_lattice", "avf_sphere_centres", "avf_cell_setup", "avf_create", "avf_destroy",
    "avf_last_error", "avf_upload", "avf_download", "avf_set_phase", "avf_get_time",
    "avf_regrid", "avf_get_cell", "avf_initial_forces", "avf_step", "avf_get_forces",
    "avf_get_clump_cells", "avf_check_cells", "avf_get_contacts", "avf_get_cell_table",
    "avf_get_sphere_keys", "avf_get_diag", "avf_sync",
    "avf_step_profiled", "avf_step_timed", "avf_observe", "avf_observe_noalloc",
    "avf_window_run",
]
KERNELS = ["predict", "scan_a", "scan_b", "scan_c", "scatter", "detect", "force", "cellreduce",
           "finalize", "count"]
