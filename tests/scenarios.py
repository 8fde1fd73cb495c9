"""Seeded test states for the parity tests (reference geometry, cgs units).

Every builder returns an Instance whose records are self-consistent the way the reference
keeps them (sphere centres from rr, R and the radius, avalanche_hst.cu:265-293; slope-frame
copies from rotate_vectors.cu:28-46), so the oracle and the GPU start from identical bytes.
"""
from __future__ import annotations

import os

import numpy as np

from avalancheflume3d_amd import abi, dem

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DATAINPUT = open(os.path.join(GOLDEN, "DataInput.dat")).read()


def datainput(npart: int, **kv) -> str:
    """The reference DataInput.dat with npart (and optionally other lines) replaced."""
    lines = DATAINPUT.splitlines()
    order = ["npart", "side_y", "side_z", "rmin", "rmax", "v0", "dt", "eps_pw", "mu_pw",
             "eps_pp", "mu_pp", "t_coll", "Lflume", "tilt", "ntags", "print_flag"]
    vals = {"npart": npart, **kv}
    for i, name in enumerate(order):
        if name in vals:
            lines[i] = f"{vals[name]}   {name}"
    return "\n".join(lines) + "\n"


def reference_system(npart: int, **kv):
    sys, ntags, _ = dem.parse_datainput(datainput(npart, **kv))
    geo = dem.reference_geometry()
    fun = dem.derive_system(sys, geo)
    return sys, geo, fun, ntags


def rotate_vectors(pp: np.ndarray, sys: abi.System) -> None:
    """rotate_vectors.cu:28-46 in numpy (same operation order)."""
    s, c = sys.sint, sys.cost
    for a, b in (("rr", "rrot"), ("rr_a", "rrot_a"), ("rr_b", "rrot_b"), ("rr_c", "rrot_c"),
                 ("rr_d", "rrot_d")):
        v = pp[a]
        out = np.empty_like(v)
        out[:, 0] = v[:, 0] * c - v[:, 2] * s
        out[:, 1] = v[:, 1]
        out[:, 2] = v[:, 0] * s + v[:, 2] * c
        pp[b] = out


def _finish_records(pp, sys, geo):
    dem.sphere_centres(pp, geo)
    rotate_vectors(pp, sys)


def hopper_state(n: int, seed: int, squeeze: float = 0.93, spin: float = 40.0,
                 npart_sys: int | None = None, rmin: float | None = None) -> dem.Instance:
    """Reference filling state: generate_particles + calculate_particle_parameters +
    place_particles + set_initial_velocity, then the lattice squeezed towards its centre so
    neighbouring clumps overlap, random orientations, and random spins.  rmin below the
    input's rmax gives the reference's bidisperse radii (generate_particles.cu:26: every
    20th clump rmax, the others rmin)."""
    sys, geo, fun, ntags = reference_system(n) if rmin is None else reference_system(n, rmin=rmin)
    pp = dem.new_particles(n)
    dem.generate_particles(pp, sys)
    dem.calculate_particle_parameters(pp, sys, geo)
    dem.place_particles(pp, sys, geo)
    dem.set_initial_velocity(pp, sys)
    rng = np.random.default_rng(seed)
    c = pp["rr"].mean(axis=0)
    pp["rr"] = c + (pp["rr"] - c) * squeeze
    # keep the bottom layer just above the gate and the sides near the walls
    pp["rr"][:, 2] += 25.0 + 0.12 - pp["rr"][:, 2].min()
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    pp["Rmatx"] = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], 1)
    pp["Rmaty"] = np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], 1)
    pp["Rmatz"] = np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1)
    pp["ww"] = rng.normal(scale=spin, size=(n, 3))
    _finish_records(pp, sys, geo)
    return dem.Instance(sys=sys, geo=geo, fun=fun, particles=pp, ntags=ntags,
                        phase=abi.PHASE_FILL)


def _slope_to_lab(q: np.ndarray, sys: abi.System) -> np.ndarray:
    s, c = sys.sint, sys.cost
    out = np.empty_like(q)
    out[:, 0] = q[:, 0] * c + q[:, 2] * s
    out[:, 1] = q[:, 1]
    out[:, 2] = q[:, 2] * c - q[:, 0] * s
    return out


def avalanche_state(n_per_region: int, seed: int, time: float = 0.01,
                    spacing_factor: float = 0.88) -> dem.Instance:
    """Avalanche-phase state (gate open, side.x = xfinal) with clumps placed to touch every
    boundary module: the hopper x/y walls, the funnel planes, the funnel-exit corners, the
    flume side walls and floor, plus clumps past xfinal (frozen)."""
    regions_slope = [
        # (lo, hi) in the slope frame (x', y, z')
        ((20.0, 1.2, -0.3), (24.0, 4.2, 1.4)),     # left funnel plane + floor
        ((20.0, 25.8, -0.3), (24.0, 28.8, 1.4)),   # right funnel plane + floor
        ((68.4, 6.6, -0.3), (71.6, 9.0, 1.4)),     # left exit corner region
        ((68.4, 21.0, -0.3), (71.6, 23.4, 1.4)),   # right exit corner region
        ((100.0, 7.3, -0.3), (104.0, 10.0, 1.4)),  # left flume wall
        ((100.0, 20.0, -0.3), (104.0, 22.7, 1.4)), # right flume wall
        ((212.0, 10.0, -0.3), (215.2, 13.2, 1.4)), # past xfinal: frozen
    ]
    n_corner = 8
    n = n_per_region * (len(regions_slope) + 1) + 2 * n_corner
    sys, geo, fun, ntags = reference_system(n)
    rmax = sys.rmax
    a = spacing_factor * 2.0 * (rmax + rmax * np.sqrt(3.0) / geo.sphere_div)
    pp = dem.new_particles(n)
    dem.generate_particles(pp, sys)
    dem.calculate_particle_parameters(pp, sys, geo)
    rng = np.random.default_rng(seed)
    k = 0
    for r, (lo, hi) in enumerate(regions_slope):
        part = pp[k:k + n_per_region]
        dem.place_lattice(part, geo, lo, hi, a, 0.02 * a, seed * 31 + r + 1)
        part["rr"] = _slope_to_lab(part["rr"], sys)
        pp[k:k + n_per_region] = part
        k += n_per_region
    # hopper corner (lab frame): rear + left walls
    part = pp[k:k + n_per_region]
    dem.place_lattice(part, geo, (-0.1, 7.4, 25.05), (3.2, 10.6, 26.8), a, 0.02 * a, seed * 31 + 99)
    pp[k:k + n_per_region] = part
    k += n_per_region
    # columns of clumps straddling the two vertical corner edges at the funnel exit
    # (x' = LFUNNEL, y = dy2 and y = side.y + dy2)
    for yc in (sys.dy2, sys.side.y + sys.dy2):
        part = pp[k:k + n_corner]
        dem.place_lattice(part, geo, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0 * n_corner), 1.0, 0.0,
                          seed * 31 + int(yc))
        off = rng.uniform(-0.22, 0.22, size=(n_corner, 2))
        q = np.stack([geo.lfunnel + off[:, 0], yc + off[:, 1],
                      0.3 + 0.85 * np.arange(n_corner)], 1)
        part["rr"] = _slope_to_lab(q, sys)
        pp[k:k + n_corner] = part
        k += n_corner
    pp["vv"] = rng.normal(scale=20.0, size=(n, 3))
    pp["ww"] = rng.normal(scale=60.0, size=(n, 3))
    _finish_records(pp, sys, geo)
    sys.side.x = (geo.lfunnel + sys.Lflume) * sys.cost  # xfinal, avalanche_hst.cu:116,607
    sys.time = time
    return dem.Instance(sys=sys, geo=geo, fun=fun, particles=pp, ntags=max(ntags, 40),
                        phase=abi.PHASE_AVALANCHE)


def poly_flume(n: int, seed: int) -> dem.Instance:
    """The large configuration's material and clumps at a small size: radii U[0.09, 0.11] cm
    per clump (offset 0.75 r), mu 0.6, eps 0.4, a lattice tighter than the clump diameter in
    a release box above a 35 deg flume, so neighbours of unequal radii overlap."""
    from avalancheflume3d_amd import workloads
    side = 4.0
    per = int(side / 0.34) ** 2
    return workloads.build(workloads.FlumeSpec(
        n=n, width=side, length=30.0, box=(side, side, 0.34 * ((n + per - 1) // per) + 0.01),
        mu=0.6, eps=0.4, rmin=0.09, rmax=0.11, spacing=0.34, jitter=0.02, seed=seed))
