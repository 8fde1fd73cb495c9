"""Synthetic workloads of the benchmark configurations, in the reference's cgs units.

Clump: 4 equal spheres of radius r on a regular tetrahedron with centres d = 0.75 r from the
centroid (sphere_div = sqrt(3) r / d); density 2.5 g/cm^3.  Contact stiffness k_n is given
instead of the contact time: the reference's normal spring is km = 0.5 kappa m_eff
(force_part.cu:121), so kappa = 2 k_n / m_s and t_coll = sqrt((pi^2 + ln^2 eps) / kappa);
dt = t_coll / 50.

    small (S)     2,000 clumps, 4x4x8 cm release box on a 35 deg flume 4 cm wide, 30 cm long
    headline (H)  250,000 clumps, funnel release into a 35 deg flume 20 cm wide, 2 m long,
                  flat deposition zone
    large (L)     2,000,000 clumps, r ~ U[0.09, 0.11] cm per clump, hopper discharge onto an
                  incline, mu 0.6, eps 0.4
    batch         64 instances, N ~ log-uniform[5e3, 6e4], mu, eps ~ U[0.3, 0.7],
                  slope ~ U[28, 40] deg
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import abi, dem

R_SPH = 0.1        # cm
D_OFF = 0.075      # cm, centre offset
RHO = 2.5          # g/cm^3
K_N = 1.0e7        # dyn/cm (1e4 N/m)


def sphere_div(r: float = R_SPH, d: float = D_OFF) -> float:
    return math.sqrt(3.0) * r / d


def t_coll_for(k_n: float, eps: float, m_s: float) -> float:
    kappa = 2.0 * k_n / m_s
    return math.sqrt((math.pi ** 2 + math.log(eps) ** 2) / kappa)


def sphere_mass(r: float, rho: float = RHO) -> float:
    return abi.SPH_FACT * rho * r * r * r  # force_part.cu:58 form


@dataclass
class FlumeSpec:
    n: int
    tilt_deg: float = 35.0
    width: float = 4.0            # flume width side.y
    length: float = 30.0          # Lflume
    mu: float = 0.5
    eps: float = 0.5
    k_n: float = K_N
    rmin: float = R_SPH
    rmax: float = R_SPH
    box: tuple = (4.0, 4.0, 8.0)  # release box (x, y, z) lengths
    funnel_width: float | None = None  # YFUNNEL; None: no funnel
    funnel_length: float = 0.0         # LFUNNEL
    deposit_length: float = 0.0        # flat zone past the flume end (0: none)
    spacing: float = 0.4
    jitter: float = 0.005
    seed: int = 1
    dt_frac: float = 50.0
    # release layer resting on the slope (x' >= 0.2 cm, full width, layer_depth deep)
    # instead of the lattice in the release box
    slope_layer: bool = False
    layer_depth: float = 4.0


def build(spec: FlumeSpec) -> dem.Instance:
    """Config-mode instance: avalanche phase from t = 0 (gate open), zero velocity.

    The release box is the hopper of the geometry (x in [0, box x], y between yhop_l and
    yhop_r, above z = 0), but its walls act only while filling: released clumps leave the box
    underneath them, and force_hop.cu applies its walls to any sphere above the hopper
    bottom, so a clump pushed back up past that edge from outside would meet the wall penalty
    at full penetration.  The funnel planes, or without a funnel the flume's side walls
    (then along the whole slope, up-slope of the box too), confine the released flow."""
    geo = dem.reference_geometry()
    geo.sphere_div = sphere_div(spec.rmax, 0.75 * spec.rmax)
    geo.clump_mode = 1
    bx, by, bz = spec.box
    fw = spec.funnel_width if spec.funnel_width is not None else spec.width
    geo.yfunnel = fw
    # without a funnel the side walls start far up-slope (they act where x' > lfunnel)
    geo.lfunnel = spec.funnel_length if spec.funnel_width is not None else -1.0e6
    geo.hop_width = bx
    geo.yhop_l = 0.5 * (fw - by)
    geo.yhop_r = geo.yhop_l + by
    geo.hop_height = 0.0
    mods = abi.MOD_HOP | abi.MOD_WALLS | abi.MOD_PP | abi.MOD_FLUME
    if spec.funnel_width is not None and spec.funnel_length > 0:
        mods |= abi.MOD_FUNNEL | abi.MOD_FCORNERS
    tilt = math.radians(spec.tilt_deg)
    x_end = (spec.funnel_length + spec.length) * math.cos(tilt)
    if spec.deposit_length > 0:
        mods |= abi.MOD_DEP
        geo.dep_sin, geo.dep_cos = 0.0, 1.0
        geo.dep_x0, geo.dep_x1 = 0.0, 0.0
        geo.dep_z0 = -(spec.funnel_length + spec.length) * math.sin(tilt)
    geo.modules_fill = mods
    geo.modules_avalanche = mods & ~abi.MOD_HOP

    ms = sphere_mass(spec.rmin)
    t_c = t_coll_for(spec.k_n, spec.eps, ms)
    sys = abi.System()
    sys.npart = spec.n
    sys.side.y = spec.width
    sys.side.z = bz
    sys.rmin, sys.rmax = spec.rmin, spec.rmax
    sys.v0 = 0.0
    sys.dt = t_c / spec.dt_frac
    sys.eps_pw = sys.eps_pp = spec.eps
    sys.mu_pw = sys.mu_pp = spec.mu
    sys.t_coll = t_c
    sys.Lflume = spec.length
    sys.tilt = spec.tilt_deg
    fun = dem.derive_system(sys, geo)
    # gate open: avalanche phase from t = 0, clumps past the deposition zone removed
    sys.time = 0.0
    sys.side.x = x_end + spec.deposit_length * (1.0 if spec.deposit_length > 0 else 0.0)

    pp = dem.new_particles(spec.n)
    pp["density"] = RHO
    if spec.rmax > spec.rmin:
        rng = np.random.default_rng(spec.seed)
        pp["radius"] = rng.uniform(spec.rmin, spec.rmax, spec.n)
    else:
        pp["radius"] = spec.rmin
    # per-clump offset 0.75 r: sphere_div is the same for every radius
    dem.calculate_particle_parameters(pp, sys, geo)
    if spec.slope_layer:
        place_slope_layer(pp, geo, spec, tilt)
    else:
        lo = (0.0, geo.yhop_l, geo.hop_height)
        hi = (bx, geo.yhop_r, geo.hop_height + bz)
        dem.place_lattice(pp, geo, lo, hi, spec.spacing, spec.jitter, spec.seed)
    sys.tot_mass = float(pp["mass"].sum())
    return dem.Instance(sys=sys, geo=geo, fun=fun, particles=pp, ntags=9,
                        phase=abi.PHASE_AVALANCHE)


def place_slope_layer(pp, geo, spec: FlumeSpec, tilt: float) -> None:
    """Jittered lattice of randomly oriented clumps in the slope frame (x' down the slope,
    z' normal to it), resting just above the flume floor, mapped to the lab frame."""
    a, margin = spec.spacing, spec.jitter + 0.01
    ext = spec.rmax * (1.0 + math.sqrt(3.0) / geo.sphere_div)  # clump radius
    ny = int((spec.width - 2 * margin) / a + 1e-9)
    nz = int(spec.layer_depth / a + 1e-9)
    rows = math.ceil(spec.n / (ny * nz))
    z0 = ext + margin - 0.5 * a
    lo = (0.2, margin, z0)
    hi = (0.2 + rows * a + 1e-6, margin + ny * a + 1e-6, z0 + nz * a + 1e-6)
    dem.place_lattice(pp, geo, lo, hi, a, spec.jitter, spec.seed)
    c, s = math.cos(tilt), math.sin(tilt)
    q = pp["rr"].copy()
    pp["rr"][:, 0] = q[:, 0] * c + q[:, 2] * s
    pp["rr"][:, 2] = q[:, 2] * c - q[:, 0] * s
    dem.sphere_centres(pp, geo)


def slope_release(n: int = 20_000, seed: int = 1, width: float = 10.0, depth: float = 4.0,
                  mu: float = 0.5, eps: float = 0.5) -> dem.Instance:
    """A layer of n clumps released from rest on a 35 deg flume 10 cm wide and 1 m long
    with a flat 50 cm deposit (the differentiable-window workload)."""
    return build(FlumeSpec(n=n, width=width, length=100.0, box=(0.2, width, 1.0),
                           deposit_length=50.0, spacing=0.37, jitter=0.005, seed=seed,
                           slope_layer=True, layer_depth=depth, mu=mu, eps=eps))


def adjoint_batch(count: int = 16, n: int = 20_000, seed: int = 3) -> list[dem.Instance]:
    """The window benchmark: count slope releases of n clumps with different seeds."""
    return [slope_release(n=n, seed=seed * 100 + i) for i in range(count)]


def small(seed: int = 1) -> dem.Instance:
    """S: 2,000 clumps, 10x10x20 lattice at 0.4 cm in the 4x4x8 cm box."""
    return build(FlumeSpec(n=2000, seed=seed))


def headline(seed: int = 1, n: int = 250_000) -> dem.Instance:
    """H: 250,000 clumps in a 20 x 20 cm funnel-fed release box (100 layers at 0.4 cm,
    jitter +-5 % of the spacing), 35 deg flume 20 cm wide and 2 m long after a 40 cm
    funnel narrowing 40 -> 20 cm, flat 50 cm deposition zone."""
    layers = math.ceil(n / 2500)
    return build(FlumeSpec(n=n, width=20.0, length=200.0, box=(20.0, 20.0, 0.4 * layers + 0.01),
                           funnel_width=40.0, funnel_length=40.0, deposit_length=50.0,
                           spacing=0.4, jitter=0.02, seed=seed))


def large(seed: int = 1, n: int = 2_000_000, spacing: float = 0.44) -> dem.Instance:
    """L: 2,000,000 polydisperse clumps (r ~ U[0.09, 0.11]) in a 40 x 40 cm hopper above a
    wide incline; mu 0.6, eps 0.4; dt from the smallest pair.  spacing: the lattice pitch
    (0.44 cm keeps the clumps apart; below the clump diameter, 0.385 cm at most, neighbours
    overlap from the start)."""
    a = spacing
    per_layer = int(40.0 / a) ** 2
    layers = math.ceil(n / per_layer)
    return build(FlumeSpec(n=n, width=40.0, length=300.0, box=(40.0, 40.0, a * layers + 0.01),
                           mu=0.6, eps=0.4, rmin=0.09, rmax=0.11, spacing=a, jitter=0.02,
                           seed=seed))


def batch(count: int = 64, seed: int = 7, scale: float = 1.0) -> list[dem.Instance]:
    """Batched ensemble: per-instance sizes, seeds, friction, restitution and slope."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(round(math.exp(rng.uniform(math.log(5e3), math.log(6e4))) * scale))
        n = max(n, 8)
        mu, eps = rng.uniform(0.3, 0.7), rng.uniform(0.3, 0.7)
        tilt = rng.uniform(28.0, 40.0)
        side = 4.0 * max(1.0, math.sqrt(n / 2000.0))  # keep ~20 layers in a square box
        per = int(side / 0.4) ** 2
        layers = math.ceil(n / per)
        out.append(build(FlumeSpec(n=n, tilt_deg=tilt, width=side, length=30.0, mu=mu, eps=eps,
                                   box=(side, side, 0.4 * layers + 0.01), seed=seed * 1000 + i)))
    return out


def _slope_lattice_near_floor(pp, geo, sys, lo_xy, per, a, seed, gap):
    """A jittered lattice of per[0] x per[1] x per[2] clumps in the slope frame, its lowest
    sphere gap above the flume floor."""
    dem.place_lattice(pp, geo, (lo_xy[0], lo_xy[1], 1.0), (lo_xy[0] + per[0] * a,
                                                         lo_xy[1] + per[1] * a, 1.0 + per[2] * a),
                      a, 0.01, seed)
    c, s = sys.cost, sys.sint
    q = pp["rr"].copy()
    pp["rr"][:, 0] = q[:, 0] * c + q[:, 2] * s
    pp["rr"][:, 2] = q[:, 2] * c - q[:, 0] * s
    dem.sphere_centres(pp, geo)
    sph = np.concatenate([pp[f] for f in ("rr_a", "rr_b", "rr_c", "rr_d")])
    zmin = (sph[:, 0] * s + sph[:, 2] * c).min()
    shift = zmin - (sys.rmax + gap)
    pp["rr"][:, 0] -= shift * s
    pp["rr"][:, 2] -= shift * c


def colliding_packets(seed: int, per=(4, 4, 4), a: float = 0.34, v: float = 50.0) -> dem.Instance:
    """Two packets of clumps (lattices tighter than the clump diameter) meeting head-on at
    relative speed 2 v in free space (no gravity, far from every wall): a window dominated
    by particle-particle contacts."""
    n = 2 * per[0] * per[1] * per[2]
    inst = build(FlumeSpec(n=n, width=20.0, length=100.0, box=(8, 8, 8),
                                               seed=seed))
    inst.geo.grav = 0.0
    pp = inst.particles
    half = n // 2
    L = per[0] * a
    lo = np.array([0.3, 8.0, 5.0])
    for k, x0 in enumerate((0.0, L + 0.01)):
        part = pp[k * half:(k + 1) * half]
        dem.place_lattice(part, inst.geo, tuple(lo + [x0, 0, 0]),
                          tuple(lo + [x0 + L, per[1] * a, per[2] * a]), a, 0.01, seed * 7 + k)
        pp[k * half:(k + 1) * half] = part
    pp["vv"] = 0.0
    pp["vv"][:half, 0] = v
    pp["vv"][half:, 0] = -v
    dem.sphere_centres(pp, inst.geo)
    inst.sys.time = 0.01
    inst.sys.side.x = 1.0e9
    return inst


def impacting_packet(seed: int, per=(6, 6, 2), a: float = 0.34, v: float = 40.0) -> dem.Instance:
    """A packet of clumps moving down the 35 deg slope at v and into the floor at 0.4 v, its
    lowest sphere 0.004 cm above it: impacts, sliding friction and the neighbours' contacts."""
    n = per[0] * per[1] * per[2]
    inst = build(FlumeSpec(n=n, width=20.0, length=100.0, box=(8, 8, 8),
                                               deposit_length=50.0, seed=seed))
    pp = inst.particles
    _slope_lattice_near_floor(pp, inst.geo, inst.sys, (0.3, 5.0), per, a, seed, 0.004)
    c, s = inst.sys.cost, inst.sys.sint
    vs = np.array([v, 0.0, -0.4 * v])  # slope frame
    pp["vv"] = [vs[0] * c + vs[2] * s, 0.0, vs[2] * c - vs[0] * s]
    dem.sphere_centres(pp, inst.geo)
    inst.sys.time = 0.01
    inst.sys.side.x = 1.0e9
    return inst
