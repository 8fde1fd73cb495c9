"""Autotuner of the launch parameters that do not change results.

The candidate-list skin, the sphere-cell size (a multiple of the largest sphere diameter),
the spatial reorder interval and the block shape of the force kernel change only speed:
every per-clump sum is sequential in the reference order and the candidate lists always hold
every touching pair, so results are bit-identical across them.  The tuner times a short run
of each candidate on the caller's instances (through the captured-graph path) and returns the
fastest options.
"""
from __future__ import annotations

import itertools
import time

from . import dem

SKINS = (0.0, 0.005, 0.01, 0.02)   # cm; 0: sphere cells rebuilt every step
FACTORS = (1.0, 1.25)
REORDERS = (50, 200, 400)
BLOCKS = (64, 128)


def autotune(instances: list[dem.Instance], steps: int = 200, warmup: int = 20,
             skins=SKINS, factors=FACTORS, reorders=REORDERS, blocks=BLOCKS,
             verbose: bool = False):
    """Returns (best avf_options, [(skin, factor, reorder, block, seconds_per_step), ...])."""
    table = []
    best, best_t = None, float("inf")
    for sk, f, r, b in itertools.product(skins, factors, reorders, blocks):
        o = dem.default_options()
        o.verlet_skin, o.sphere_cell_factor, o.reorder_interval, o.block_clumps = sk, f, r, b
        ctx = dem.Context(instances, o)
        ctx.regrid()
        ctx.initial_forces()
        ctx.step(warmup)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step(steps)
        ctx.sync()
        dt = (time.perf_counter() - t0) / steps
        ctx.close()
        table.append((sk, f, r, b, dt))
        if verbose:
            print(f"skin {sk:.3f} factor {f:.2f} reorder {r:4d} block {b:4d}: "
                  f"{dt * 1e6:9.1f} us/step", flush=True)
        if dt < best_t:
            best, best_t = o, dt
    return best, table
