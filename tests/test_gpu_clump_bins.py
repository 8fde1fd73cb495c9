"""Per-step binning by clump (avf_options.step_binning 0, the default) against per-step
binning by sphere (step_binning 1), through the C ABI.

The clump sweep lists candidate pairings with a float test and a margin; the force stage
re-tests each in FP64 and applies the reference's visit rule, so both binnings must give
the same contacts in the same order: every bit of the state, the accelerations, the last
step's forces and torques, the time and the step counter.  The cases cover hopper filling,
the avalanche with its walls, unequal radii (the reference's bidisperse generator and
U[0.09, 0.11]), a flowing layer with spatial reorders and clump-cell regrids, pairs placed
at sep (1 +- 1e-12 .. 1e-4), and packings closer than the clump diameter (several clumps per
clump cell, many touching clumps per clump).
"""
import copy

import numpy as np
import pytest

from avalancheflume3d_amd import abi, workloads
from tests import scenarios
from tests.test_gpu_parity import STATE, gpu_ctx, state_of
from tests.test_gpu_sweep_edges import margin_pairs

pytestmark = pytest.mark.gpu


def flow_layer(seed, regrid=None):
    inst = workloads.build(workloads.FlumeSpec(n=300, width=2.0, length=4.0, box=(0.2, 2.0, 1.0),
                                               deposit_length=6.0, spacing=0.37, jitter=0.005,
                                               seed=seed, slope_layer=True, layer_depth=1.2))
    if regrid:
        inst = copy.copy(inst)
        inst.geo = abi.Geometry.from_buffer_copy(inst.geo)
        inst.geo.regrid_interval = regrid
    return inst


def same_run(inst, chunks, reorder=50):
    """Steps inst with both binnings; asserts every bit equal after every chunk.  Returns the
    clump-binning context's diagnostics."""
    a = gpu_ctx(inst, binning=0, reorder=reorder)
    b = gpu_ctx(inst, binning=1, reorder=reorder)
    a.initial_forces()
    b.initial_forces()
    for chunk in chunks:
        a.step(chunk)
        b.step(chunk)
        pa, aa, ga = a.download()
        pb, ab, gb = b.download()
        for fld in STATE:
            assert np.array_equal(pa[fld].view(np.uint64), pb[fld].view(np.uint64)), fld
        assert np.array_equal(aa.view(np.uint64), ab.view(np.uint64))
        assert np.array_equal(ga.view(np.uint64), gb.view(np.uint64))
        fa, ta = a.forces()
        fb, tb = b.forces()
        assert np.array_equal(fa.view(np.uint64), fb.view(np.uint64))
        assert np.array_equal(ta.view(np.uint64), tb.view(np.uint64))
        assert a.time() == b.time()
    d = a.diag()
    assert d["neigh_overflow"] == 0 and b.diag()["neigh_overflow"] == 0
    a.close()
    b.close()
    return d


@pytest.mark.parametrize("kind", ["hopper", "avalanche", "bidisperse", "poly"])
def test_clump_binning_equals_sphere_binning(gpu, kind):
    d = same_run(state_of(kind, 41), (1, 37, 120))
    print(kind, "max touching clumps", d["max_neighbours"])


def test_clump_binning_flowing_layer_reorders_and_regrids(gpu):
    """A flowing layer on the slope: reorders every 13 steps and a clump-cell regrid every
    7 (the keys, grid and lists follow the storage order and the device-side parameters)."""
    same_run(flow_layer(81, regrid=7), (3, 40, 200), reorder=13)


def test_clump_binning_at_the_float_margin(gpu):
    """Clump pairs whose closest spheres sit at sep (1 + delta), delta = +-1e-12 .. 1e-4:
    the listed superset re-tested in FP64 gives the sphere binning's contacts."""
    inst, deltas, touching = margin_pairs(33)
    assert np.array_equal(touching, deltas < 0)
    same_run(inst, (1, 1, 5))


@pytest.mark.parametrize("a", [0.26, 0.3])
def test_clump_binning_dense_packings(gpu, a):
    """Packings closer than the clump diameter (0.35 cm): several clumps per clump cell and
    many touching clumps per clump."""
    inst = workloads.colliding_packets(5, per=(10, 10, 10), a=a, v=20.0)
    d = same_run(inst, (1, 2, 3))
    print(f"spacing {a}: most touching clumps of a clump", d["max_neighbours"])
    assert d["max_neighbours"] >= 1


def test_clump_binning_small_configuration(gpu):
    """The small configuration (2,000 clumps) over 300 steps."""
    same_run(workloads.small(), (100, 200))
