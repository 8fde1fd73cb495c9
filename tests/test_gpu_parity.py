"""GPU parity: the gfx950 path (through the C ABI) against the CPU oracle on seeded states.

Bars: contact sets and clump-cell assignment bit-exact; per-step F/T
within 1e-12 relative; 100-step trajectories within 1e-9 relative.  The kernels keep the
reference's operation order, so the tests also report whether results are bit-identical.
"""
import numpy as np
import pytest

from avalancheflume3d_amd import abi, dem
from oracle import oracle
from tests import scenarios

pytestmark = pytest.mark.gpu

F_RTOL = 1e-12   # per-step force / torque (every compared step)
T_RTOL = 1e-9    # trajectories over 100 steps
STATE = ["rr", "vv", "ww", "Rmatx", "Rmaty", "Rmatz", "rr_a", "rr_b", "rr_c", "rr_d", "radius"]


def opts(graph=1, reorder=50, factor=1.0, skin=0.0, lean=0, closed=0, binning=0):
    o = dem.default_options()
    o.step_binning = binning
    o.use_graph = graph
    o.reorder_interval = reorder
    o.sphere_cell_factor = factor
    o.verlet_skin = skin
    o.lean_state = lean
    o.closed_forms = closed
    return o


def gpu_ctx(insts, **kw):
    ctx = dem.Context(insts if isinstance(insts, list) else [insts], opts(**kw))
    ctx.regrid()  # the initial clump-cell grid, avalanche_hst.cu:311-333
    return ctx


def oracle_sim(inst):
    return oracle.OracleSim(inst.sys, inst.geo, inst.fun, inst.ntags, inst.phase, inst.particles)


def vec_close(a, b, rtol, what):
    """|a - b| <= rtol * |b| row-wise (vector norms), NaN-aware."""
    a = np.asarray(a, np.float64).reshape(len(a), -1)
    b = np.asarray(b, np.float64).reshape(len(b), -1)
    nan_ok = np.isnan(a).all(1) & np.isnan(b).all(1)
    err = np.linalg.norm(a - b, axis=1)
    scale = np.maximum(np.linalg.norm(b, axis=1), 1e-300)
    bad = ~nan_ok & ~(err <= rtol * scale)
    assert not bad.any(), f"{what}: {bad.sum()} rows off, worst rel {np.nanmax(err / scale):.3e}"
    return bool((a.view(np.uint64) == b.view(np.uint64)).all())


def contact_keys(c):
    return set(zip(c["l"].tolist(), c["m"].tolist(), c["sl"].tolist(), c["sm"].tolist(),
                   c["repulsive"].tolist()))


def state_of(kind, seed):
    """hopper / avalanche: equal radii; bidisperse: the reference's generator with
    rmin < rmax; poly: radii U[0.09, 0.11] (the large configuration's clumps)."""
    if kind == "hopper":
        return scenarios.hopper_state(400, seed)
    if kind == "bidisperse":
        return scenarios.hopper_state(400, seed, rmin=0.12)
    if kind == "poly":
        return scenarios.poly_flume(1500, seed)
    return scenarios.avalanche_state(16, seed)


KINDS = ["hopper", "avalanche", "bidisperse", "poly"]


@pytest.mark.parametrize("kind", KINDS)
def test_contact_set_and_cells_bit_exact(gpu, kind):
    inst = state_of(kind, 11 if kind != "avalanche" else 7)
    ctx = gpu_ctx(inst)
    ctx.initial_forces()
    got = ctx.contacts()
    o = oracle_sim(inst)
    o.record(200000)
    o.forces()
    ref = o.recorded()
    assert len(ref) > 20
    assert contact_keys(got) == contact_keys(ref)
    if kind in ("bidisperse", "poly"):  # contacts between spheres of unequal radii
        rad = inst.particles["radius"]
        assert (rad[ref["l"]] != rad[ref["m"]]).sum() > 10
    brute = oracle.brute_contacts(inst.particles, inst.sys)
    assert set(zip(brute["l"], brute["m"], brute["sl"], brute["sm"])) == \
        set(zip(ref["l"], ref["m"], ref["sl"], ref["sm"]))
    # clump-cell assignment: the cell_locate index (cells.cu:69-91) of every clump
    cell = o.get()["cell"]
    pp = inst.particles
    ii = np.trunc(cell.cs_inv.x * (pp["rr"][:, 0] - cell.rr_min.x)).astype(np.int64)
    jj = np.trunc(cell.cs_inv.y * (pp["rr"][:, 1] - cell.rr_min.y)).astype(np.int64)
    kk = np.trunc(cell.cs_inv.z * (pp["rr"][:, 2] - cell.rr_min.z)).astype(np.int64)
    ii[ii == cell.ncell.x] -= 1
    jj[jj == cell.ncell.y] -= 1
    kk[kk == cell.ncell.z] -= 1
    want = ii + cell.ncell.x * (jj + kk * cell.ncell.y)
    frozen = (inst.sys.time > 0) & (pp["rr"][:, 0] > inst.sys.side.x)
    want[frozen] = -1
    assert np.array_equal(ctx.clump_cells(), want)
    ctx.close()


@pytest.mark.parametrize("kind", ["hopper", "avalanche"])
def test_reference_cell_table_and_sphere_keys(gpu, kind):
    """The debug exports: nocup / cell_vec in the reference layout against the oracle's
    cell_locate on the same state and grid, and the sphere keys and sort order of the
    binning against the grid arithmetic applied to the downloaded sphere centres."""
    inst = scenarios.hopper_state(400, 13) if kind == "hopper" else scenarios.avalanche_state(16, 9)
    ctx = gpu_ctx(inst)
    ctx.initial_forces()
    cell_vec, nocup = ctx.cell_table()
    pp = ctx.download()[0]
    sys = abi.System.from_buffer_copy(inst.sys)
    sys.time = ctx.time()[0]
    want_vec, want_oc, rc, bad = oracle.cell_table(ctx.cell(), sys, pp)
    assert rc == 0, (rc, bad)
    assert np.array_equal(nocup, want_oc)
    assert np.array_equal(cell_vec, want_vec)
    assert nocup.sum() > 0
    keys, order, grid = ctx.sphere_keys()
    frozen = (sys.time > 0) & (pp["rr"][:, 0] > sys.side.x)
    want = np.empty((len(pp), 4), dtype=np.int64)
    for s_, fld in enumerate(("rr_a", "rr_b", "rr_c", "rr_d")):
        idx = []
        for ax in range(3):
            f = (pp[fld][:, ax] - grid.lo[ax]) * grid.inv_h
            i = np.where(f >= grid.n[ax], grid.n[ax] - 1, np.trunc(np.where(f >= 0, f, 0)))
            idx.append(i.astype(np.int64))
        want[:, s_] = idx[0] + grid.n[0] * (idx[1] + grid.n[1] * idx[2])
    want[frozen] = 0xffffffff
    assert np.array_equal(keys.astype(np.int64), want.reshape(-1))
    binned = np.flatnonzero(want.reshape(-1) != 0xffffffff)
    assert np.array_equal(order, binned[np.lexsort((binned, want.reshape(-1)[binned]))])
    ctx.step(3)  # candidate lists are off here: per-step binning keeps the export valid
    assert ctx.sphere_keys()[1].size > 0
    ctx.close()


@pytest.mark.parametrize("kind", KINDS)
def test_initial_forces(gpu, kind):
    inst = state_of(kind, 12 if kind != "avalanche" else 8)
    ctx = gpu_ctx(inst)
    ctx.initial_forces()
    f, t = ctx.forces()
    _, aa, anga = ctx.download()
    o = oracle_sim(inst)
    o.initial()
    r = o.get()
    assert np.abs(r["ff"]).max() > 0
    exact = [vec_close(f, r["ff"], F_RTOL, "force"), vec_close(t, r["tq"], F_RTOL, "torque")]
    live = ~((inst.sys.time > 0) & (inst.particles["rr"][:, 0] > inst.sys.side.x))
    exact.append(vec_close(aa[live], r["aa_i"][live], F_RTOL, "accel"))
    exact.append(vec_close(anga[live], r["anga_i"][live], F_RTOL, "ang accel"))
    print(f"bit-exact initial forces ({kind}):", all(exact))
    ctx.close()


def run_pair(inst, steps, chunk, **kw):
    ctx = gpu_ctx(inst, **kw)
    ctx.initial_forces()
    o = oracle_sim(inst)
    o.initial()
    done, exact = 0, True
    while done < steps:
        k = min(chunk, steps - done)
        ctx.step(k)
        o.step(k)
        done += k
        pp, aa, anga = ctx.download()
        f, t = ctx.forces()
        r = o.get()
        for fld in STATE:
            exact &= vec_close(pp[fld], r["pp"][fld], T_RTOL, f"{fld} after {done} steps")
        exact &= vec_close(f, r["ff"], F_RTOL, f"force after {done} steps")
        exact &= vec_close(t, r["tq"], F_RTOL, f"torque after {done} steps")
        tg, cc = ctx.time()
        assert tg == r["sys"].time
    ctx.close()
    return exact


def test_trajectory_fill_100_steps(gpu):
    exact = run_pair(scenarios.hopper_state(500, 21), 100, 25)
    print("bit-exact 100-step fill trajectory:", exact)


@pytest.mark.parametrize("kind", ["bidisperse", "poly"])
def test_trajectory_unequal_radii_100_steps(gpu, kind):
    """Unequal radii (sep = r_l + r_m, m_eff of two sphere masses, the contact point on
    l's sphere) over 100 steps: state within 1e-9, forces and torques within 1e-12 at
    every compared step."""
    exact = run_pair(state_of(kind, 23), 100, 20)
    print(f"bit-exact 100-step trajectory ({kind}):", exact)


def test_trajectory_avalanche_100_steps(gpu):
    exact = run_pair(scenarios.avalanche_state(16, 22), 100, 50)
    print("bit-exact 100-step avalanche trajectory:", exact)


@pytest.mark.parametrize("lean", [-1, 1])
def test_trajectory_candidate_lists_100_steps(gpu, lean):
    """The candidate-list path (verlet_skin > 0) against the oracle, with the clump state
    carried in the state arrays every step (-1) or only in the stage records (1)."""
    exact = run_pair(scenarios.avalanche_state(16, 23), 100, 50, skin=0.02, lean=lean)
    print(f"bit-exact 100-step avalanche trajectory on candidate lists (lean {lean}):", exact)


@pytest.mark.parametrize("skin,lean", [(0.03, -1), (0.002, -1), (0.03, 1), (0.002, 1)])
def test_candidate_lists_equal_per_step_binning(gpu, skin, lean):
    """verlet_skin > 0 changes no bit of the state or of the last step's forces, with list
    rebuilds, chunk undo (the small skin forces them) and spatial reorders in between, in
    both state layouts of the list chunks."""
    from avalancheflume3d_amd import workloads
    cases = [scenarios.hopper_state(500, 61), scenarios.avalanche_state(16, 62)]
    flow = workloads.build(workloads.FlumeSpec(n=300, width=2.0, length=4.0, box=(0.2, 2.0, 1.0),
                                               deposit_length=6.0, spacing=0.37, jitter=0.005,
                                               seed=63, slope_layer=True, layer_depth=1.2))
    cases.append(flow)
    undone = 0
    for inst in cases:
        a = gpu_ctx(inst)
        b = gpu_ctx(inst, skin=skin, reorder=40, lean=lean)
        a.initial_forces()
        b.initial_forces()
        for chunk in (1, 37, 120, 300):
            a.step(chunk)
            b.step(chunk)
            pa, aa, ga = a.download()
            pb, ab, gb = b.download()
            for fld in STATE:
                assert np.array_equal(pa[fld].view(np.uint64), pb[fld].view(np.uint64)), fld
            assert np.array_equal(aa.view(np.uint64), ab.view(np.uint64))
            assert np.array_equal(a.forces()[0].view(np.uint64), b.forces()[0].view(np.uint64))
            assert a.time() == b.time()
        da, db = a.diag(), b.diag()
        assert da["neigh_overflow"] == db["neigh_overflow"] == 0
        print("max touching clumps", da["max_neighbours"], db["max_neighbours"],
              "steps undone", db["steps_undone"])
        undone += db["steps_undone"]
        a.close()
        b.close()
    if skin < 0.01:  # the small skin makes chunks overrun their displacement bound
        assert undone > 0


@pytest.mark.parametrize("lean", [-1, 1])
def test_candidate_list_chunks_across_regrids(gpu, lean):
    """Chunks of list steps apply their per-step parameter advances at their end and end on
    the steps whose clump-cell regrid is due; with a regrid every 7 steps the candidate-list
    path still equals per-step binning bit for bit."""
    import copy
    from avalancheflume3d_amd import workloads
    flow = workloads.build(workloads.FlumeSpec(n=300, width=2.0, length=4.0, box=(0.2, 2.0, 1.0),
                                               deposit_length=6.0, spacing=0.37, jitter=0.005,
                                               seed=64, slope_layer=True, layer_depth=1.2))
    for inst in (scenarios.avalanche_state(16, 65), flow):
        inst = copy.copy(inst)
        inst.geo = abi.Geometry.from_buffer_copy(inst.geo)
        inst.geo.regrid_interval = 7
        a = gpu_ctx(inst)
        b = gpu_ctx(inst, skin=0.02, reorder=40, lean=lean)
        a.initial_forces()
        b.initial_forces()
        for chunk in (3, 40, 101):
            a.step(chunk)
            b.step(chunk)
            pa, aa, _ = a.download()
            pb, ab, _ = b.download()
            for fld in STATE:
                assert np.array_equal(pa[fld].view(np.uint64), pb[fld].view(np.uint64)), fld
            assert np.array_equal(aa.view(np.uint64), ab.view(np.uint64))
            assert a.time() == b.time()
        d = b.diag()
        # 144 steps with a reorder every 40: at least 4 list builds, every step counted
        assert d["steps"] == 144 and d["list_builds"] >= 4 and d["steps_undone"] >= 0
        assert a.diag()["list_builds"] == 0  # no candidate lists without a skin
        a.close()
        b.close()


def test_graph_eager_and_reorder_invariance(gpu):
    inst = scenarios.hopper_state(300, 31)
    outs = []
    for graph, reorder, factor in ((1, 50, 1.0), (0, 0, 1.0), (1, 1, 1.0), (1, 7, 2.5)):
        ctx = gpu_ctx(inst, graph=graph, reorder=reorder, factor=factor)
        ctx.initial_forces()
        ctx.step(20)
        pp, aa, anga = ctx.download()
        outs.append((pp, aa, ctx.forces()[0]))
        ctx.close()
    for pp, aa, f in outs[1:]:
        for fld in STATE:
            assert np.array_equal(pp[fld].view(np.uint64), outs[0][0][fld].view(np.uint64)), fld
        assert np.array_equal(aa.view(np.uint64), outs[0][1].view(np.uint64))
        assert np.array_equal(f.view(np.uint64), outs[0][2].view(np.uint64))


@pytest.mark.parametrize("skin", [0.0, 0.02])
def test_closed_form_mass_and_inertia_change_no_bit(gpu, skin):
    """Mass and inertia recomputed from radius and density (closed_forms 1) against loaded
    (-1): reference clumps (clump_mode 0), config clumps (mode 1), polydisperse radii, and an
    instance whose stored masses are not the closed form (checked on upload: loaded)."""
    import copy

    from avalancheflume3d_amd import workloads
    poly = workloads.build(workloads.FlumeSpec(n=200, width=2.0, length=4.0, box=(2.0, 2.0, 3.21),
                                               rmin=0.09, rmax=0.11, seed=5))
    odd = copy.deepcopy(scenarios.hopper_state(200, 72))
    odd.particles["mass"][7] *= 1.0 + 2.0 ** -40
    for inst in (scenarios.hopper_state(300, 71), scenarios.avalanche_state(12, 73),
                 workloads.small(), poly, odd):
        outs = []
        for closed in (1, -1):
            ctx = gpu_ctx(inst, skin=skin, closed=closed)
            ctx.initial_forces()
            ctx.step(30)
            pp, aa, anga = ctx.download()
            outs.append((pp, aa, anga))
            # recomputed only where asked for and the upload check passed
            assert ctx.diag()["closed_instances"] == int(closed == 1 and inst is not odd)
            ctx.close()
        for fld in STATE:
            assert np.array_equal(outs[0][0][fld].view(np.uint64), outs[1][0][fld].view(np.uint64)), fld
        assert np.array_equal(outs[0][1].view(np.uint64), outs[1][1].view(np.uint64))
        assert np.array_equal(outs[0][2].view(np.uint64), outs[1][2].view(np.uint64))
    # one batch: instances that pass the upload check next to one that does not
    batch = [scenarios.hopper_state(150, 74), odd, poly]
    outs = []
    for closed in (1, -1):
        ctx = gpu_ctx(batch, skin=skin, closed=closed)
        ctx.initial_forces()
        ctx.step(30)
        outs.append([ctx.download(k) for k in range(len(batch))])
        assert ctx.diag()["closed_instances"] == (2 if closed == 1 else 0)
        ctx.close()
    for a, b in zip(*outs):
        for fld in STATE:
            assert np.array_equal(a[0][fld].view(np.uint64), b[0][fld].view(np.uint64)), fld
        assert np.array_equal(a[1].view(np.uint64), b[1].view(np.uint64))


def test_large_configuration_default_path(gpu):
    """2,000,000 polydisperse clumps (lean chunks, and by default the mass and inertia
    recomputed from radius and density) against the same run with them loaded: 20
    candidate-list steps, bit for bit, with no clump outside the clump-cell grid and no
    neighbour-list overflow."""
    from avalancheflume3d_amd import workloads
    inst = workloads.large()
    outs = []
    for closed in (0, -1):
        ctx = gpu_ctx(inst, skin=0.02, reorder=400, closed=closed)
        ctx.initial_forces()
        ctx.step(20)
        outs.append(ctx.download())
        d = ctx.diag()
        assert d["cell_below"] == 0 and d["cell_beyond"] == 0 and d["neigh_overflow"] == 0
        assert d["closed_instances"] == int(closed == 0)
        ctx.close()
    for fld in STATE:
        assert np.array_equal(outs[0][0][fld].view(np.uint64), outs[1][0][fld].view(np.uint64)), fld
    assert np.array_equal(outs[0][1].view(np.uint64), outs[1][1].view(np.uint64))
    assert np.array_equal(outs[0][2].view(np.uint64), outs[1][2].view(np.uint64))


def test_batch_equals_separate_runs(gpu):
    insts = [scenarios.hopper_state(120, 41), scenarios.avalanche_state(10, 42),
             scenarios.hopper_state(333, 43)]
    insts[2].sys.mu_pp = 0.3
    insts[2].sys.fr_fact_pp = 0.3 / 1.0e12
    ctx = gpu_ctx(insts)
    ctx.initial_forces()
    ctx.step(30)
    for k, inst in enumerate(insts):
        one = gpu_ctx(inst)
        one.initial_forces()
        one.step(30)
        a, b = ctx.download(k), one.download(0)
        for fld in STATE:
            assert np.array_equal(a[0][fld].view(np.uint64), b[0][fld].view(np.uint64)), (k, fld)
        one.close()
    ctx.close()


def test_headline_size_lists_equal_per_step_binning(gpu):
    """At the headline size (250,000 clumps, 1M spheres): 150 steps with per-step binning by
    sphere, and on candidate lists in both state layouts with a list rebuild forced by a
    reorder every 64 steps, equal per-step binning by clump bit for bit (state,
    accelerations, last forces, time and counter)."""
    from avalancheflume3d_amd import workloads
    inst = workloads.headline()
    outs = []
    for kw in ({}, {"binning": 1}, {"skin": 0.01, "reorder": 64, "lean": -1},
               {"skin": 0.01, "reorder": 64, "lean": 1}):
        ctx = gpu_ctx(inst, **kw)
        ctx.initial_forces()
        ctx.step(150)
        pp, aa, ga = ctx.download()
        outs.append((pp, aa, ga, ctx.forces()[0], ctx.time()))
        assert ctx.diag()["neigh_overflow"] == 0
        ctx.close()
    ref = outs[0]
    for got in outs[1:]:
        for fld in STATE:
            assert np.array_equal(ref[0][fld].view(np.uint64), got[0][fld].view(np.uint64)), fld
        for k in (1, 2, 3):
            assert np.array_equal(ref[k].view(np.uint64), got[k].view(np.uint64))
        assert ref[4] == got[4]


def test_tiny_and_ragged_batch_on_candidate_lists(gpu):
    """Instances of 1, 2 and 7 clumps batched with a 300-clump flume, stepped on candidate
    lists, equal each instance stepped alone with per-step binning bit for bit."""
    from avalancheflume3d_amd import workloads
    insts = [workloads.build(workloads.FlumeSpec(n=k, width=2.0, length=4.0,
                                                 box=(2.0, 2.0, 0.4 * max(1, k // 25 + 1) + 0.01),
                                                 deposit_length=4.0, seed=70 + k))
             for k in (1, 2, 7, 300)]
    batch = gpu_ctx(insts, skin=0.01, reorder=30)
    batch.initial_forces()
    batch.step(250)
    for k, inst in enumerate(insts):
        one = gpu_ctx(inst)
        one.initial_forces()
        one.step(250)
        pa = one.download()[0]
        pb = batch.download(k)[0]
        for fld in STATE:
            assert np.array_equal(pa[fld].view(np.uint64), pb[fld].view(np.uint64)), (k, fld)
        assert one.time() == batch.time(k)
        one.close()
    batch.close()


@pytest.mark.parametrize("skin", [0.0, 0.01])
def test_small_configuration_100_steps_vs_oracle(gpu, skin):
    """The S configuration (2,000 clumps in a 4x4x8 cm box above a 35 deg flume) against the
    oracle over 100 steps: state within 1e-9, last forces within 1e-12 (per-step binning and
    candidate lists)."""
    from avalancheflume3d_amd import workloads
    exact = run_pair(workloads.small(), 100, 50, skin=skin)
    print(f"bit-exact S trajectory (skin {skin}):", exact)


def test_headline_configuration_vs_oracle(gpu):
    """The headline configuration (250,000 clumps): initial forces within 1e-12 and three
    steps on candidate lists within 1e-9 of the oracle."""
    from avalancheflume3d_amd import workloads
    inst = workloads.headline()
    ctx = gpu_ctx(inst, skin=0.01, reorder=200)
    ctx.initial_forces()
    o = oracle_sim(inst)
    o.initial()
    f, t = ctx.forces()
    r = o.get()
    exact = [vec_close(f, r["ff"], F_RTOL, "force"), vec_close(t, r["tq"], F_RTOL, "torque")]
    ctx.step(3)
    o.step(3)
    pp = ctx.download()[0]
    r = o.get()
    for fld in STATE:
        exact.append(vec_close(pp[fld], r["pp"][fld], T_RTOL, f"{fld} after 3 steps"))
    f, t = ctx.forces()
    exact.append(vec_close(f, r["ff"], F_RTOL, "force after 3 steps"))
    exact.append(vec_close(t, r["tq"], F_RTOL, "torque after 3 steps"))
    print("bit-exact headline initial forces and 3 steps:", all(exact))
    ctx.close()


def test_large_configuration_vs_oracle(gpu):
    """The large configuration at full size (2,000,000 clumps, radii U[0.09, 0.11], mu 0.6,
    eps 0.4), packed 0.36 cm apart so that neighbours of unequal radii overlap from the
    start: initial forces and torques within 1e-12 of the oracle, then two steps with the
    sphere cells rebuilt every step: state within 1e-9, forces and torques within 1e-12."""
    from avalancheflume3d_amd import workloads
    inst = workloads.large(spacing=0.36)
    ctx = gpu_ctx(inst)
    ctx.initial_forces()
    o = oracle_sim(inst)
    o.initial()
    f, t = ctx.forces()
    r = o.get()
    assert np.abs(r["tq"]).max() > 0
    exact = [vec_close(f, r["ff"], F_RTOL, "force"), vec_close(t, r["tq"], F_RTOL, "torque")]
    ctx.step(2)
    o.step(2)
    pp = ctx.download()[0]
    f, t = ctx.forces()
    r = o.get()
    for fld in STATE:
        exact.append(vec_close(pp[fld], r["pp"][fld], T_RTOL, f"{fld} after 2 steps"))
    exact.append(vec_close(f, r["ff"], F_RTOL, "force after 2 steps"))
    exact.append(vec_close(t, r["tq"], F_RTOL, "torque after 2 steps"))
    print("bit-exact large configuration, initial forces and 2 steps:", all(exact))
    ctx.close()
