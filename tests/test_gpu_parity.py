"""GPU parity of the WCSPH step against the CPU oracle (through the C ABI).

Bit-exact: Morton sort order, sorted positions, boundary order, fluid and
boundary neighbour lists (reference enumeration order).
Tolerance (stated per test): density <= 1e-5 relative to the FP64 oracle
for one step from an identical state; pressure error / k <= 1e-4;
N-step trajectories within 10x the CPU FP32-vs-FP64 divergence (+1e-6 m).
"""
import ctypes

import numpy as np
import pytest

import wcsph
from oracle import oracle as O
from tests.helpers import (SMALL, SMALL_REF, baseline_params, by_id, cell_size, reference_params, rel_err,
                           small_scene)

pytestmark = pytest.mark.gpu


def stirred(fl, seed=3, squeeze=0.88, sigma=0.3):
    """A compressed, moving state so pressure and viscosity are active."""
    rng = np.random.default_rng(seed)
    lo = fl[:, :3].min(axis=0)
    pos = fl.copy()
    pos[:, :3] = (lo + (fl[:, :3] - lo) * np.float32(squeeze)).astype(np.float32)
    vel = np.zeros_like(fl)
    vel[:, :3] = rng.normal(0.0, sigma, size=(fl.shape[0], 3)).astype(np.float32)
    return pos, vel


def set_mode(sim, mode, **kw):
    v = sim.get_variant()
    v.neighbor_mode = mode
    for k, val in kw.items():
        setattr(v, k, val)
    sim.set_variant(v)


MODES = [wcsph.abi.NEIGHBORS_CELL, wcsph.abi.NEIGHBORS_PARTICLE, wcsph.abi.NEIGHBORS_CELL_LOOP,
         wcsph.abi.NEIGHBORS_BRICK]


@pytest.mark.parametrize("mode", MODES)
def test_sort_order_and_neighbors_bit_exact(mode):
    fl, bd, box = small_scene()
    prm = baseline_params()
    cs = cell_size(prm)
    h = cs
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        set_mode(sim, mode)
        # boundary: Morton-sorted once at creation
        bgpu = sim.boundary()
        bord = O.sort_order(bd, cs)
        np.testing.assert_array_equal(bgpu[:, :3], bd[bord, :3])
        # a state whose order differs from the Morton order
        pos, vel = stirred(fl)
        rng = np.random.default_rng(5)
        shuf = rng.permutation(fl.shape[0])
        pos, vel = pos[shuf], vel[shuf]
        sim.set_state(pos, vel)
        sim.update(1e-4)
        perm = sim.permutation()
        np.testing.assert_array_equal(perm, O.sort_order(pos, cs))
        sp = sim.sorted_positions()
        np.testing.assert_array_equal(sp, pos[perm])
        counts, lst = sim.neighbors(0)
        oc, ol = O.neighbors(sp, sp, h, cs, True)
        np.testing.assert_array_equal(counts, oc)
        np.testing.assert_array_equal(lst, ol)
        bcounts, blst = sim.neighbors(1)
        obc, obl = O.neighbors(sp, bgpu, h, cs, False)
        np.testing.assert_array_equal(bcounts, obc)
        np.testing.assert_array_equal(blst, obl)
        assert bcounts.sum() > 0
        # second step: ties keep the previous order (stable), from the GPU's own order
        st = sim.get_state()
        sim.update(1e-4)
        np.testing.assert_array_equal(sim.permutation(), O.sort_order(st["pos"], cs))


@pytest.mark.parametrize("mode", MODES)
def test_exact_lattice_neighbors_bit_exact(mode):
    """Unjittered lattice: pairs at exactly 2s = h, where the inclusive
    predicate, its FMA rounding and the cell culling all meet."""
    from wcsph import scenes
    desc = scenes.DamBreak(dims=(14, 14, 14), box_max=(0.64, 0.5, 0.5), jitter_frac=0.0)
    fl, bd, box = small_scene(desc)
    prm = baseline_params()
    cs = cell_size(prm)
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        set_mode(sim, mode)
        sim.update(1e-5)
        sp = sim.sorted_positions()
        np.testing.assert_array_equal(sp, fl[O.sort_order(fl, cs)])
        counts, lst = sim.neighbors(0)
        oc, ol = O.neighbors(sp, sp, cs, cs, True)
        np.testing.assert_array_equal(counts, oc)
        np.testing.assert_array_equal(lst, ol)
        bcounts, blst = sim.neighbors(1)
        obc, obl = O.neighbors(sp, sim.boundary(), cs, cs, False)
        np.testing.assert_array_equal(bcounts, obc)
        np.testing.assert_array_equal(blst, obl)


@pytest.mark.parametrize("mode", MODES)
def test_neighbor_lists_with_clusters_bit_exact(mode):
    """Dense clusters next to sparse particles: lanes of one wavefront differ
    by hundreds of neighbours, which drives the LDS staging window into its
    direct-to-HBM path; lists must still equal the oracle's."""
    fl, bd, box = small_scene()
    prm = baseline_params()
    cs = cell_size(prm)
    rng = np.random.default_rng(11)
    n = fl.shape[0]
    pos = fl.copy()
    lo, hi = fl[:, :3].min(axis=0), fl[:, :3].max(axis=0)
    pos[:, :3] = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    for c in range(3):  # three tight clusters of 300 particles each
        centre = rng.uniform(lo + 0.05, hi - 0.05).astype(np.float32)
        sel = slice(c * 300, (c + 1) * 300)
        pos[sel, :3] = centre + rng.normal(0, 0.004, size=(300, 3)).astype(np.float32)
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        set_mode(sim, mode, fluid_capacity=512, neighbor_groups=37)
        sim.set_state(pos)
        st = sim.update(1e-5)
        assert st.flags == 0 and st.max_fluid_neighbors > 250
        sp = sim.sorted_positions()
        np.testing.assert_array_equal(sp, pos[O.sort_order(pos, cs)])
        counts, lst = sim.neighbors(0)
        oc, ol = O.neighbors(sp, sp, cs, cs, True)
        np.testing.assert_array_equal(counts, oc)
        np.testing.assert_array_equal(lst, ol)
        bcounts, blst = sim.neighbors(1)
        obc, obl = O.neighbors(sp, sim.boundary(), cs, cs, False)
        np.testing.assert_array_equal(bcounts, obc)
        np.testing.assert_array_equal(blst, obl)


@pytest.mark.parametrize("capf,capb", [(64, 32), (56, 24)])
def test_cell_builder_with_capacities_inside_its_stage(capf, capb):
    """Capacities that fit the cell builder's LDS stage select its variant
    that never stores to HBM directly; (56, 24) is the exact fit of this
    state's largest counts (56 fluid, 21 boundary)."""
    fl, bd, box = small_scene()
    prm = baseline_params()
    cs = cell_size(prm)
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        set_mode(sim, wcsph.abi.NEIGHBORS_CELL, fluid_capacity=capf, boundary_capacity=capb, neighbor_groups=61)
        pos, vel = stirred(fl)
        sim.set_state(pos, vel)
        st = sim.update(1e-5)
        assert st.flags == 0 and st.max_fluid_neighbors == 56
        sp = sim.sorted_positions()
        counts, lst = sim.neighbors(0)
        oc, ol = O.neighbors(sp, sp, cs, cs, True)
        np.testing.assert_array_equal(counts, oc)
        np.testing.assert_array_equal(lst, ol)
        bcounts, blst = sim.neighbors(1)
        obc, obl = O.neighbors(sp, sim.boundary(), cs, cs, False)
        np.testing.assert_array_equal(bcounts, obc)
        np.testing.assert_array_equal(blst, obl)


def test_boundary_mass_vs_fp64():
    fl, bd, box = small_scene()
    prm = baseline_params()
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        g = sim.boundary()
    ref = O.Solver(prm, fl, bd, box, precision=64).boundary()
    np.testing.assert_array_equal(g[:, :3], ref[:, :3].astype(np.float32))
    assert rel_err(g[:, 3], ref[:, 3]) <= 1e-5


def _one_step_compare(prm, fl, bd, box, pos, vel, dt):
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        sim.set_state(pos, vel)
        sim.update(dt)
        g = by_id(sim.get_state())
    ref = O.Solver(prm, pos, bd, box, fluid_vel=vel, precision=64)
    ref.step(dt)
    r = by_id(ref.state())
    return g, r


@pytest.mark.parametrize("preset", ["baseline", "reference"])
def test_one_step_vs_fp64(preset):
    if preset == "baseline":
        prm, desc = baseline_params(), SMALL
        dt = 1e-4
    else:
        prm, desc = reference_params(), SMALL_REF
        dt = 1e-4
    fl, bd, box = small_scene(desc)
    pos, vel = stirred(fl)
    g, r = _one_step_compare(prm, fl, bd, box, pos, vel, dt)
    # density: <= 1e-5 relative to FP64 (stated tolerance)
    assert rel_err(g["rho"], r["rho"]) <= 1e-5
    # pressure: error normalised by the stiffness
    assert np.max(np.abs(g["p"] - r["p"])) / prm.stiffness <= 1e-4
    assert np.max(r["p"]) > 0.0  # the state is compressed: pressure is active
    # positions / velocities after one step
    assert np.max(np.abs(g["pos"][:, :3] - r["pos"][:, :3])) <= 1e-6
    vscale = np.max(np.abs(r["vel"][:, :3]))
    assert np.max(np.abs(g["vel"][:, :3] - r["vel"][:, :3])) <= 2e-5 * vscale + 1e-6


def test_n_steps_within_fp32_divergence():
    fl, bd, box = small_scene()
    prm = baseline_params()
    pos, vel = stirred(fl, sigma=0.2)
    steps, dt = 60, 1e-4
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        sim.set_state(pos, vel)
        st = sim.step(dt, steps)
        g = by_id(sim.get_state())
        assert st.steps == steps and abs(st.sim_time - steps * np.float32(dt)) < 1e-9
    r64 = O.Solver(prm, pos, bd, box, fluid_vel=vel, precision=64)
    r64.step(dt, steps)
    r32 = O.Solver(prm, pos, bd, box, fluid_vel=vel, precision=32)
    r32.step(dt, steps)
    a, b = by_id(r64.state()), by_id(r32.state())
    e_cpu = np.max(np.abs(b["pos"][:, :3] - a["pos"][:, :3]))
    e_gpu = np.max(np.abs(g["pos"][:, :3] - a["pos"][:, :3]))
    assert e_gpu <= 10 * e_cpu + 1e-6, (e_gpu, e_cpu)
    ev_cpu = np.max(np.abs(b["vel"][:, :3] - a["vel"][:, :3]))
    ev_gpu = np.max(np.abs(g["vel"][:, :3] - a["vel"][:, :3]))
    assert ev_gpu <= 10 * ev_cpu + 1e-5, (ev_gpu, ev_cpu)
    assert rel_err(g["rho"], a["rho"]) <= 1e-4


def test_cfl_dt_matches_oracle():
    fl, bd, box = small_scene()
    prm = baseline_params(adaptive_dt=1)
    pos, vel = stirred(fl, sigma=8.0)  # vmax ~ 35 m/s: the CFL branch decides dt
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        sim.set_state(pos, vel)
        st = sim.step(1e-3, 3)
    ref = O.Solver(prm, pos, bd, box, fluid_vel=vel, precision=64)
    ref.step(1e-3, 3)
    t, last = ref.time()
    assert last < 5e-4  # CFL-limited, below dt_max
    assert abs(st.last_dt - last) <= 1e-6 * last
    assert abs(st.sim_time - t) <= 1e-6 * t


def test_cfl_at_rest_uses_dt_max():
    fl, bd, box = small_scene()
    prm = baseline_params(adaptive_dt=1)
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        st = sim.update(1e-2)  # zero velocities: dt = dt_max
        assert st.last_dt == pytest.approx(np.float32(5e-4), rel=1e-7)


@pytest.mark.parametrize("preset", ["baseline", "reference"])
def test_neighbor_modes_give_identical_states(preset):
    """The list builders and the cell-loop form (no lists) sum the same
    pair terms in the same order: states are bit-identical."""
    if preset == "baseline":
        prm, desc, dt = baseline_params(), SMALL, 1e-4
    else:
        prm, desc, dt = reference_params(), SMALL_REF, 1e-3
    fl, bd, box = small_scene(desc)
    pos, vel = stirred(fl)
    out = []
    for mode in MODES:
        with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
            set_mode(sim, mode)
            sim.set_state(pos, vel)
            st = sim.step(dt, 8)
            out.append((sim.get_state(), st))
    for m, (state, st) in enumerate(out[1:], 1):
        for k in ("rho", "p", "pos", "vel", "ids"):
            np.testing.assert_array_equal(out[0][0][k], state[k], err_msg=f"mode {MODES[m]}: {k}")
        assert st.max_fluid_neighbors == out[0][1].max_fluid_neighbors
        assert st.max_boundary_neighbors == out[0][1].max_boundary_neighbors


def test_cell_loop_beyond_lds_capacity():
    """A cluster of 2500 particles inside half a cell: its bricks' halos hold
    more candidates than LDS stages, so cells are read from HBM.  States
    still equal the list form's."""
    fl, bd, box = small_scene()
    prm = baseline_params()
    rng = np.random.default_rng(4)
    pos = fl.copy()
    centre = fl[:, :3].mean(axis=0)
    d = rng.normal(size=(2500, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos[:2500, :3] = (centre + 0.02 * d * rng.uniform(0, 1, (2500, 1)) ** (1 / 3)).astype(np.float32)
    out = []
    for mode in (wcsph.abi.NEIGHBORS_PARTICLE, wcsph.abi.NEIGHBORS_CELL_LOOP, wcsph.abi.NEIGHBORS_BRICK):
        with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
            set_mode(sim, mode, fluid_capacity=4096)
            sim.set_state(pos)
            st = sim.update(1e-7)
            assert st.flags == 0 and st.max_fluid_neighbors > 2400
            out.append(sim.get_state())
    for m in (1, 2):
        for k in ("rho", "p", "pos", "vel", "ids"):
            np.testing.assert_array_equal(out[0][k], out[m][k], err_msg=f"{m}: {k}")


def test_graph_and_eager_bit_identical():
    fl, bd, box = small_scene()
    prm = baseline_params()
    pos, vel = stirred(fl)
    out = []
    for use_graph in (0, 1):
        with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
            v = sim.get_variant()
            v.use_graph = use_graph
            sim.set_variant(v)
            sim.set_state(pos, vel)
            sim.step(1e-4, 10)
            sim.step(1e-4, 5)
            out.append(sim.get_state())
    for k in ("rho", "p", "pos", "vel", "ids"):
        np.testing.assert_array_equal(out[0][k], out[1][k], err_msg=k)


@pytest.mark.parametrize("mode", [wcsph.abi.NEIGHBORS_CELL, wcsph.abi.NEIGHBORS_PARTICLE,
                                  wcsph.abi.NEIGHBORS_CELL_LOOP])
def test_every_graph_replay_matches_eager(mode):
    """One graph, replayed step by step with the library's own eager work
    between the replays (state export, a neighbour-list rebuild for export,
    a metrics reduction): the state after every replay equals the eager
    run's state after the same step."""
    fl, bd, box = small_scene()
    prm = baseline_params()
    pos, vel = stirred(fl)
    runs = []
    for use_graph in (0, 1):
        states = []
        with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
            v = sim.get_variant()
            v.use_graph, v.neighbor_mode = use_graph, mode
            sim.set_variant(v)
            sim.set_state(pos, vel)
            for _ in range(6):
                st = sim.step(1e-4, 1)
                assert st.flags == 0
                states.append(sim.get_state())
                sim.neighbors(0)
                sim.compute_state_metrics()
        runs.append(states)
    for step, (a, b) in enumerate(zip(*runs)):
        for k in ("rho", "p", "pos", "vel", "ids"):
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"step {step + 1}: {k}")


def test_block_sizes_agree():
    fl, bd, box = small_scene()
    prm = baseline_params()
    pos, vel = stirred(fl)
    ref = None
    for bs in (64, 256, 1024):
        with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
            v = sim.get_variant()
            v.block_size = bs
            sim.set_variant(v)
            sim.set_state(pos, vel)
            sim.step(1e-4, 5)
            st = sim.get_state()
        if ref is None:
            ref = st
        else:
            np.testing.assert_array_equal(st["pos"], ref["pos"])


def converging(fl, rate=150.0):
    """The stirred state plus a flow towards the block's centre, so that the
    neighbour counts grow from step to step."""
    pos, vel = stirred(fl)
    c = pos[:, :3].mean(axis=0)
    vel[:, :3] += (-rate * (pos[:, :3] - c)).astype(np.float32)
    return pos, vel


def _run_lists(mode, capf, capb, pos, vel, fl, bd, box, per_step=False):
    """Two async calls with different dt hints (2 + 3 steps) and one sync, or
    (per_step) one sync per step; returns the stats, state, lists, variant
    and sorted positions."""
    prm = baseline_params()
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        set_mode(sim, mode, fluid_capacity=capf, boundary_capacity=capb)
        sim.set_state(pos, vel)
        if per_step:
            maxima = []
            for hint in (1e-4, 1e-4, 2e-4, 2e-4, 2e-4):
                maxima.append(sim.step(hint, 1).max_fluid_neighbors)
            return maxima
        sim.step_async(1e-4, 2)
        sim.step_async(2e-4, 3)
        st = sim.sync()
        return st, sim.get_state(), (sim.neighbors(0), sim.neighbors(1)), sim.get_variant(), sim.sorted_positions()


@pytest.mark.parametrize("mode,capb", [(wcsph.abi.NEIGHBORS_PARTICLE, 64), (wcsph.abi.NEIGHBORS_CELL, 64),
                                        (wcsph.abi.NEIGHBORS_CELL, 8), (wcsph.abi.NEIGHBORS_BRICK, 64)])
@pytest.mark.parametrize("first", ["first_step", "later_step"])
def test_neighbor_overflow_recovers(mode, capb, first):
    """Lists too small for the state, for every form that builds lists.  The
    step that overflows stores nothing, the later ones are skipped, and
    wcsph_sync grows the lists and runs the lost steps again with the dt
    hints they were enqueued with (two async calls, hints 1e-4 and 2e-4).
    The final state, time, step count, maxima and lists equal a run started
    with capacity 128, bit for bit, and the lists equal the oracle's.
    first_step: fluid capacity 8 (the first step overflows); later_step: the
    capacity just above the first step's largest count, on a converging flow
    whose counts grow, so a later step overflows after earlier ones
    completed.  The cell builder with capb 8 runs its all-LDS variant."""
    fl, bd, box = small_scene()
    pos, vel = converging(fl)
    ref_st, ref, ref_lists, _, _ = _run_lists(mode, 128, 64, pos, vel, fl, bd, box)
    assert ref_st.flags == 0
    if first == "first_step":
        capf = 8
    else:
        maxima = _run_lists(mode, 128, 64, pos, vel, fl, bd, box, per_step=True)
        capf = (maxima[0] + 3) // 4 * 4
        assert max(maxima[1:]) > capf, maxima  # a later step needs more than the first
        capb = 64
    st, got, lists, var, spos = _run_lists(mode, capf, capb, pos, vel, fl, bd, box)
    assert st.flags == wcsph.abi.FLAG_LISTS_GROWN
    assert var.fluid_capacity >= ref_st.max_fluid_neighbors and var.fluid_capacity % 4 == 0
    assert (st.steps, st.sim_time, st.last_dt) == (ref_st.steps, ref_st.sim_time, ref_st.last_dt)
    assert st.max_fluid_neighbors == ref_st.max_fluid_neighbors
    assert st.max_boundary_neighbors == ref_st.max_boundary_neighbors
    for k in ("pos", "vel", "rho", "p", "ids"):
        np.testing.assert_array_equal(got[k], ref[k])
    for w in (0, 1):
        np.testing.assert_array_equal(lists[w][0], ref_lists[w][0])
        np.testing.assert_array_equal(lists[w][1], ref_lists[w][1])
    h = cell_size(baseline_params())
    oc, ol = O.neighbors(spos, spos, h, h, True)
    np.testing.assert_array_equal(lists[0][0], oc)
    np.testing.assert_array_equal(lists[0][1], ol)


def test_neighbor_overflow_beyond_the_largest_capacity():
    """5,000 particles packed inside a ball of radius 0.3 h: every particle
    has 4,999 neighbours, past the largest list capacity (4,096).  The step
    fails with WCSPH_ERR_NEIGHBOR_OVERFLOW and the state, time and step count
    stay those the step started from."""
    prm = baseline_params()
    h = float(cell_size(prm))
    rng = np.random.default_rng(8)
    d = rng.normal(size=(5000, 3))
    d *= (rng.uniform(0, 1, 5000) ** (1 / 3) * 0.3 * h / np.linalg.norm(d, axis=1))[:, None]
    pos = np.ones((5000, 4), np.float32)
    pos[:, :3] = (d + 1.0).astype(np.float32)
    box = wcsph.make_box((0.0, 0.0, 0.0), (2.0, 2.0, 2.0))
    with wcsph.FluidSimulator(prm, pos, None, box) as sim:
        st = wcsph.abi.Stats()
        rc = sim.lib.wcsph_step(sim.handle, 1e-4, 2, ctypes.byref(st))
        assert rc == wcsph.abi.WCSPH_ERR_NEIGHBOR_OVERFLOW
        assert b"4096" in sim.lib.wcsph_last_error(sim.handle)
        assert st.flags & wcsph.abi.FLAG_FLUID_OVERFLOW and st.max_fluid_neighbors == 4999
        assert st.steps == 0 and st.sim_time == 0.0
        g = sim.get_state()
        np.testing.assert_array_equal(g["pos"], pos)
        np.testing.assert_array_equal(g["ids"], np.arange(5000, dtype=np.uint32))


def test_nonfinite_state_is_reported():
    """An external acceleration of 1e38 m/s^2 over dt = 10 s drives the
    velocities to infinity in the first step: the step reports
    WCSPH_FLAG_NONFINITE and returns WCSPH_ERR_NONFINITE."""
    fl, bd, box = small_scene()
    prm = baseline_params()
    wcsph.abi.set_force_rules(prm, [wcsph.abi.force_rule(value=(0.0, 1e38, 0.0))])
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        st = wcsph.abi.Stats()
        rc = sim.lib.wcsph_step(sim.handle, 10.0, 1, ctypes.byref(st))
        assert rc == wcsph.abi.WCSPH_ERR_NONFINITE
        assert st.flags & wcsph.abi.FLAG_NONFINITE
        assert b"non-finite" in sim.lib.wcsph_last_error(sim.handle)
        assert not np.all(np.isfinite(sim.get_state()["vel"]))


def test_domain_error_on_set_state():
    fl, bd, box = small_scene()
    prm = baseline_params()
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        bad = fl.copy()
        bad[0, 0] = 50.0
        with pytest.raises(wcsph.WcsphError) as e:
            sim.set_state(bad)
        assert e.value.code == wcsph.abi.WCSPH_ERR_DOMAIN


def test_state_metrics_fp64():
    fl, bd, box = small_scene()
    prm = baseline_params()
    pos, vel = stirred(fl)
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        sim.set_state(pos, vel)
        ms, ke = sim.compute_state_metrics()
    v = vel[:, :3].astype(np.float64)
    sp2 = np.sum(v * v, axis=1)
    mass = np.float32(1000.0) * (np.float32(0.02) ** 3 * np.float32(0.8))
    assert ms == pytest.approx(np.mean(np.sqrt(sp2)), rel=1e-12)
    assert ke == pytest.approx(0.5 * float(mass) * np.sum(sp2), rel=1e-6)


def test_reset_restores_lattice():
    fl, bd, box = small_scene()
    prm = baseline_params()
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        sim.step(1e-4, 3)
        sim.reset()
        st = sim.get_state()
        np.testing.assert_array_equal(st["pos"], fl)
        assert np.all(st["vel"] == 0)


def test_scene_without_fluid_steps_its_clock():
    """No fluid particles: a step only advances the scene's time (dt_max in
    adaptive mode, the vmax < 1e-9 branch of sph.cu:115-130); every export is
    empty and differentiable runs are refused."""
    _, bd, box = small_scene()
    prm = baseline_params(adaptive_dt=1)
    empty = np.zeros((0, 4), np.float32)
    with wcsph.FluidSimulator(prm, empty, bd, box) as sim:
        st = sim.step(1e-4, 3)
        assert st.steps == 3 and st.flags == 0
        assert st.sim_time == pytest.approx(3 * float(np.float32(prm.dt_max)), rel=1e-6)
        assert sim.get_state()["pos"].shape == (0, 4)
        counts, lst = sim.neighbors(0)
        assert counts.shape == (0,) and lst.shape == (0,)
        assert sim.compute_state_metrics() == (0.0, 0.0)
        assert sim.boundary().shape == bd.shape
    with wcsph.FluidSimulator(baseline_params(), empty, bd, box) as sim:
        out = np.zeros(4, np.float32)
        rc = sim.lib.wcsph_grad_forward(sim.handle, None, 1e-4, 2, 1, out.ctypes.data)
        assert rc == wcsph.abi.WCSPH_ERR_INVALID_ARGUMENT
        assert b"need fluid" in sim.lib.wcsph_last_error(sim.handle)
