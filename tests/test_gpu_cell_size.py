"""The fluid grid at the reference's other cell sizes (CELL_SIZE_MULT of its
KTT space, step_tuner.cuh:88: cell size h * mult, walk range
RANGE = (int)(h / cs + 0.5), common.cuh:126-151), with the per-particle list
builder, the one form that takes a multiplier other than 1.

Bit-exact against the oracle at each multiplier: the Morton sort order on
the fp32(h) * fp32(mult) grid, the fluid lists in (2R+1)^3 walk order, the
boundary order and lists on the boundary grid (cell size h whatever the
multiplier).  Tolerance: density <= 1e-5 relative to the FP64 oracle on the
same grid after one step.  Form and range checks of wcsph_set_variant.
"""
import numpy as np
import pytest

import wcsph
from oracle import oracle as O
from tests.helpers import baseline_params, by_id, cell_size, rel_err, small_scene
from tests.test_gpu_parity import converging, stirred

pytestmark = pytest.mark.gpu

F32 = np.float32
MULTS = [F32(2.0) / F32(3.0), F32(0.5), F32(1.5), F32(2.0)]
IDS = ["2/3", "1/2", "3/2", "2"]


def with_mult(sim, mult, mode=wcsph.abi.NEIGHBORS_PARTICLE, **kw):
    v = sim.get_variant()
    v.neighbor_mode = mode
    v.cell_size_mult = float(mult)
    for k, val in kw.items():
        setattr(v, k, val)
    sim.set_variant(v)
    got = sim.get_variant()
    assert F32(got.cell_size_mult) == F32(mult)


def check_lists(sim, h, cs):
    sp = sim.sorted_positions()
    counts, lst = sim.neighbors(0)
    oc, ol = O.neighbors(sp, sp, h, cs, True)
    np.testing.assert_array_equal(counts, oc)
    np.testing.assert_array_equal(lst, ol)
    bcounts, blst = sim.neighbors(1)
    obc, obl = O.neighbors(sp, sim.boundary(), h, h, False)
    np.testing.assert_array_equal(bcounts, obc)
    np.testing.assert_array_equal(blst, obl)
    return counts, bcounts


@pytest.mark.parametrize("mult", MULTS, ids=IDS)
def test_sort_order_and_lists_bit_exact(mult):
    fl, bd, box = small_scene()
    prm = baseline_params()
    h = cell_size(prm)
    cs = F32(h * mult)
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        with_mult(sim, mult)
        # the boundary keeps its order on the grid of cell size h
        np.testing.assert_array_equal(sim.boundary()[:, :3], bd[O.sort_order(bd, h), :3])
        pos, vel = stirred(fl)
        shuf = np.random.default_rng(5).permutation(fl.shape[0])
        pos, vel = pos[shuf], vel[shuf]
        sim.set_state(pos, vel)
        st = sim.update(1e-4)
        assert st.flags == 0
        perm = sim.permutation()
        np.testing.assert_array_equal(perm, O.sort_order(pos, cs))
        np.testing.assert_array_equal(sim.sorted_positions(), pos[perm])
        counts, bcounts = check_lists(sim, h, cs)
        assert bcounts.sum() > 0 and counts.max() > 30
        # second step from the GPU's own order (stable ties)
        g = sim.get_state()
        sim.update(1e-4)
        np.testing.assert_array_equal(sim.permutation(), O.sort_order(g["pos"], cs))
        check_lists(sim, h, cs)


@pytest.mark.parametrize("mult", MULTS[:2], ids=IDS[:2])
def test_exact_lattice_and_clusters_bit_exact(mult):
    """Pairs at exactly h on an unjittered lattice (the inclusive predicate
    and the cell culling of the range-2 walk meet), then dense clusters
    whose hundreds of neighbours drive the LDS window into its direct-to-HBM
    path."""
    from wcsph import scenes
    prm = baseline_params()
    h = cell_size(prm)
    cs = F32(h * mult)
    desc = scenes.DamBreak(dims=(14, 14, 14), box_max=(0.64, 0.5, 0.5), jitter_frac=0.0)
    fl, bd, box = small_scene(desc)
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        with_mult(sim, mult)
        sim.update(1e-5)
        np.testing.assert_array_equal(sim.sorted_positions(), fl[O.sort_order(fl, cs)])
        check_lists(sim, h, cs)
    fl, bd, box = small_scene()
    rng = np.random.default_rng(11)
    pos = fl.copy()
    lo, hi = fl[:, :3].min(axis=0), fl[:, :3].max(axis=0)
    pos[:, :3] = rng.uniform(lo, hi, size=pos[:, :3].shape).astype(F32)
    for c in range(3):
        centre = rng.uniform(lo + 0.05, hi - 0.05).astype(F32)
        pos[c * 300:(c + 1) * 300, :3] = centre + rng.normal(0, 0.004, size=(300, 3)).astype(F32)
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        with_mult(sim, mult, fluid_capacity=512)
        sim.set_state(pos)
        st = sim.update(1e-5)
        assert st.flags == 0 and st.max_fluid_neighbors > 250
        np.testing.assert_array_equal(sim.sorted_positions(), pos[O.sort_order(pos, cs)])
        check_lists(sim, h, cs)


@pytest.mark.parametrize("mult", MULTS[:2], ids=IDS[:2])
def test_one_step_vs_fp64_on_the_same_grid(mult):
    fl, bd, box = small_scene()
    prm = baseline_params()
    pos, vel = stirred(fl)
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        with_mult(sim, mult)
        sim.set_state(pos, vel)
        sim.update(1e-4)
        g = by_id(sim.get_state())
    ref = O.Solver(prm, pos, bd, box, fluid_vel=vel, precision=64)
    ref.set_cell_size_mult(float(mult))
    ref.step(1e-4)
    r = by_id(ref.state())
    assert rel_err(g["rho"], r["rho"]) <= 1e-5
    assert np.max(np.abs(g["p"] - r["p"])) / prm.stiffness <= 1e-4
    assert np.max(np.abs(g["pos"][:, :3] - r["pos"][:, :3])) <= 1e-6


def test_cell_size_changes_only_the_summation_order():
    """Ten CFL steps at cell sizes h, 2h/3 and h/2 from one state: the same
    neighbour counts per particle at every step (the sets do not depend on
    the grid), densities within FP32 summation-order noise."""
    fl, bd, box = small_scene()
    prm = baseline_params(adaptive_dt=1)
    pos, vel = stirred(fl)
    out = []
    for mult in (F32(1.0),) + tuple(MULTS[:2]):
        with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
            with_mult(sim, mult)
            sim.set_state(pos, vel)
            per = []
            for _ in range(10):
                st = sim.step(2e-4, 1)
                assert st.flags == 0
                s = by_id(sim.get_state())
                c = np.empty(fl.shape[0], np.uint32)
                c[sim.get_state()["ids"]] = sim.neighbors(0)[0]
                per.append((s, c, st.last_dt))
            out.append(per)
    for per in out[1:]:
        for k, ((s, c, dt), (s0, c0, dt0)) in enumerate(zip(per, out[0])):
            if k == 0:
                np.testing.assert_array_equal(c, c0)
                assert rel_err(s["rho"], s0["rho"]) <= 5e-6
            assert rel_err(dt, dt0) <= 1e-3
        assert np.max(np.abs(per[-1][0]["pos"][:, :3] - out[0][-1][0]["pos"][:, :3])) <= 1e-4


def test_switching_cell_size_mid_run_and_overflow_recovery():
    """The grid is rebuilt by wcsph_set_variant between steps (h -> h/2 ->
    2h/3 -> h), lists stay bit-exact; at cell size 2h/3 a list overflow is
    recovered as at cell size h (same state, time and lists as a run with
    room)."""
    fl, bd, box = small_scene()
    prm = baseline_params()
    h = cell_size(prm)
    pos, vel = stirred(fl)
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        sim.set_state(pos, vel)
        for mult in (F32(1.0), F32(0.5), MULTS[0], F32(1.0)):
            with_mult(sim, mult)
            assert sim.step(1e-4, 2).flags == 0
            check_lists(sim, h, F32(h * mult))
    pos, vel = converging(fl)
    res = []
    for capf in (128, 8):
        with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
            with_mult(sim, MULTS[0], fluid_capacity=capf)
            sim.set_state(pos, vel)
            sim.step_async(1e-4, 2)
            sim.step_async(2e-4, 3)
            st = sim.sync()
            res.append((st, sim.get_state(), sim.neighbors(0), sim.get_variant()))
            check_lists(sim, h, F32(h * MULTS[0]))
    (st0, g0, l0, _), (st1, g1, l1, v1) = res
    assert st0.flags == 0 and st1.flags == wcsph.abi.FLAG_LISTS_GROWN
    assert F32(v1.cell_size_mult) == MULTS[0] and v1.fluid_capacity >= st0.max_fluid_neighbors
    assert (st1.steps, st1.sim_time) == (st0.steps, st0.sim_time)
    for k in ("pos", "vel", "rho", "ids"):
        np.testing.assert_array_equal(g1[k], g0[k])
    np.testing.assert_array_equal(l1[1], l0[1])


def test_batched_scenes_at_cell_size():
    """Several uneven scenes batched at cell size 2h/3 equal each scene
    stepped alone at that cell size."""
    from tests.test_gpu_batch import _scenes
    prm = baseline_params(adaptive_dt=1)
    sc = _scenes(5, 9)
    with wcsph.BatchSimulator(prm, sc) as b:
        with_mult(b, MULTS[0])
        b.step(1e-3, 6)
        st = b.get_state()
    for k, (fl, bd, box, vel) in enumerate(sc):
        with wcsph.FluidSimulator(prm, fl, bd, box, fluid_vel=vel) as one:
            with_mult(one, MULTS[0])
            one.step(1e-3, 6)
            g = one.get_state()
        sl = b.ranges[k][0], b.ranges[k][0] + b.ranges[k][1]
        np.testing.assert_array_equal(st["pos"][sl[0]:sl[1]], g["pos"])
        np.testing.assert_array_equal(st["rho"][sl[0]:sl[1]], g["rho"])


def test_set_variant_checks_the_multiplier():
    fl, bd, box = small_scene()
    prm = baseline_params()
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        before = sim.get_variant()
        for mode in (wcsph.abi.NEIGHBORS_CELL, wcsph.abi.NEIGHBORS_CELL_LOOP, wcsph.abi.NEIGHBORS_BRICK):
            with pytest.raises(wcsph.WcsphError, match="per-particle builder"):
                with_mult(sim, MULTS[0], mode=mode)
        for bad in (0.25, 0.49, 2.5, float("nan")):
            with pytest.raises(wcsph.WcsphError, match=r"cell_size_mult must be in \[0.5, 2\]"):
                with_mult(sim, bad)
        after = sim.get_variant()
        assert [getattr(after, f) for f, _ in after._fields_] == [getattr(before, f) for f, _ in before._fields_]
        # 0 means 1; the grid then serves every form again
        with_mult(sim, MULTS[1])
        v = sim.get_variant()
        v.cell_size_mult = 0.0
        v.neighbor_mode = wcsph.abi.NEIGHBORS_CELL
        sim.set_variant(v)
        assert sim.get_variant().cell_size_mult == 1.0
        assert sim.step(1e-4, 1).flags == 0
