"""GPU parity on the reference's own saved states (the .sphs fixtures under
tests/golden/), snapshot round trips through the GPU state, and the
autotuner (state unchanged, winner applied, results unchanged)."""
import os

import numpy as np
import pytest

import wcsph
from oracle import oracle as O
from tests.helpers import baseline_params, by_id, cell_size, reference_params, rel_err, small_scene
from wcsph import autotune, scenes, snapshot

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STATES = ["dragon_falling", "dragon_settled", "surface_settled", "tilting_chaotic", "fountain_active", "vortex_active"]
MODES = [wcsph.abi.NEIGHBORS_CELL, wcsph.abi.NEIGHBORS_PARTICLE]


def reference_state_sim(state, mode):
    _, pos, vel = snapshot.load(os.path.join(GOLD, f"snap_{state}_2048.sphs"))
    lo = pos[:, :3].min(axis=0) - 0.5
    hi = pos[:, :3].max(axis=0) + 0.5
    prm = reference_params()
    sim = wcsph.FluidSimulator(prm, pos, np.zeros((0, 4), np.float32), scenes.make_box(lo, hi))
    v = sim.get_variant()
    v.neighbor_mode = mode
    v.fluid_capacity = 256
    sim.set_variant(v)
    sim.set_state(pos, vel)
    return sim, prm, pos


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("state", STATES)
def test_reference_states_sort_and_neighbors_bit_exact(state, mode):
    sim, prm, pos = reference_state_sim(state, mode)
    with sim:
        cs = cell_size(prm)
        sim.update(1e-6)
        perm = sim.permutation()
        np.testing.assert_array_equal(perm, O.sort_order(pos, cs))
        sp = sim.sorted_positions()
        np.testing.assert_array_equal(sp, pos[perm])
        counts, lst = sim.neighbors(0)
        oc, ol = O.neighbors(sp, sp, cs, cs, True)
        np.testing.assert_array_equal(counts, oc)
        np.testing.assert_array_equal(lst, ol)
        assert counts.max() > 20


@pytest.mark.parametrize("state", STATES)
def test_reference_states_one_step_density(state):
    """One reference-physics step from a saved reference state against the
    FP64 oracle: density <= 1e-5 relative, positions <= 1e-6 m."""
    _, pos, vel = snapshot.load(os.path.join(GOLD, f"snap_{state}_2048.sphs"))
    sim, prm, _ = reference_state_sim(state, wcsph.abi.NEIGHBORS_PARTICLE)
    box = scenes.make_box(pos[:, :3].min(axis=0) - 0.5, pos[:, :3].max(axis=0) + 0.5)
    with sim:
        sim.update(1e-5)
        g = by_id(sim.get_state())
    orc = O.Solver(prm, pos, np.zeros((0, 4), np.float32), box, fluid_vel=vel, precision=64)
    orc.step(1e-5)
    r = O.by_id(orc.state())
    assert rel_err(g["rho"], r["rho"]) <= 1e-5
    assert np.max(np.abs(g["pos"][:, :3] - r["pos"][:, :3])) <= 1e-6


FULL = np.load(os.path.join(GOLD, "ref_states_positions.npz"))


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("state", STATES)
def test_reference_states_full_size_bit_exact(state, mode):
    """Every particle of each saved reference state (27,000 to 125,000):
    sort order, sorted positions and fluid neighbour lists equal the
    oracle's, bit for bit, and one reference-physics step from rest matches
    the FP64 oracle's density (1e-5 relative) and positions (1e-6 m)."""
    xyz = FULL[state]
    pos = np.ones((xyz.shape[0], 4), np.float32)
    pos[:, :3] = xyz
    prm = reference_params()
    box = scenes.make_box(xyz.min(axis=0) - 0.5, xyz.max(axis=0) + 0.5)
    with wcsph.FluidSimulator(prm, pos, np.zeros((0, 4), np.float32), box) as sim:
        v = sim.get_variant()
        v.neighbor_mode = mode
        v.fluid_capacity = 256
        sim.set_variant(v)
        st = sim.update(1e-5)
        assert st.flags == 0
        cs = cell_size(prm)
        perm = sim.permutation()
        np.testing.assert_array_equal(perm, O.sort_order(pos, cs))
        sp = sim.sorted_positions()
        np.testing.assert_array_equal(sp, pos[perm])
        counts, lst = sim.neighbors(0)
        oc, ol = O.neighbors(sp, sp, cs, cs, True)
        np.testing.assert_array_equal(counts, oc)
        np.testing.assert_array_equal(lst, ol)
        g = by_id(sim.get_state())
    if mode == wcsph.abi.NEIGHBORS_PARTICLE:
        orc = O.Solver(prm, pos, np.zeros((0, 4), np.float32), box, precision=64)
        orc.step(1e-5)
        r = O.by_id(orc.state())
        assert rel_err(g["rho"], r["rho"]) <= 1e-5
        assert np.max(np.abs(g["pos"][:, :3] - r["pos"][:, :3])) <= 1e-6


def test_snapshot_roundtrip_through_gpu_state(tmp_path):
    fl, bd, box = small_scene()
    prm = baseline_params()
    path = tmp_path / "dam.sphs"
    with wcsph.FluidSimulator(prm, fl, bd, box) as a, wcsph.FluidSimulator(prm, fl, bd, box) as b:
        a.step(5e-4, 20)
        st = a.get_state()
        snapshot.save(path, "dam_break", st["pos"], st["vel"])
        _, pos, vel = snapshot.load(path, app_name="dam_break", fluid_n=fl.shape[0])
        b.set_state(pos, vel)
        a.set_state(st["pos"], st["vel"])
        a.step(5e-4, 10)
        b.step(5e-4, 10)
        sa, sb = a.get_state(), b.get_state()
        np.testing.assert_array_equal(sa["pos"], sb["pos"])
        np.testing.assert_array_equal(sa["vel"], sb["vel"])


def test_autotune_keeps_state_and_applies_winner(tmp_path):
    fl, bd, box = small_scene()
    prm = baseline_params()
    cache = tmp_path / "tune.json"
    with wcsph.FluidSimulator(prm, fl, bd, box) as a, wcsph.FluidSimulator(prm, fl, bd, box) as b:
        a.step(5e-4, 5)
        b.step(5e-4, 5)
        before = a.get_state()
        best, results = autotune.tune(a, 5e-4, steps_per_trial=3, cache_path=str(cache))
        after = a.get_state()
        for k in ("pos", "vel", "ids"):
            np.testing.assert_array_equal(before[k], after[k])
        assert len(results) == 10 and all(ms > 0 for _, ms in results)
        assert sorted(d["cell_size_mult"] for d, _ in results)[:2] == pytest.approx([0.5, 2 / 3])
        assert autotune.variant_to_dict(a.get_variant()) == autotune.variant_to_dict(best)
        assert min(ms for _, ms in results) == pytest.approx(
            [ms for d, ms in results if d == autotune.variant_to_dict(best)][0])
        # the tuned variant steps to the same bits as the default one (every
        # form at cell size h gives the same bits; another cell size sums in
        # another order, so b then takes it too)
        if best.cell_size_mult != 1.0:
            b.set_variant(best)
        a.step(5e-4, 10)
        b.step(5e-4, 10)
        np.testing.assert_array_equal(a.get_state()["pos"], b.get_state()["pos"])
        # the cache reapplies the winner on a fresh simulator
        with wcsph.FluidSimulator(prm, fl, bd, box) as c:
            v = autotune.apply_cached(c, str(cache))
            assert v is not None
            assert autotune.variant_to_dict(c.get_variant()) == autotune.variant_to_dict(best)
