"""Edge sizes, in every neighbour form.

- Scenes smaller than a few wavefronts: 1, 2, 7, 63, 65 and 129 fluid
  particles (partial wavefronts and list quads; a lone particle with no
  fluid neighbour), taken as the particles nearest a corner of the walls, so
  boundary neighbours are in play.  The sorted order and the fluid and
  boundary lists equal the oracle's bit for bit; after five steps the state
  matches the FP64 oracle (positions 2e-6 m, density 1e-5 relative).
- A batch holding a scene without fluid and a one-particle scene beside an
  ordinary one: every scene ends exactly as when stepped alone, and the
  empty scene's clock advances.
"""
import numpy as np
import pytest

import wcsph
from oracle import oracle as O
from tests.helpers import baseline_params, by_id, cell_size, rel_err, small_scene

pytestmark = pytest.mark.gpu

MODES = [wcsph.abi.NEIGHBORS_CELL, wcsph.abi.NEIGHBORS_PARTICLE, wcsph.abi.NEIGHBORS_CELL_LOOP,
         wcsph.abi.NEIGHBORS_BRICK]
SIZES = [1, 2, 7, 63, 65, 129]
DT = 1e-4


def corner_cluster(n, seed):
    fl, bd, box = small_scene()
    lo = fl[:, :3].min(axis=0)
    order = np.argsort(np.linalg.norm(fl[:, :3] - lo, axis=1), kind="stable")
    pos = np.ascontiguousarray(fl[order[:n]])
    rng = np.random.default_rng(seed)
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.normal(0.0, 0.3, size=(n, 3)).astype(np.float32)
    return pos, vel, bd, box


def set_mode(sim, mode):
    v = sim.get_variant()
    v.neighbor_mode = mode
    sim.set_variant(v)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n", SIZES)
def test_tiny_scene_lists_and_state(n, mode):
    pos, vel, bd, box = corner_cluster(n, seed=n)
    prm = baseline_params()
    cs = cell_size(prm)
    with wcsph.FluidSimulator(prm, pos, bd, box, fluid_vel=vel) as sim:
        set_mode(sim, mode)
        st = sim.step(DT, 1)
        assert st.flags == 0
        sp = sim.sorted_positions()
        np.testing.assert_array_equal(sim.permutation(), O.sort_order(pos, cs))
        counts, lst = sim.neighbors(0)
        oc, ol = O.neighbors(sp, sp, cs, cs, True)
        np.testing.assert_array_equal(counts, oc)
        np.testing.assert_array_equal(lst, ol)
        bc, bl = sim.neighbors(1)
        obc, obl = O.neighbors(sp, sim.boundary(), cs, cs, False)
        np.testing.assert_array_equal(bc, obc)
        np.testing.assert_array_equal(bl, obl)
        st = sim.step(DT, 4)
        assert st.flags == 0 and st.steps == 5
        g = by_id(sim.get_state())
    if n >= 7:
        assert counts.max() > 0 and bc.max() > 0  # the cluster has fluid and wall neighbours
    ref = O.Solver(prm, pos, bd, box, fluid_vel=vel, precision=64)
    ref.step(DT, 5)
    r = by_id(ref.state())
    assert np.max(np.abs(g["pos"][:, :3] - r["pos"][:, :3])) <= 2e-6
    assert rel_err(g["rho"], r["rho"]) <= 1e-5


@pytest.mark.parametrize("mode", MODES)
def test_batch_with_empty_and_one_particle_scenes(mode):
    fl, bd, box = small_scene()
    rng = np.random.default_rng(11)
    vel = np.zeros_like(fl)
    vel[:, :3] = rng.normal(0.0, 0.3, size=(fl.shape[0], 3)).astype(np.float32)
    empty = np.zeros((0, 4), np.float32)
    one = np.ascontiguousarray(fl[:1])
    sc = [(empty, bd, box, empty), (one, bd, box, np.ascontiguousarray(vel[:1])), (fl, bd, box, vel),
          (empty, bd, box, empty)]
    prm = baseline_params(adaptive_dt=1)
    with wcsph.BatchSimulator(prm, sc) as b:
        set_mode(b, mode)
        b.step(1e-3, 6)
        st = b.get_state()
        per = [b.scene_stats(k) for k in range(len(sc))]
        assert b.n == 1 + fl.shape[0]
    for k, (f, bdk, bx, v) in enumerate(sc):
        with wcsph.FluidSimulator(prm, f, bdk, bx, fluid_vel=v if len(v) else None) as one_sim:
            set_mode(one_sim, mode)
            s1 = one_sim.step(1e-3, 6)
            g = one_sim.get_state()
        a, c = b.ranges[k][0], b.ranges[k][0] + b.ranges[k][1]
        np.testing.assert_array_equal(st["pos"][a:c], g["pos"])
        np.testing.assert_array_equal(st["vel"][a:c], g["vel"])
        np.testing.assert_array_equal(st["rho"][a:c], g["rho"])
        assert per[k].sim_time == s1.sim_time and per[k].steps == 6 and per[k].flags == 0


def test_grid_limits_are_refused():
    """Scenes too large for the cell grids are refused at creation, before
    any grid is allocated: beyond the reference's 21-bit Morton cell range
    (2^20 cells of h = 4 cm either side of 0), more than 2^28 cells in all
    (a 30 m cube), and a finite but huge particle coordinate."""
    from wcsph import scenes

    fl, _, _ = small_scene()
    prm = baseline_params()
    none = np.zeros((0, 4), np.float32)
    with pytest.raises(wcsph.WcsphError, match="21-bit Morton cell range") as e:
        wcsph.FluidSimulator(prm, fl, none, scenes.make_box((0.0, 0.0, 0.0), (1e5, 1.0, 1.0)))
    assert e.value.code == wcsph.abi.WCSPH_ERR_INVALID_ARGUMENT
    with pytest.raises(wcsph.WcsphError, match=r"at most 2\^28") as e:
        wcsph.FluidSimulator(prm, fl, none, scenes.make_box((0.0, 0.0, 0.0), (30.0, 30.0, 30.0)))
    assert e.value.code == wcsph.abi.WCSPH_ERR_INVALID_ARGUMENT
    far = fl.copy()
    far[0, 0] = 1e30
    with pytest.raises(wcsph.WcsphError, match="21-bit Morton cell range"):
        wcsph.FluidSimulator(prm, far, none, scenes.make_box((0.0, 0.0, 0.0), (1.0, 1.0, 1.0)))
