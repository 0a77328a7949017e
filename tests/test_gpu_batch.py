"""Batched form: many uneven scenes stepped together must give, per scene,
exactly the state of stepping that scene alone (same kernels, same order),
including per-scene CFL dt."""
import numpy as np
import pytest

import wcsph
from tests.helpers import baseline_params
from wcsph import scenes

pytestmark = pytest.mark.gpu


def _scenes(n, seed):
    descs = scenes.batch_dam_breaks(n_scenes=n, lo=300, hi=6000, seed=seed)
    out = []
    rng = np.random.default_rng(seed)
    for d in descs:
        fl, bd, box = scenes.dam_break(d)
        vel = np.zeros_like(fl)
        vel[:, :3] = rng.normal(0, rng.uniform(0.1, 6.0), size=(fl.shape[0], 3)).astype(np.float32)
        out.append((fl, bd, box, vel))
    return out


@pytest.mark.parametrize("mode", [wcsph.abi.NEIGHBORS_PARTICLE, wcsph.abi.NEIGHBORS_CELL,
                                  wcsph.abi.NEIGHBORS_CELL_LOOP, wcsph.abi.NEIGHBORS_BRICK])
def test_batch_equals_scenes_stepped_alone(mode):
    prm = baseline_params(adaptive_dt=1)
    sc = _scenes(7, 5)
    with wcsph.BatchSimulator(prm, sc) as b:
        v = b.get_variant()
        v.neighbor_mode = mode
        b.set_variant(v)
        b.step(1e-3, 12)
        st = b.get_state()
        per = [b.scene_stats(k) for k in range(len(sc))]
        assert b.n == sum(x[0].shape[0] for x in sc)
    dts = set()
    for k, (fl, bd, box, vel) in enumerate(sc):
        with wcsph.FluidSimulator(prm, fl, bd, box, fluid_vel=vel) as one:
            v = one.get_variant()
            v.neighbor_mode = mode
            one.set_variant(v)
            s1 = one.step(1e-3, 12)
            g = one.get_state()
        sl = b.ranges[k][0], b.ranges[k][0] + b.ranges[k][1]
        np.testing.assert_array_equal(st["pos"][sl[0]:sl[1]], g["pos"])
        np.testing.assert_array_equal(st["vel"][sl[0]:sl[1]], g["vel"])
        np.testing.assert_array_equal(st["rho"][sl[0]:sl[1]], g["rho"])
        np.testing.assert_array_equal(st["ids"][sl[0]:sl[1]] - sl[0], g["ids"])
        assert per[k].sim_time == s1.sim_time and per[k].last_dt == s1.last_dt
        dts.add(per[k].last_dt)
    assert len(dts) > 1  # the scenes really ran with their own CFL dt


def test_batch_graph_replay_and_reset():
    prm = baseline_params(adaptive_dt=1)
    sc = _scenes(5, 9)
    out = []
    for use_graph in (0, 1):
        with wcsph.BatchSimulator(prm, sc) as b:
            v = b.get_variant()
            v.use_graph = use_graph
            b.set_variant(v)
            b.step(1e-3, 6)
            b.reset()
            b.step(1e-3, 6)
            out.append(b.get_state())
    for k in ("pos", "vel", "rho", "ids"):
        np.testing.assert_array_equal(out[0][k], out[1][k])
