"""Batched form: many uneven scenes stepped together must give, per scene,
exactly the state of stepping that scene alone (same kernels, same order),
including per-scene CFL dt; each batched scene also matches the FP64
oracle stepping that scene (the tolerances of the single-scene tests), and
a seeded subset of the 256-scene benchmark set matches stepping alone."""
import numpy as np
import pytest

import wcsph
from oracle import oracle as O
from tests.helpers import baseline_params, by_id, rel_err
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


def _stirred_scenes(seed):
    """Three scenes of 1k-4k particles, compressed and moving (pressure and
    viscosity active), each with its own speed so the CFL dt differs."""
    out = []
    rng = np.random.default_rng(seed)
    for k, d in enumerate(scenes.batch_dam_breaks(n_scenes=3, lo=1000, hi=4000, seed=seed)):
        fl, bd, box = scenes.dam_break(d)
        lo = fl[:, :3].min(axis=0)
        pos = fl.copy()
        pos[:, :3] = (lo + (fl[:, :3] - lo) * np.float32(0.88)).astype(np.float32)
        vel = np.zeros_like(fl)
        vel[:, :3] = rng.normal(0.0, 0.3 * (k + 1), size=(fl.shape[0], 3)).astype(np.float32)
        out.append((pos, bd, box, vel))
    return out


def test_batched_scenes_vs_fp64_oracle():
    """One CFL step of a 3-scene batch against the FP64 oracle stepping each
    scene: density <= 1e-5 relative, pressure / k <= 1e-4, positions <= 1e-6 m;
    then 30 steps within 10x the CPU FP32-vs-FP64 divergence (+1e-6 m)."""
    prm = baseline_params(adaptive_dt=1)
    sc = _stirred_scenes(21)
    with wcsph.BatchSimulator(prm, sc) as b:
        b.step(1e-3, 1)
        one = b.get_state()
        b.step(1e-3, 29)
        many = b.get_state()
        ranges = list(b.ranges)
    for k, (pos, bd, box, vel) in enumerate(sc):
        p0, cnt = ranges[k][0], ranges[k][1]
        sl = slice(p0, p0 + cnt)

        def scene_state(st):
            return by_id({key: st[key][sl] for key in ("pos", "vel", "rho", "p")} | {"ids": st["ids"][sl] - p0})

        g1, gn = scene_state(one), scene_state(many)
        r64 = O.Solver(prm, pos, bd, box, fluid_vel=vel, precision=64)
        r64.step(1e-3, 1)
        a1 = O.by_id(r64.state())
        assert rel_err(g1["rho"], a1["rho"]) <= 1e-5
        assert np.max(np.abs(g1["p"] - a1["p"])) / prm.stiffness <= 1e-4
        assert np.max(a1["p"]) > 0.0
        assert np.max(np.abs(g1["pos"][:, :3] - a1["pos"][:, :3])) <= 1e-6
        r64.step(1e-3, 29)
        r32 = O.Solver(prm, pos, bd, box, fluid_vel=vel, precision=32)
        r32.step(1e-3, 30)
        an, bn = O.by_id(r64.state()), O.by_id(r32.state())
        e_cpu = np.max(np.abs(bn["pos"][:, :3] - an["pos"][:, :3]))
        e_gpu = np.max(np.abs(gn["pos"][:, :3] - an["pos"][:, :3]))
        assert e_gpu <= 10 * e_cpu + 1e-6, (k, e_gpu, e_cpu)


def test_benchmark_batch_subset_equals_scenes_alone():
    """32 scenes of the 256-scene benchmark set (every 8th by size, 2k to
    200k particles), stepped together for 5 CFL steps from rest, equal each
    scene stepped alone, bit for bit (state, time, dt)."""
    descs = scenes.batch_dam_breaks()
    order = np.argsort([int(np.prod(d.dims)) for d in descs], kind="stable")
    pick = [descs[i] for i in order[::8]]
    sizes = [int(np.prod(d.dims)) for d in pick]
    assert len(pick) == 32 and sizes[0] < 3000 and sizes[-1] > 150000
    prm = baseline_params(adaptive_dt=1)
    sc = [scenes.dam_break(d) for d in pick]
    with wcsph.BatchSimulator(prm, sc) as b:
        b.step(5e-4, 5)
        st = b.get_state()
        per = [b.scene_stats(k) for k in range(len(sc))]
        ranges = list(b.ranges)
    for k, (fl, bd, box) in enumerate(sc):
        with wcsph.FluidSimulator(prm, fl, bd, box) as one:
            s1 = one.step(5e-4, 5)
            g = one.get_state()
            assert s1.flags == 0
        p0, cnt = ranges[k][0], ranges[k][1]
        for key in ("pos", "vel", "rho", "p"):
            np.testing.assert_array_equal(st[key][p0:p0 + cnt], g[key])
        np.testing.assert_array_equal(st["ids"][p0:p0 + cnt] - p0, g["ids"])
        assert (per[k].sim_time, per[k].last_dt, per[k].steps) == (s1.sim_time, s1.last_dt, s1.steps)
