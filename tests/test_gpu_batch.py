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


def _run_batch(prm, sc, modes=None, steps=10, variant_mode=None):
    with wcsph.BatchSimulator(prm, sc) as b:
        if variant_mode is not None:
            v = b.get_variant()
            v.neighbor_mode = variant_mode
            b.set_variant(v)
        if modes is not None:
            b.set_scene_neighbor_modes(modes)
            np.testing.assert_array_equal(b.get_scene_neighbor_modes(), modes)
        st = b.step(1e-3, steps)
        assert st.flags == 0
        return b.get_state(), b.neighbors(0), b.neighbors(1)


def test_per_scene_builders_equal_one_builder():
    """Scenes of one batch built by different list builders (alternating the
    cell and per-particle builders, and each one alone) give the bits of the
    batch built by one builder: the lists and every state array."""
    prm = baseline_params(adaptive_dt=1)
    sc = _scenes(7, 12)
    ref_state, ref_f, ref_b = _run_batch(prm, sc, variant_mode=wcsph.abi.NEIGHBORS_PARTICLE)
    C, P = wcsph.abi.NEIGHBORS_CELL, wcsph.abi.NEIGHBORS_PARTICLE
    for modes in ([C, P, C, P, C, P, C], [P] * 7, [C] * 7, [P, P, P, C, C, C, P]):
        for variant_mode in (C, P):
            st, lf, lb = _run_batch(prm, sc, np.array(modes, np.uint32), variant_mode=variant_mode)
            for k in ("pos", "vel", "rho", "p", "ids"):
                np.testing.assert_array_equal(st[k], ref_state[k])
            for got, ref in ((lf, ref_f), (lb, ref_b)):
                np.testing.assert_array_equal(got[0], ref[0])
                np.testing.assert_array_equal(got[1], ref[1])


def test_per_scene_builders_are_checked():
    prm = baseline_params(adaptive_dt=1)
    sc = _scenes(3, 13)
    with wcsph.BatchSimulator(prm, sc) as b:
        with pytest.raises(wcsph.WcsphError, match="CELL or WCSPH_NEIGHBORS_PARTICLE"):
            b.set_scene_neighbor_modes([wcsph.abi.NEIGHBORS_CELL_LOOP] * 3)
        with pytest.raises(ValueError, match="one builder per scene"):
            b.set_scene_neighbor_modes([wcsph.abi.NEIGHBORS_CELL] * 2)
        # variant's builder everywhere until set
        v = b.get_variant()
        np.testing.assert_array_equal(b.get_scene_neighbor_modes(), [v.neighbor_mode] * 3)
        b.set_scene_neighbor_modes([wcsph.abi.NEIGHBORS_CELL, wcsph.abi.NEIGHBORS_PARTICLE,
                                    wcsph.abi.NEIGHBORS_PARTICLE])
        v.neighbor_mode, v.cell_size_mult = wcsph.abi.NEIGHBORS_PARTICLE, 2.0 / 3.0
        with pytest.raises(wcsph.WcsphError, match="per-particle builder for every scene"):
            b.set_variant(v)
        b.set_scene_neighbor_modes(None)
        b.set_variant(v)  # now every scene uses the per-particle builder
        with pytest.raises(wcsph.WcsphError, match="per-particle builder for every scene"):
            b.set_scene_neighbor_modes([wcsph.abi.NEIGHBORS_CELL] * 3)
        assert b.step(1e-3, 2).flags == 0


def test_size_bucket_tuner_applies_builders():
    """The size-bucket tuner times both builders on each bucket's scenes (a
    batch of their own) and sets each scene's builder; the batch then steps
    to the bits of one builder."""
    from wcsph import autotune
    prm = baseline_params(adaptive_dt=1)
    sc = [scenes.dam_break(d) for d in scenes.batch_dam_breaks(n_scenes=6, lo=3000, hi=90000, seed=44)]
    counts = [s[0].shape[0] for s in sc]
    idx, groups = autotune.scene_buckets(counts)
    assert len(groups) >= 2  # the scenes span size buckets
    with wcsph.BatchSimulator(prm, sc) as b:
        modes, report = autotune.tune_scene_buckets(prm, sc, 5e-4, batch=b, steps_per_trial=2)
        assert len(modes) == len(sc) and set(modes) <= {wcsph.abi.NEIGHBORS_CELL, wcsph.abi.NEIGHBORS_PARTICLE}
        buckets = [r for r in report if "bucket" in r]
        assert sorted(r["bucket"] for r in buckets) == sorted(groups)
        for r in buckets:
            assert len(r["trials"]) == 3 and all(ms > 0 for _, ms in r["trials"])
        mixed = [r for r in report if "mixed_ms_per_step" in r]
        if mixed and len(set(modes)) == 1:  # the mixed builders lost to one builder on the whole batch
            assert mixed[0]["mixed_ms_per_step"] >= mixed[0]["one_builder_ms_per_step"]
            assert modes[0] == mixed[0]["one_builder"]
        else:
            for r in buckets:
                assert all(m == r["builder"] for k, m in enumerate(modes) if idx[k] == r["bucket"])
        np.testing.assert_array_equal(b.get_scene_neighbor_modes(), modes)
        assert b.step(5e-4, 5).flags == 0
        got = b.get_state()
    ref, _, _ = _run_batch(prm, sc, steps=5, variant_mode=wcsph.abi.NEIGHBORS_PARTICLE)
    for k in ("pos", "vel", "rho"):
        np.testing.assert_array_equal(got[k], ref[k])
