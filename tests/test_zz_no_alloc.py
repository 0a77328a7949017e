"""wcsph_autotune_noalloc, the autotuner that allocates no device memory,
against wcsph_autotune, through the C ABI.

Two calls in a row on one GPU, the second on a smaller handle (the scratch
array then still holds the first call's larger state).  For each, two
handles with the same state and candidates are tuned, one by each entry
point.  Afterwards both hold exactly the state they had before (positions,
velocities, ids, time and step count: bit-exact), both report finite
positive trial times and a winner from the candidates, and once both apply
the same variant, ten more steps give the same bits on both and on a third
handle that was never tuned.  The refusals (a candidate that changes the
list capacities, a handle larger than the scratch array) are checked too;
they return before any trial runs.
"""
import ctypes

import numpy as np
import pytest

import wcsph
from tests.helpers import baseline_params
from wcsph import abi, scenes

pytestmark = pytest.mark.gpu

HINT = 1e-3


def _state(dims, box_max, seed):
    fl, bd, box = scenes.dam_break(scenes.DamBreak(dims=dims, box_max=box_max, seed=seed))
    rng = np.random.default_rng(seed)
    lo = fl[:, :3].min(axis=0)
    pos = fl.copy()
    pos[:, :3] = (lo + (fl[:, :3] - lo) * np.float32(0.88)).astype(np.float32)
    vel = np.zeros_like(fl)
    vel[:, :3] = rng.normal(0.0, 0.4, size=(fl.shape[0], 3)).astype(np.float32)
    return fl, bd, box, pos, vel


def _sim(prm, st):
    fl, bd, box, pos, vel = st
    sim = wcsph.FluidSimulator(prm, fl, bd, box)
    sim.set_state(pos, vel)
    sim.step(HINT, 2)
    return sim


def _candidates(base):
    out = []
    for mode, bs, groups in ((abi.NEIGHBORS_PARTICLE, 128, base.neighbor_groups),
                             (abi.NEIGHBORS_PARTICLE, 256, base.neighbor_groups),
                             (abi.NEIGHBORS_CELL, base.block_size, 4096),
                             (abi.NEIGHBORS_CELL_LOOP, 256, base.neighbor_groups)):
        v = abi.Variant()
        for f, _ in abi.Variant._fields_:
            setattr(v, f, getattr(base, f))
        v.neighbor_mode, v.block_size, v.neighbor_groups = mode, bs, groups
        out.append(v)
    return out


def _tune(sim, entry, cands):
    arr = (abi.Variant * len(cands))(*cands)
    ms = (ctypes.c_double * len(cands))()
    best = abi.Variant()
    rc = getattr(sim.lib, entry)(sim.handle, HINT, 3, ctypes.cast(arr, ctypes.c_void_p), len(cands), ms,
                                 ctypes.byref(best))
    return rc, list(ms), best


def _snapshot(sim):
    g = sim.get_state()
    st = sim.lib.wcsph_scene_stats
    s = abi.Stats()
    assert st(sim.handle, 0, ctypes.byref(s)) == 0
    return g, (s.sim_time, s.last_dt, s.steps)


def _fields(v):
    return tuple(getattr(v, f) for f, _ in abi.Variant._fields_)


def test_noalloc_autotune_matches_autotune_on_two_calls():
    prm = baseline_params(adaptive_dt=1)
    for dims, box_max, seed in (((24, 16, 16), (0.96, 0.5, 0.5), 3), ((10, 10, 8), (0.5, 0.4, 0.4), 4)):
        print(f"no-alloc autotune: lattice {dims}", flush=True)
        st = _state(dims, box_max, seed)
        a, b, c = _sim(prm, st), _sim(prm, st), _sim(prm, st)
        try:
            cands = _candidates(a.get_variant())
            before_a, before_b = _snapshot(a), _snapshot(b)
            rc_a, ms_a, best_a = _tune(a, "wcsph_autotune", cands)
            rc_b, ms_b, best_b = _tune(b, "wcsph_autotune_noalloc", cands)
            assert rc_a == 0 and rc_b == 0, (rc_a, rc_b, b.lib.wcsph_last_error(b.handle))
            for ms in (ms_a, ms_b):
                assert all(np.isfinite(m) and m > 0 for m in ms), ms
            assert _fields(best_b) in {_fields(v) for v in cands}
            assert _fields(b.get_variant()) == _fields(best_b)
            after_a, after_b = _snapshot(a), _snapshot(b)
            for (g0, t0), (g1, t1) in ((before_a, after_a), (before_b, after_b), (after_a, after_b)):
                for k in ("pos", "vel", "ids"):
                    np.testing.assert_array_equal(g0[k], g1[k])
                assert t0 == t1
            # the same variant on all three handles: the same bits after 10 steps
            for s in (a, b, c):
                s.set_variant(best_a)
                s.step(HINT, 10)
            ga, gb, gc = a.get_state(), b.get_state(), c.get_state()
            for k in ("pos", "vel", "rho", "ids"):
                np.testing.assert_array_equal(ga[k], gb[k])
                np.testing.assert_array_equal(ga[k], gc[k])
        finally:
            for s in (a, b, c):
                s.close()


def test_noalloc_autotune_refusals():
    prm = baseline_params()
    st = _state((10, 10, 8), (0.5, 0.4, 0.4), 5)
    sim = _sim(prm, st)
    try:
        before, t_before = _snapshot(sim)
        cands = _candidates(sim.get_variant())
        cands[1].fluid_capacity = cands[1].fluid_capacity + 4
        rc, _, _ = _tune(sim, "wcsph_autotune_noalloc", cands)
        assert rc == abi.WCSPH_ERR_INVALID_ARGUMENT
        assert b"capacities" in sim.lib.wcsph_last_error(sim.handle)
        after, t_after = _snapshot(sim)
        np.testing.assert_array_equal(before["pos"], after["pos"])
        assert t_before == t_after
    finally:
        sim.close()
    # 128 x 128 x 128 = 2,097,152 particles: 75.5 MB of state, above the array
    fl, bd, box = scenes.dam_break(scenes.DamBreak(dims=(128, 128, 128), box_max=(5.2, 2.6, 2.6), seed=6))
    with wcsph.FluidSimulator(prm, fl, bd, box) as big:
        rc, _, _ = _tune(big, "wcsph_autotune_noalloc", _candidates(big.get_variant()))
        assert rc == abi.WCSPH_ERR_INVALID_ARGUMENT
        assert b"64 MiB" in big.lib.wcsph_last_error(big.handle)
