"""The 32^3 dam break (32,768 fluid particles, fixed dt 1e-4) over its whole
1000-step run, against the FP64 oracle.

Every 100 steps: the GPU's kinetic energy, centre of mass and mean density
against the FP64 oracle's, and at the end the per-particle positions.  The
tolerance of each is 10x how far the oracle's own FP32 step drifts from its
FP64 step over the same steps (plus a small absolute floor), so the test
measures the HIP step against the FP32 rounding that any FP32 implementation
of the reference's step carries.  Over this run the FP32 drift stays at
about 1e-5 in kinetic energy and 1e-7 m in the centre of mass; single
particles can drift by millimetres after a wall collision decides
differently, so positions are compared by their 99th percentile.
"""
import numpy as np
import pytest

import wcsph
from oracle import oracle as O
from tests.helpers import baseline_params, by_id
from wcsph import scenes

pytestmark = pytest.mark.gpu

DT, CHUNK, CHUNKS = 1e-4, 100, 10


def aggregates(st):
    v = st["vel"][:, :3].astype(np.float64)
    x = st["pos"][:, :3].astype(np.float64)
    return 0.5 * np.sum(v * v), x.mean(axis=0), np.asarray(st["rho"], np.float64).mean()


def test_c1_full_run_vs_fp64():
    fl, bd, box = scenes.dam_break(scenes.DAM_BREAK_32K)
    assert fl.shape[0] == 32768
    prm = baseline_params()
    assert prm.adaptive_dt == 0  # the benchmark preset steps the hint as given
    gpu = []
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        for _ in range(CHUNKS):
            st = sim.step(DT, CHUNK)
            assert st.flags == 0
            gpu.append(by_id(sim.get_state()))
        assert st.steps == CHUNK * CHUNKS
    r64 = O.Solver(prm, fl, bd, box, precision=64)
    r32 = O.Solver(prm, fl, bd, box, precision=32)
    for c in range(CHUNKS):
        r64.step(DT, CHUNK)
        r32.step(DT, CHUNK)
        a, b, g = by_id(r64.state()), by_id(r32.state()), gpu[c]
        ka, ca, da = aggregates(a)
        kb, cb, db = aggregates(b)
        kg, cg, dg = aggregates(g)
        step = CHUNK * (c + 1)
        assert abs(kg - ka) <= 10 * abs(kb - ka) + 1e-4 * ka, (step, kg, ka, kb)
        assert np.max(np.abs(cg - ca)) <= 10 * np.max(np.abs(cb - ca)) + 1e-6, (step, cg, ca, cb)
        assert abs(dg - da) <= 10 * abs(db - da) + 1e-6 * da, (step, dg, da, db)
    dx_cpu = np.max(np.abs(b["pos"][:, :3] - a["pos"][:, :3]), axis=1)
    dx_gpu = np.max(np.abs(g["pos"][:, :3] - a["pos"][:, :3]), axis=1)
    p_cpu, p_gpu = np.percentile(dx_cpu, 99), np.percentile(dx_gpu, 99)
    assert p_gpu <= 10 * p_cpu + 1e-6, (p_gpu, p_cpu)
    assert np.median(dx_gpu) <= 10 * np.median(dx_cpu) + 1e-7
