"""The 1,048,576-particle dam break (CFL dt, lambda 0.4, hint 5e-4) over the
first 250 steps of its 500-step run, against the FP64 oracle.

Every 50 steps: kinetic energy, centre of mass, mean density and the
simulated time (the sum of the CFL dt actually taken) of the GPU against
the FP64 oracle's; at the end the per-particle positions (99th percentile
and median of the error).  Each tolerance is 10x how far the oracle's own
FP32 step drifts from its FP64 step over the same steps, plus a small
absolute floor.  The oracle runs both precisions over the 250 steps (about
50 s on 16 host threads; the whole run would take the test past the
per-test time limits GPU runs use).
"""
import numpy as np
import pytest

import wcsph
from oracle import oracle as O
from tests.helpers import by_id
from tests.test_gpu_c1_run import aggregates
from wcsph import abi, scenes

pytestmark = pytest.mark.gpu

HINT, CHUNK, CHUNKS = 5e-4, 50, 5


def test_c2_first_250_steps_vs_fp64():
    fl, bd, box = scenes.dam_break(scenes.DAM_BREAK_1M)
    assert fl.shape[0] == 1048576
    prm = scenes.preset_params(abi.PRESET_BENCHMARK, adaptive_dt=1, cfl_factor=0.4)
    gpu = []
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        for _ in range(CHUNKS):
            st = sim.step(HINT, CHUNK)
            assert st.flags == 0
            gpu.append((by_id(sim.get_state()), st.sim_time))
    r64 = O.Solver(prm, fl, bd, box, precision=64)
    r32 = O.Solver(prm, fl, bd, box, precision=32)
    for c in range(CHUNKS):
        r64.step(HINT, CHUNK)
        r32.step(HINT, CHUNK)
        a, b = by_id(r64.state()), by_id(r32.state())
        g, tg = gpu[c]
        ta, tb = r64.time()[0], r32.time()[0]
        ka, ca, da = aggregates(a)
        kb, cb, db = aggregates(b)
        kg, cg, dg = aggregates(g)
        step = CHUNK * (c + 1)
        assert abs(tg - ta) <= 10 * abs(tb - ta) + 1e-6 * ta, (step, tg, ta, tb)
        assert abs(kg - ka) <= 10 * abs(kb - ka) + 1e-4 * ka, (step, kg, ka, kb)
        assert np.max(np.abs(cg - ca)) <= 10 * np.max(np.abs(cb - ca)) + 1e-6, (step, cg, ca, cb)
        assert abs(dg - da) <= 10 * abs(db - da) + 1e-6 * da, (step, dg, da, db)
    dx_cpu = np.max(np.abs(b["pos"][:, :3] - a["pos"][:, :3]), axis=1)
    dx_gpu = np.max(np.abs(g["pos"][:, :3] - a["pos"][:, :3]), axis=1)
    p_cpu, p_gpu = np.percentile(dx_cpu, 99), np.percentile(dx_gpu, 99)
    assert p_gpu <= 10 * p_cpu + 1e-6, (p_gpu, p_cpu)
    assert np.median(dx_gpu) <= 10 * np.median(dx_cpu) + 1e-7
