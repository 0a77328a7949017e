"""Full-size GPU checks of the two large single-scene workloads (sizes the
oracle cannot step in test time, so size-independent properties and sampled
parity against the oracle on the GPU's own exported state):

- the 1,048,576-particle dam break (128 x 128 x 64 lattice, CFL dt), which
  runs the per-particle list builder (the default above 131,072 particles)
  and the 32-bit offsets of every particle array at full size;
- the 8.3M-particle obstacle collapse (clustered block, sphere + torus
  boundary), whose 10.8M-cell grid exercises the scan's clearing of the
  previous step's occupied cells and a 62-neighbour maximum.

For 2,048 seeded sample particles of the last step: the fluid and boundary
neighbour lists equal the oracle's reference-order lists over the whole
exported sorted state (bit-exact), and the density equals an FP64 sum over
those lists with the exported positions and boundary masses (1e-5
relative).  Every step reports no flag.
"""
import math

import numpy as np
import pytest

import wcsph
from oracle import oracle as O
from tests.helpers import cell_size
from wcsph import abi, scenes

pytestmark = pytest.mark.gpu

N_SAMPLE = 2048


def cubic_w(r, h):
    q = np.asarray(r, np.float64) / h
    sig = 8.0 / (math.pi * h ** 3)
    return np.where(q <= 0.5, sig * (6 * q ** 3 - 6 * q ** 2 + 1), np.where(q <= 1.0, 2 * sig * (1 - q) ** 3, 0.0))


def check_sampled_parity(sim, prm, seed):
    h = float(cell_size(prm))
    r = np.float32(prm.particle_radius)
    s = np.float32(2.0) * r
    mass = float(np.float32(prm.rest_density) * (s * s * s * np.float32(0.8)))
    sp = sim.sorted_positions()
    bpos = sim.boundary()
    rho = sim.get_state()["rho"]
    lists = [sim.neighbors(0), sim.neighbors(1)]
    n = sp.shape[0]
    idx = np.sort(np.random.default_rng(seed).choice(n, N_SAMPLE, replace=False))
    q = np.ascontiguousarray(sp[idx])
    offs = [np.concatenate([[0], np.cumsum(c.astype(np.int64))]) for c, _ in lists]
    # fluid lists: the oracle keeps the self entry when the query is not the
    # source array itself; it is dropped from the reference list here
    oc, ol = O.neighbors(q, sp, h, h, False)
    ooff = np.concatenate([[0], np.cumsum(oc.astype(np.int64))])
    bsrc = np.ascontiguousarray(bpos)
    bc, bl = O.neighbors(q, bsrc, h, h, False)
    boff = np.concatenate([[0], np.cumsum(bc.astype(np.int64))])
    for k, i in enumerate(idx):
        ref = ol[ooff[k]:ooff[k + 1]]
        ref = ref[ref != i]
        got = lists[0][1][offs[0][i]:offs[0][i + 1]]
        np.testing.assert_array_equal(got, ref)
        gotb = lists[1][1][offs[1][i]:offs[1][i + 1]]
        np.testing.assert_array_equal(gotb, bl[boff[k]:boff[k + 1]])
        xi = sp[i, :3].astype(np.float64)
        df = cubic_w(np.linalg.norm(sp[got, :3].astype(np.float64) - xi, axis=1), h).sum()
        db = float(np.dot(cubic_w(np.linalg.norm(bpos[gotb, :3].astype(np.float64) - xi, axis=1), h),
                          bpos[gotb, 3].astype(np.float64)))
        expect = mass * (cubic_w(0.0, h) + df) + db
        assert abs(float(rho[i]) - expect) <= 1e-5 * expect, (i, rho[i], expect)
    return lists


def test_dam_break_1m_full_size():
    desc = scenes.DAM_BREAK_1M
    prm = scenes.preset_params(abi.PRESET_BENCHMARK, adaptive_dt=1, cfl_factor=0.4)
    fl, bd, box = scenes.dam_break(desc)
    assert fl.shape[0] == 1 << 20
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        v = sim.get_variant()
        assert v.neighbor_mode == abi.NEIGHBORS_PARTICLE  # the size-based default at this size
        st = sim.step(5e-4, 20)
        assert st.flags == 0 and st.steps == 20
        assert 0 < st.max_fluid_neighbors <= v.fluid_capacity
        lists = check_sampled_parity(sim, prm, seed=1)
        # every particle has neighbours and the lists cover the whole state
        assert lists[0][0].min() > 0
        assert (lists[1][0] > 0).sum() > 0  # the wall layers are reached
        g = sim.get_state()
        assert np.all(np.isfinite(g["pos"])) and np.array_equal(np.sort(g["ids"]), np.arange(fl.shape[0]))


def test_obstacle_collapse_8m_full_size():
    desc = scenes.OBSTACLE_COLLAPSE_8M
    prm = scenes.preset_params(abi.PRESET_BENCHMARK, adaptive_dt=1, cfl_factor=0.4, particle_radius=desc.r)
    fl, bd, box = scenes.obstacle_collapse(desc)
    assert fl.shape[0] > 8_000_000
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        st = sim.step(5e-4, 4)
        assert st.flags == 0 and st.steps == 4
        assert st.max_fluid_neighbors > 45  # the clusters at 1.8 r
        check_sampled_parity(sim, prm, seed=2)
