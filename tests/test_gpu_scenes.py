"""GPU parity of the scene extras: external force rules (the reference's
fountain and vortex lambdas), a box moved every step (tilting box), and
boundary particles sampled from meshes around a clustered fluid block.
Tolerances: density <= 1e-5 relative to the FP64 oracle after one step;
positions <= 1e-6 m after one step; N-step trajectories within 10x the CPU
FP32-vs-FP64 divergence (+1e-6 m); neighbour lists bit-exact."""
import math

import numpy as np
import pytest

import wcsph
from oracle import oracle as O
from tests.helpers import SMALL_REF, baseline_params, by_id, cell_size, reference_params, rel_err, small_scene
from tests.test_scene_geometry import FOUNTAIN, VORTEX
from wcsph import abi, scenes

pytestmark = pytest.mark.gpu


def centred_ref_scene(shift=(0.6, 0.55, 0.5)):
    """The 12^3 reference-physics dam break moved so that it straddles the
    fountain/vortex force regions (y < 0, x^2 + z^2 < 0.2)."""
    fl, bd, box = small_scene(SMALL_REF)
    t = np.array([*shift, 0.0], np.float32)
    fl = (fl - t).astype(np.float32)
    bd = (bd - t).astype(np.float32)
    box = scenes.make_box(np.array(box.min[:]) - t[:3], np.array(box.max[:]) - t[:3])
    return fl, bd, box


@pytest.mark.parametrize("rules", ["fountain", "vortex"])
def test_external_force_rules_one_step_vs_fp64(rules):
    prm = abi.set_force_rules(reference_params(), FOUNTAIN if rules == "fountain" else VORTEX)
    fl, bd, box = centred_ref_scene()
    assert (fl[:, 1] < -0.3).sum() > 100 and ((fl[:, 0] ** 2 + fl[:, 2] ** 2 < 0.2) & (fl[:, 1] < 0)).sum() > 100
    dt = 1e-3
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        sim.update(dt)
        g = by_id(sim.get_state())
    ref = O.Solver(prm, fl, bd, box, precision=64)
    ref.step(dt)
    r = by_id(ref.state())
    assert rel_err(g["rho"], r["rho"]) <= 1e-5
    assert np.max(np.abs(g["pos"][:, :3] - r["pos"][:, :3])) <= 1e-6
    vscale = np.max(np.abs(r["vel"][:, :3]))
    # the rules act over the non-pressure step dt_np = 1e-3: the fountain jet
    # lifts (25 - 9.81 m/s^2), the vortex swirls (2 m/s^2 in xz)
    if rules == "fountain":
        assert np.max(r["vel"][:, 1]) > 0.01
    else:
        assert np.max(np.hypot(r["vel"][:, 0], r["vel"][:, 2])) > 1.5e-3
    assert np.max(np.abs(g["vel"][:, :3] - r["vel"][:, :3])) <= 2e-5 * vscale + 1e-6


def test_external_force_rules_n_steps():
    prm = abi.set_force_rules(reference_params(), FOUNTAIN)
    fl, bd, box = centred_ref_scene()
    steps, dt = 40, 2e-3
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        sim.step(dt, steps)
        g = by_id(sim.get_state())
    r64 = O.Solver(prm, fl, bd, box, precision=64)
    r64.step(dt, steps)
    r32 = O.Solver(prm, fl, bd, box, precision=32)
    r32.step(dt, steps)
    a, b = by_id(r64.state()), by_id(r32.state())
    e_cpu = np.max(np.abs(b["pos"][:, :3] - a["pos"][:, :3]))
    e_gpu = np.max(np.abs(g["pos"][:, :3] - a["pos"][:, :3]))
    assert e_gpu <= 10 * e_cpu + 1e-6, (e_gpu, e_cpu)


def tilted_box(lo, hi, angle):
    """Box [lo, hi] rotated by `angle` about z through its centre
    (Box::set_model of the tilting-box scene, app/tilting_box/application.cu)."""
    c = (np.array(lo) + np.array(hi)) / 2
    m = np.eye(4)
    m[:2, :2] = [[math.cos(angle), -math.sin(angle)], [math.sin(angle), math.cos(angle)]]
    m[:3, 3] = c - m[:3, :3] @ c
    return scenes.make_box(lo, hi, model=m)


def test_moving_box_matches_oracle():
    prm = baseline_params()
    desc = scenes.DamBreak(dims=(14, 14, 14), box_max=(0.64, 0.5, 0.5), seed=5)
    fl, _, box0 = small_scene(desc)
    lo, hi = (0.0, 0.0, 0.0), (0.64, 0.5, 0.5)
    half_diag = 0.5 * math.sqrt(0.64 ** 2 + 0.5 ** 2 + 0.5 ** 2)
    c = np.array([0.32, 0.25, 0.25])
    domain = (c - half_diag - 0.05, c + half_diag + 0.05)
    empty = np.zeros((0, 4), np.float32)
    steps, dt = 60, 5e-4
    ref64 = O.Solver(prm, fl, empty, box0, precision=64)
    ref32 = O.Solver(prm, fl, empty, box0, precision=32)
    with wcsph.FluidSimulator(prm, fl, empty, box0, domain=domain) as sim:
        for k in range(steps):
            box = tilted_box(lo, hi, -0.01 * k)  # the floor under the fluid rises
            sim.set_box(box)
            ref64.set_box(box)
            ref32.set_box(box)
            sim.update(dt)
            ref64.step(dt)
            ref32.step(dt)
        g = by_id(sim.get_state())
        got = sim.get_box()
        np.testing.assert_array_equal(np.array(got.model[:]), np.array(box.model[:]))
        # a box leaving the grid is refused and left unchanged
        with pytest.raises(wcsph.WcsphError) as e:
            sim.set_box(scenes.make_box((-5, 0, 0), (0.64, 0.5, 0.5)))
        assert e.value.code == abi.WCSPH_ERR_DOMAIN
        np.testing.assert_array_equal(np.array(sim.get_box().model[:]), np.array(box.model[:]))
    a, b = by_id(ref64.state()), by_id(ref32.state())
    e_cpu = np.max(np.abs(b["pos"][:, :3] - a["pos"][:, :3]))
    e_gpu = np.max(np.abs(g["pos"][:, :3] - a["pos"][:, :3]))
    assert e_gpu <= 10 * e_cpu + 1e-6, (e_gpu, e_cpu)
    # the tilted walls were hit: particles sit r inside the rotated floor
    inv = np.array(box.model_inv[:]).reshape(4, 4).T
    local = (inv @ np.c_[a["pos"][:, :3], np.ones(len(fl))].T).T
    assert np.sum(np.isclose(local[:, 1], float(np.float32(prm.particle_radius)), atol=1e-4)) > 0


def sphere_in_cluster_scene():
    """A clustered block around a mesh-sampled sphere (fluid inside the
    sphere and its voxel shell removed)."""
    r = 0.003125
    fl = scenes.clustered_block(r, (-0.1, -0.1, -0.1), (0.1, 0.1, 0.1), 0.1, 0.025, seed=9)
    R = 0.05
    keep = np.linalg.norm(fl[:, :3], axis=1) > R + 2.5 * r
    fl = np.ascontiguousarray(fl[keep])
    mesh = scenes.procedural_mesh(abi.MESH_SPHERE, (R,), 24)
    bd = scenes.sample_mesh(mesh, 2 * r)
    box = scenes.make_box((-0.11, -0.11, -0.11), (0.11, 0.11, 0.11))
    return r, fl, bd, box


@pytest.mark.parametrize("mode", [abi.NEIGHBORS_CELL, abi.NEIGHBORS_PARTICLE])
def test_mesh_boundary_and_clusters_bit_exact_and_density(mode):
    r, fl, bd, box = sphere_in_cluster_scene()
    prm = baseline_params(particle_radius=r)
    cs = cell_size(prm)
    with wcsph.FluidSimulator(prm, fl, bd, box) as sim:
        v = sim.get_variant()
        v.neighbor_mode = mode
        v.fluid_capacity = 128
        sim.set_variant(v)
        sim.update(1e-5)
        g = by_id(sim.get_state())
        sp = sim.sorted_positions()
        np.testing.assert_array_equal(sim.permutation(), O.sort_order(fl, cs))
        counts, lst = sim.neighbors(0)
        oc, ol = O.neighbors(sp, sp, cs, cs, True)
        np.testing.assert_array_equal(counts, oc)
        np.testing.assert_array_equal(lst, ol)
        bgpu = sim.boundary()
        bcounts, blst = sim.neighbors(1)
        obc, obl = O.neighbors(sp, bgpu, cs, cs, False)
        np.testing.assert_array_equal(bcounts, obc)
        np.testing.assert_array_equal(blst, obl)
        assert bcounts.max() > 8 and counts.max() > counts.mean() * 1.3
    ref = O.Solver(prm, fl, bd, box, precision=64)
    ref.step(1e-5)
    rr = by_id(ref.state())
    assert rel_err(g["rho"], rr["rho"]) <= 1e-5
