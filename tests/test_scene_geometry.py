"""Scene set-up on the host: external force rules (in the oracle),
procedural meshes, the mesh voxel sampler and the clustered fluid block."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import baseline_params, reference_params
from wcsph import abi, scenes

F32 = np.float32

FOUNTAIN = [abi.force_rule(abi.REGION_XZ_DISC_Y_BELOW, abi.FORCE_CONSTANT, y_below=0.0, radius2=0.2,
                           value=(0.0, 25.0, 0.0)),
            abi.force_rule(abi.REGION_Y_BELOW, abi.FORCE_TOWARD_Y_AXIS, y_below=-0.3, magnitude=2.0)]
VORTEX = [abi.force_rule(abi.REGION_Y_BELOW, abi.FORCE_SWIRL_Y_AXIS, y_below=-0.5, magnitude=2.0)]


def fountain_lambda(p):
    """The fountain scene's EXTERNAL_FORCE (app/fountain/application.cu:41-46)."""
    x, y, z = p
    if x * x + z * z < 0.2 and y < 0.0:
        return np.array([0.0, 25.0, 0.0])
    if y < -0.3:
        d = np.array([-x, 0.0, -z])
        return 2.0 * d / np.linalg.norm(d)
    return np.zeros(3)


def vortex_lambda(p):
    """The vortex scene's EXTERNAL_FORCE (app/vortex/application.cu:45)."""
    x, y, z = p
    if y < -0.5:
        d = np.array([z, 0.0, -x])
        return 2.0 * d / np.linalg.norm(d)
    return np.zeros(3)


def one_step_accel(prm, pos, dt=1e-3):
    p = np.array([[*pos, 1.0]], F32)
    s = O.Solver(prm, p, None, scenes.make_box((-10, -10, -10), (10, 10, 10)), precision=64, parallel=False)
    s.step(dt, 1)
    st = s.state()
    return st["vel"][0, :3] / float(F32(dt))


@pytest.mark.parametrize("pos", [(0.1, -0.2, 0.1), (0.5, -0.2, 0.1), (0.5, -0.5, -0.3), (-0.2, 0.5, 0.1),
                                 (0.3, -1.0, 0.4), (0.0, -0.1, 0.0)])
def test_fountain_force_matches_scene_lambda(pos):
    prm = abi.set_force_rules(baseline_params(gravity=(0.0, -9.81, 0.0)), FOUNTAIN)
    a = one_step_accel(prm, pos)
    p = np.array(pos, F32).astype(np.float64)
    expect = np.array([0.0, float(F32(-9.81)), 0.0]) + fountain_lambda(p)
    np.testing.assert_allclose(a, expect, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("pos", [(0.3, -0.6, 0.4), (-0.7, -0.9, 0.1), (0.3, -0.4, 0.4)])
def test_vortex_force_matches_scene_lambda(pos):
    prm = abi.set_force_rules(baseline_params(gravity=(0.0, 0.0, 0.0)), VORTEX)
    a = one_step_accel(prm, pos)
    expect = vortex_lambda(np.array(pos, F32).astype(np.float64))
    np.testing.assert_allclose(a, expect, rtol=1e-9, atol=1e-9)
    if pos[1] < -0.5:  # tangential: perpendicular to the radius in xz, magnitude 2
        p32 = np.array(pos, F32).astype(np.float64)
        assert abs(a[0] * p32[0] + a[2] * p32[2]) < 1e-9
        assert math.hypot(a[0], a[2]) == pytest.approx(2.0, rel=1e-9)


def test_force_on_the_axis_is_zero_and_rules_cancel_gravity():
    prm = abi.set_force_rules(baseline_params(gravity=(0.0, 0.0, 0.0)), VORTEX)
    np.testing.assert_array_equal(one_step_accel(prm, (0.0, -1.0, 0.0)), np.zeros(3))
    # surface_tension scene: a constant (0, 9.81, 0) cancels gravity
    # (app/surface_tension/application.cu)
    up = [abi.force_rule(abi.REGION_ALL, abi.FORCE_CONSTANT, value=(0.0, 9.81, 0.0))]
    prm = abi.set_force_rules(reference_params(), up)
    assert np.max(np.abs(one_step_accel(prm, (0.1, 0.2, 0.3)))) < 1e-9


# ------------------------------------------------------------- meshes
def test_cube_mesh_is_the_box_surface():
    tri = scenes.procedural_mesh(abi.MESH_CUBE, (0.5, 0.25, 1.0))
    assert tri.shape == (12, 3, 3)
    v = tri.reshape(-1, 3)
    assert set(np.unique(np.abs(v[:, 0]))) == {0.5} and set(np.unique(np.abs(v[:, 2]))) == {1.0}
    # every face plane carries two triangles of total area = face area
    area = 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1)
    assert area.sum() == pytest.approx(2 * (1.0 * 0.5 + 1.0 * 2.0 + 0.5 * 2.0), rel=1e-6)


def test_sphere_and_torus_meshes():
    res = 16
    c = np.array([0.3, -0.2, 0.5])
    sph = scenes.procedural_mesh(abi.MESH_SPHERE, (0.7,), res, scenes.translation(c))
    assert sph.shape[0] == 4 * res * res - 4 * res
    np.testing.assert_allclose(np.linalg.norm(sph.reshape(-1, 3) - c, axis=1), 0.7, rtol=1e-6)
    tor = scenes.procedural_mesh(abi.MESH_TORUS, (0.8, 0.2), res, scenes.translation(c))
    assert tor.shape[0] == 4 * res * res
    v = tor.reshape(-1, 3) - c
    rho = np.hypot(v[:, 0], v[:, 2])
    np.testing.assert_allclose((rho - 0.8) ** 2 + v[:, 1] ** 2, 0.04, rtol=1e-5)
    # closed, consistently sized: total area close to the analytic surfaces
    def area(t):
        return 0.5 * np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1).sum()
    assert area(sph) == pytest.approx(4 * math.pi * 0.49, rel=0.02)
    assert area(tor) == pytest.approx(4 * math.pi ** 2 * 0.8 * 0.2, rel=0.02)


# ------------------------------------------------------------- sampler
def test_sampler_on_a_square_gives_the_plane_lattice():
    sq = np.array([[[0, 0, 0], [1, 0, 0], [1, 1, 0]], [[0, 0, 0], [1, 1, 0], [0, 1, 0]]], F32)
    pts = scenes.sample_mesh(sq, 0.1)
    # grid origin (-0.05, -0.05, -0.05): voxel centres on the plane z = 0,
    # 11 x 11 of them covering [0, 1]^2 (the cubes beside them do not touch z = 0)
    assert pts.shape == (121, 4)
    np.testing.assert_allclose(pts[:, 2], 0.0, atol=1e-7)
    got = {(round(x / 0.1), round(y / 0.1)) for x, y in pts[:, :2].astype(np.float64)}
    assert got == {(i, j) for i in range(11) for j in range(11)}
    np.testing.assert_array_equal(pts[:, 3], 1.0)


def test_sampler_on_a_box_is_the_voxel_shell():
    v = 0.1
    tri = scenes.procedural_mesh(abi.MESH_CUBE, (0.5, 0.3, 0.2))
    pts = scenes.sample_mesh(tri, v)
    n = [11, 7, 5]  # 2*half/v + 1 voxel centres per axis, the outer ones on the faces
    assert pts.shape[0] == n[0] * n[1] * n[2] - (n[0] - 2) * (n[1] - 2) * (n[2] - 2)
    onface = np.isclose(np.abs(pts[:, 0]), 0.5, atol=1e-6) | np.isclose(np.abs(pts[:, 1]), 0.3, atol=1e-6) | \
        np.isclose(np.abs(pts[:, 2]), 0.2, atol=1e-6)
    assert onface.all()


def brute_overlap(c, hs, t):
    """Independent check by sampling: the triangle and the cube overlap if a
    dense set of triangle points falls inside the (slightly grown) cube."""
    u, w = np.meshgrid(np.linspace(0, 1, 60), np.linspace(0, 1, 60))
    m = (u + w) <= 1
    p = t[0] + u[m, None] * (t[1] - t[0]) + w[m, None] * (t[2] - t[0])
    return np.any(np.all(np.abs(p - c) <= hs * (1 + 1e-9), axis=1))


@pytest.mark.parametrize("kind,prm", [(abi.MESH_SPHERE, (0.3,)), (abi.MESH_TORUS, (0.3, 0.1))])
def test_sampler_on_curved_surfaces(kind, prm):
    v = 0.04
    tri = scenes.procedural_mesh(kind, prm, 12).astype(np.float64)
    pts = scenes.sample_mesh(tri, v).astype(np.float64)[:, :3]
    # every sample lies within half a voxel diagonal of the surface mesh
    if kind == abi.MESH_SPHERE:
        d = np.abs(np.linalg.norm(pts, axis=1) - prm[0])
    else:
        d = np.abs(np.hypot(np.hypot(pts[:, 0], pts[:, 2]) - prm[0], pts[:, 1]) - prm[1])
    sag = prm[-1] * (1 - math.cos(math.pi / 12))  # chord error of the mesh
    assert d.max() <= math.sqrt(3) / 2 * v + sag + 1e-6
    # every vertex's own voxel is sampled
    lo = tri.reshape(-1, 3).min(axis=0) - v / 2
    keys = {tuple(k) for k in np.round((pts - lo - v / 2) / v).astype(int)}
    own = np.floor((tri.reshape(-1, 3) - lo) / v).astype(int)
    assert all(tuple(k) in keys for k in own)
    # sampled voxels overlap a triangle, unsampled neighbours of samples do not
    # (checked on a subset against dense point sampling of the triangles)
    rng = np.random.default_rng(0)
    cents = lo + v / 2 + v * np.array(sorted(keys))
    for c in cents[rng.choice(len(cents), 20, replace=False)]:
        near = tri[np.min(np.linalg.norm(tri - c, axis=2), axis=1) < 3 * v]
        assert any(brute_overlap(c, v / 2, t) for t in near) or len(near) == 0
    probe = cents[rng.choice(len(cents), 40, replace=False)] + v * np.array([0, 0, 3])
    for c in probe:
        key = tuple(np.round((c - lo - v / 2) / v).astype(int))
        if key in keys:
            continue
        near = tri[np.min(np.linalg.norm(tri - c, axis=2), axis=1) < 3 * v]
        assert not any(brute_overlap(c, v / 2 * 0.98, t) for t in near)


def test_sampler_rejects_bad_input():
    with pytest.raises(ValueError):
        scenes.sample_mesh(np.zeros((1, 3, 3), F32), 0.0)
    assert scenes.sample_mesh(np.zeros((0, 3, 3), F32), 0.1).shape == (0, 4)


# ------------------------------------------------------------- clustered block
def test_clustered_block():
    r = 0.01
    fl = scenes.clustered_block(r, (0.0, 0.0, 0.0), (0.8, 0.6, 0.6), 0.1, 0.08, jitter_frac=0.0, seed=3)
    s = 2 * r
    main = np.all(np.abs((fl[:, :3] - r) / s - np.round((fl[:, :3] - r) / s)) < 1e-3, axis=1)
    frac = 1 - main.mean()
    assert 0.1 <= frac < 0.2
    assert fl[:, :3].min() >= r - 1e-6 and np.all(fl[:, :3].max(axis=0) <= np.array([0.8, 0.6, 0.6]) - r + 1e-6)
    # no two particles closer than the cluster spacing 1.8 r
    from scipy.spatial import cKDTree
    d, _ = cKDTree(fl[:, :3].astype(np.float64)).query(fl[:, :3].astype(np.float64), k=2)
    assert d[:, 1].min() >= 1.8 * r * (1 - 1e-5)
    # deterministic in the seed, different across seeds
    again = scenes.clustered_block(r, (0.0, 0.0, 0.0), (0.8, 0.6, 0.6), 0.1, 0.08, jitter_frac=0.0, seed=3)
    np.testing.assert_array_equal(fl, again)
    other = scenes.clustered_block(r, (0.0, 0.0, 0.0), (0.8, 0.6, 0.6), 0.1, 0.08, jitter_frac=0.0, seed=4)
    assert other.shape != fl.shape or not np.array_equal(other, fl)
    jit = scenes.clustered_block(r, (0.0, 0.0, 0.0), (0.8, 0.6, 0.6), 0.1, 0.08, jitter_frac=0.1, seed=3)
    assert jit.shape == fl.shape and np.max(np.abs(jit - fl)) <= 0.1 * r * 1.0001


def test_obstacle_collapse_8m_sizes():
    d = scenes.OBSTACLE_COLLAPSE_8M
    small = d.scaled(0.25)
    fl, bd, box = scenes.obstacle_collapse(small)
    assert 100_000 < fl.shape[0] < 160_000 and bd.shape[0] > 10_000
    # the obstacle lies inside the box and clear of the fluid block
    assert np.all(bd[:, :3].min(axis=0) >= np.array(small.box_min) - 1e-6)
    assert bd[:, 0].min() > fl[:, 0].max()
