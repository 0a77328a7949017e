"""Pins of the CPU oracle (no GPU).

Each check is chosen so that a plausible mistake fails it: a bit-loop vs
magic-mask Morton key; a brute-force stable order; O(N^2) neighbour sets
with exact FP32 rounding of the fused predicate; closed forms and
integrals of the kernels; the EOS clamps; an independent FP64 lattice sum
for the boundary mass; closed-form free fall; per-face collisions; the
CFL branches; momentum conservation; symmetry of a uniform lattice.
"""
import math
from fractions import Fraction

import numpy as np
import pytest
from scipy import integrate

from oracle import oracle as O
from tests.helpers import SMALL, baseline_params, reference_params, small_scene
from wcsph import scenes

F32 = np.float32


# ---------------------------------------------------------------- Morton
def magic_spread(v):
    """Published magic-mask bit spread (the method morton.cuh:7-15 uses)."""
    v &= 0x1FFFFF
    v = (v | (v << 32)) & 0x1F00000000FFFF
    v = (v | (v << 16)) & 0x1F0000FF0000FF
    v = (v | (v << 8)) & 0x100F00F00F00F00F
    v = (v | (v << 4)) & 0x10C30C30C30C30C3
    v = (v | (v << 2)) & 0x1249249249249249
    return v


def ref_key(cx, cy, cz):
    off = 1 << 20
    u = [(c + off) & 0xFFFFFFFF for c in (cx, cy, cz)]
    return magic_spread(u[0]) | (magic_spread(u[1]) << 1) | (magic_spread(u[2]) << 2)


def test_morton_key_matches_magic_masks():
    rng = np.random.default_rng(0)
    cs = rng.integers(-(1 << 20), (1 << 20), size=(2000, 3))
    for c in cs:
        assert O.morton_key(*map(int, c)) == ref_key(*map(int, c))


def test_morton_special_cases():
    assert O.morton_key(0, 0, 0) == 0b111 << 60  # bit 20 of each offset coordinate
    assert O.morton_key(-(1 << 20), -(1 << 20), -(1 << 20)) == 0
    assert O.morton_key(1, 0, 0) == (7 << 60) | 1
    assert O.morton_key(0, 1, 0) == (7 << 60) | 2
    assert O.morton_key(0, 0, 1) == (7 << 60) | 4
    # order is not translation invariant across 0: -1 sorts before 0 on x but
    # its key differs in bit 60 region
    assert O.morton_key(-1, 0, 0) < O.morton_key(0, 0, 0)
    assert O.morton_key(-1, 5, 5) < O.morton_key(0, 0, 0)


def test_cell_coord_is_floor_of_ieee_division():
    cs = F32(0.04)
    for p in [0.0, -0.0, 0.04, -0.04, 1e-9, -1e-9, 0.0399999, 0.12, -0.12]:
        p = F32(p)
        assert O.lib().orc_cell_coord_f32(p, cs) == int(np.floor(p / cs))
    rng = np.random.default_rng(1)
    ps = rng.uniform(-3, 3, 5000).astype(F32)
    got = np.array([O.lib().orc_cell_coord_f32(p, cs) for p in ps])
    np.testing.assert_array_equal(got, np.floor(ps / cs).astype(int))


# ---------------------------------------------------------------- sort
def brute_order(pos, cs):
    c = np.floor(pos[:, :3].astype(F32) / F32(cs)).astype(np.int64)
    keys = np.array([ref_key(*map(int, row)) for row in c], dtype=object)
    idx = list(range(len(pos)))
    return np.array(sorted(idx, key=lambda i: (keys[i], i)), dtype=np.uint32)


@pytest.mark.parametrize("kind", ["lattice", "straddle", "clustered"])
def test_sort_order_is_stable_key_order(kind):
    rng = np.random.default_rng(2)
    cs = F32(0.04)
    if kind == "lattice":
        fl, _, _ = small_scene()
        pos = fl[rng.permutation(len(fl))][:3000]
    elif kind == "straddle":  # coordinates on both sides of 0, like the reference snapshots
        pos = np.ones((3000, 4), F32)
        pos[:, :3] = rng.uniform(-0.5, 0.5, (3000, 3))
    else:  # many ties: few cells
        pos = np.ones((3000, 4), F32)
        pos[:, :3] = rng.uniform(-0.06, 0.06, (3000, 3))
    np.testing.assert_array_equal(O.sort_order(pos, cs), brute_order(pos, cs))
    keys = O.morton_keys(pos, cs)
    c = np.floor(pos[:, :3] / cs).astype(np.int64)
    assert all(int(keys[i]) == ref_key(*map(int, c[i])) for i in range(0, 3000, 97))


# ---------------------------------------------------------------- neighbours
def round_f32(fr):
    """Nearest float32 of an exact rational, ties to even."""
    f = F32(float(fr))
    cands = [f, np.nextafter(f, F32(np.inf)), np.nextafter(f, F32(-np.inf))]
    return min(cands, key=lambda c: (abs(Fraction(float(c)) - fr), int(np.array(c).view(np.uint32)) & 1))


def fused_d2_f32(rx, ry, rz):
    x2 = round_f32(Fraction(float(rx)) ** 2)
    t = round_f32(Fraction(float(ry)) ** 2 + Fraction(float(x2)))
    return round_f32(Fraction(float(rz)) ** 2 + Fraction(float(t)))


def brute_neighbors(q, src, h, exclude_self):
    h2 = F32(F32(h) * F32(h))
    out = []
    for i in range(len(q)):
        r = (q[i, :3][None, :] - src[:, :3]).astype(F32)
        d64 = np.sum(r.astype(np.float64) ** 2, axis=1)
        nb = set()
        for j in np.nonzero(d64 <= float(h2) * (1 + 1e-6))[0]:
            if exclude_self and j == i:
                continue
            if abs(d64[j] - float(h2)) > 1e-6 * float(h2):
                ok = d64[j] <= float(h2)
            else:
                ok = fused_d2_f32(*r[j]) <= h2
            if ok:
                nb.add(int(j))
        out.append(nb)
    return out


def cell_offset_rank(qi, sj, cs, R=1):
    ci = np.floor(qi[:3] / cs)
    cj = np.floor(sj[:3] / cs)
    d = (cj - ci).astype(int) + R
    assert np.all((d >= 0) & (d <= 2 * R))
    w = 2 * R + 1
    return d[0] * w * w + d[1] * w + d[2]


@pytest.mark.parametrize("kind", ["jittered", "exact_lattice", "straddle"])
def test_neighbors_vs_brute_force(kind):
    cs = F32(0.04)
    h = cs
    rng = np.random.default_rng(3)
    if kind == "jittered":
        fl, _, _ = small_scene(scenes.DamBreak(dims=(10, 10, 10), box_max=(0.64, 0.5, 0.5), seed=4))
        pos = fl
    elif kind == "exact_lattice":  # pairs at exactly 2s: the inclusive '<=' and FMA rounding decide
        fl, _, _ = small_scene(scenes.DamBreak(dims=(10, 10, 10), box_max=(0.64, 0.5, 0.5), jitter_frac=0.0))
        pos = fl
    else:
        pos = np.ones((1200, 4), F32)
        pos[:, :3] = rng.uniform(-0.15, 0.15, (1200, 3))
    pos = pos[O.sort_order(pos, cs)]
    counts, lst = O.neighbors(pos, pos, h, cs, True)
    ref = brute_neighbors(pos, pos, h, True)
    off = np.concatenate([[0], np.cumsum(counts.astype(np.int64))]).astype(np.int64)
    for i in range(len(pos)):
        got = lst[off[i]:off[i + 1]]
        assert set(got.tolist()) == ref[i]
        # reference enumeration order: cell offset (dx, dy, dz), then index
        rk = [(cell_offset_rank(pos[i], pos[j], cs), j) for j in got]
        assert rk == sorted(rk)


# The reference's CELL_SIZE_MULT values (step_tuner.cuh:88) and 1.5, 2: the
# fluid grid has cell size fp32(h) * fp32(mult) and walk range
# R = ceil(h / cs) (2 for 1/2 and 2/3, else 1).
MULTS = [(F32(2.0) / F32(3.0), 2), (F32(0.5), 2), (F32(1.5), 1), (F32(2.0), 1)]


@pytest.mark.parametrize("mult,R", MULTS, ids=["2/3", "1/2", "3/2", "2"])
@pytest.mark.parametrize("kind", ["exact_lattice", "straddle"])
def test_neighbors_vs_brute_force_finer_and_coarser_cells(kind, mult, R):
    """Same sets as the O(N^2) search at every cell size, and the order of a
    (2R+1)^3 walk: cell offset (dx, dy, dz) in [-R, R], then index."""
    h = F32(0.04)
    cs = F32(h * mult)
    assert math.ceil(float(h) / float(cs) - 1e-9) == R
    if kind == "exact_lattice":
        fl, _, _ = small_scene(scenes.DamBreak(dims=(9, 9, 9), box_max=(0.64, 0.5, 0.5), jitter_frac=0.0))
        pos = fl
    else:  # a ball around the origin: the cell coordinates change sign
        rng = np.random.default_rng(8)
        pos = np.ones((1500, 4), F32)
        pos[:, :3] = rng.uniform(-0.08, 0.08, (1500, 3))
        # a pair across the diagonal corner cell (2, 2, 2): near opposite
        # corners, sqrt(3) * 1.02 * cs apart
        pos[0, :3] = F32(cs) * F32(3.99)
        pos[1, :3] = F32(cs) * F32(5.01)
    pos = pos[O.sort_order(pos, cs)]
    counts, lst = O.neighbors(pos, pos, h, cs, True)
    ref = brute_neighbors(pos, pos, h, True)
    off = np.concatenate([[0], np.cumsum(counts.astype(np.int64))]).astype(np.int64)
    shell = corner = 0
    for i in range(len(pos)):
        got = lst[off[i]:off[i + 1]]
        assert set(got.tolist()) == ref[i]
        rk = [(cell_offset_rank(pos[i], pos[j], cs, R), j) for j in got]
        assert rk == sorted(rk)
        d = [np.abs(np.floor(pos[j, :3] / cs) - np.floor(pos[i, :3] / cs)) for j in got]
        shell += sum(1 for x in d if x.max() == R)
        corner += sum(1 for x in d if x.min() == R)
    if kind == "straddle":
        assert shell > 0  # the outermost cells of the walk hold neighbours
        # corner cells of a walk of range 2: nearest point sqrt(3) cs away,
        # within h for cell size h/2 only
        if R == 2:
            assert (corner > 0) == (mult == F32(0.5))


@pytest.mark.parametrize("precision", [32, 64])
def test_solver_cell_size_changes_only_the_order(precision):
    """A step on the finer fluid grids: the same neighbour set and boundary
    set per particle as at cell size h (the boundary grid keeps cell size
    h), and densities and positions equal up to the summation order."""
    fl, bd, box = small_scene(scenes.DamBreak(dims=(10, 10, 10), box_max=(0.64, 0.5, 0.5), seed=4))
    prm = baseline_params()
    out = []
    for mult in (1.0, float(F32(2.0) / F32(3.0)), 0.5):
        a = O.Solver(prm, fl, bd, box, precision=precision)
        a.set_cell_size_mult(mult)
        a.step(1e-4, 2)
        st = a.state()
        ids = st["ids"]
        sets = []
        for which in (0, 1):
            c, l = a.neighbors(which)
            o = np.concatenate([[0], np.cumsum(c.astype(np.int64))])
            bpos = a.boundary()
            key = (lambda j: int(ids[j])) if which == 0 else (lambda j: tuple(bpos[j, :3]))
            sets.append({int(ids[i]): frozenset(key(j) for j in l[o[i]:o[i + 1]]) for i in range(len(ids))})
        out.append((O.by_id(st), sets))
        a.close()
    base, base_sets = out[0]
    tol = 1e-5 if precision == 32 else 1e-12
    for st, sets in out[1:]:
        assert sets == base_sets
        assert np.max(np.abs(st["rho"] - base["rho"]) / base["rho"]) < tol
        assert np.max(np.abs(st["pos"] - base["pos"])) < tol * 0.1


# ---------------------------------------------------------------- kernels
H = 0.04


def test_cubic_spline_closed_form_and_normalisation():
    assert O.lib().orc_cubic_w(0.0, H) == pytest.approx(8 / (math.pi * H ** 3), rel=1e-15)
    assert O.lib().orc_cubic_w(1.0, H) == 0.0
    assert O.lib().orc_cubic_w(1.2, H) == 0.0
    # continuity at q = 1/2: 6/8 - 6/4 + 1 = 2 (1/2)^3
    lo = O.lib().orc_cubic_w(0.5, H)
    hi = O.lib().orc_cubic_w(0.5 + 1e-12, H)
    assert lo == pytest.approx(hi, rel=1e-9)
    # 4 pi int_0^h r^2 W(r/h) dr = 1
    val, _ = integrate.quad(lambda r: 4 * math.pi * r * r * O.lib().orc_cubic_w(r / H, H), 0, H,
                            points=[H / 2], epsabs=1e-13)
    assert val == pytest.approx(1.0, abs=1e-9)


@pytest.mark.parametrize("q", [0.1, 0.3, 0.49, 0.51, 0.7, 0.95])
def test_cubic_gradient_is_derivative_of_w(q):
    r = q * H
    e = 1e-7 * H
    dwdr = (O.lib().orc_cubic_w((r + e) / H, H) - O.lib().orc_cubic_w((r - e) / H, H)) / (2 * e)
    # grad W = coef * r_vec, so coef = (dW/dr) / |r|
    assert O.lib().orc_cubic_grad(q, H) == pytest.approx(dwdr / r, rel=1e-6)


@pytest.mark.parametrize("q", [0.1, 0.5, 0.9])
def test_spiky_reference_gradient(q):
    r = q * H
    # gradient of the reference spiky W = 15/(pi h^9)(1-q)^3 (kernel.cuh:34-41) ...
    spiky = lambda rr: 15 / (math.pi * H ** 9) * (1 - rr / H) ** 3  # noqa: E731
    e = 1e-7 * H
    dwdr = (spiky(r + e) - spiky(r - e)) / (2 * e)
    assert O.lib().orc_spiky_ref_grad(q, H) == pytest.approx(dwdr / r, rel=1e-6)
    # ... which is the textbook spiky gradient -45/(pi h^6)(h-r)^2/r times h^-6
    textbook = -45 / (math.pi * H ** 6) * (H - r) ** 2 / r
    assert O.lib().orc_spiky_ref_grad(q, H) == pytest.approx(textbook * H ** -6, rel=1e-12)
    assert O.lib().orc_spiky_ref_grad(0.0, H) == 0.0 and O.lib().orc_spiky_ref_grad(1.01, H) == 0.0


@pytest.mark.parametrize("q", [0.2, 0.5, 0.8])
def test_cohesion_is_akinci_form(q):
    r = q * H
    c = 32 / (math.pi * H ** 9)
    if 2 * r > H:
        ref = c * (H - r) ** 3 * r ** 3
    else:
        ref = c * (2 * (H - r) ** 3 * r ** 3 - H ** 6 / 64)
    assert O.lib().orc_cohesion(q, H) == pytest.approx(ref, rel=1e-12)


def test_tait_clamps():
    t = O.lib().orc_tait
    assert t(900.0, 1000.0, 1000.0, 7.0, 1.5) == 0.0  # below rest density
    assert t(1000.0, 1000.0, 1000.0, 7.0, 1.5) == 0.0
    assert t(1100.0, 1000.0, 1000.0, 7.0, 1.5) == pytest.approx(1000.0 * (1.1 ** 7 - 1), rel=1e-14)
    assert t(2000.0, 1000.0, 1000.0, 7.0, 1.5) == pytest.approx(1000.0 * (1.5 ** 7 - 1), rel=1e-14)


# ---------------------------------------------------------------- solver
def big_box(lo=-10.0, hi=10.0):
    return scenes.make_box((lo, lo, lo), (hi, hi, hi))


def one_particle(prm, pos, vel, box, steps, hint):
    p = np.array([[*pos, 1.0]], F32)
    v = np.array([[*vel, 0.0]], F32)
    s = O.Solver(prm, p, None, box, fluid_vel=v, precision=64, parallel=False)
    s.step(hint, steps)
    return s.state(), s.time()


def test_free_fall_closed_form():
    prm = baseline_params()
    dt, n = 1e-3, 50
    st, (t, _) = one_particle(prm, (0.0, 1.0, 0.0), (0.0, 0.0, 0.0), big_box(), n, dt)
    dt = float(F32(dt))
    g = float(F32(-9.81))
    assert st["vel"][0, 1] == pytest.approx(n * g * dt, rel=1e-12)
    assert st["pos"][0, 1] == pytest.approx(1.0 + g * dt * dt * n * (n + 1) / 2, rel=1e-12)
    assert st["rho"][0] == pytest.approx(float(baseline_params().rest_density) * 0.8 * 0.02 ** 3
                                         * 8 / (math.pi * 0.04 ** 3), rel=1e-6)
    assert t == pytest.approx(n * dt, rel=1e-12)


@pytest.mark.parametrize("axis,sign", [(0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1)])
def test_box_collision_each_face(axis, sign):
    prm = baseline_params(gravity=(0.0, 0.0, 0.0))
    r = float(F32(0.01))
    box = scenes.make_box((0, 0, 0), (1, 1, 1))
    pos = [0.5, 0.5, 0.5]
    vel = [0.0, 0.0, 0.0]
    pos[axis] = 0.5 + sign * 0.49
    vel[axis] = sign * 10.0
    st, _ = one_particle(prm, pos, vel, box, 1, 1e-3)
    expect = 0.0 + r if sign < 0 else 1.0 - r
    assert st["pos"][0, axis] == pytest.approx(expect, abs=1e-12)
    assert st["vel"][0, axis] == pytest.approx(-0.5 * sign * 10.0, rel=1e-7)


def test_box_collision_rotated_frame():
    prm = baseline_params(gravity=(0.0, 0.0, 0.0))
    ang = 0.3
    m = np.eye(4)
    m[:2, :2] = [[math.cos(ang), -math.sin(ang)], [math.sin(ang), math.cos(ang)]]
    m[:3, 3] = [0.2, -0.1, 0.0]
    box = scenes.make_box((-1, -1, -1), (1, 1, 1), model=m)
    # a particle heading out through the box's +x face (in the box frame)
    inside = np.array([0.99, 0.0, 0.0, 1.0])
    world = m @ inside
    vdir = m[:3, :3] @ np.array([5.0, 0.0, 0.0])
    st, _ = one_particle(prm, world[:3], vdir, box, 1, 1e-3)
    back = np.linalg.inv(m) @ np.array([*st["pos"][0, :3], 1.0])
    assert back[0] == pytest.approx(1.0 - float(F32(0.01)), abs=1e-6)
    vb = np.linalg.inv(m)[:3, :3] @ st["vel"][0, :3]
    assert vb[0] == pytest.approx(-2.5, rel=1e-5)


def test_cfl_branches():
    prm = baseline_params(adaptive_dt=1)
    box = big_box()
    # at rest: dt_max
    _, (t, last) = one_particle(prm, (0, 0, 0), (0, 0, 0), box, 1, 1e-2)
    assert last == pytest.approx(float(F32(5e-4)), rel=1e-7)
    # fast: lambda * s / vmax (sph.cu:124-129)
    _, (t, last) = one_particle(prm, (0, 0, 0), (30.0, 0, 0), box, 1, 1e-2)
    assert last == pytest.approx(float(F32(0.4)) * float(F32(0.02)) / 30.0, rel=1e-7)
    # hint below dt_min: clamped to dt_min
    _, (t, last) = one_particle(prm, (0, 0, 0), (1.0, 0, 0), box, 1, 1e-7)
    assert last == pytest.approx(float(F32(1e-5)), rel=1e-7)


def test_pairwise_forces_conserve_momentum():
    prm = baseline_params(gravity=(0.0, 0.0, 0.0))
    fl, _, _ = small_scene(scenes.DamBreak(dims=(8, 8, 8), box_max=(0.64, 0.5, 0.5), seed=9))
    rng = np.random.default_rng(4)
    pos = fl.copy()
    pos[:, :3] = (pos[:, :3] - 0.01) * F32(0.85) + F32(3.0)  # compressed, far from the walls
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.normal(0, 0.5, (len(pos), 3)).astype(F32)
    s = O.Solver(prm, pos, None, big_box(0.0, 6.0), fluid_vel=vel, precision=64)
    p0 = vel[:, :3].astype(np.float64).sum(axis=0)
    s.step(1e-4)
    st = s.state()
    assert st["p"].max() > 0  # pressure active
    p1 = st["vel"][:, :3].sum(axis=0)
    scale = np.abs(st["vel"][:, :3] - 0).sum()
    assert np.max(np.abs(p1 - p0)) <= 1e-12 * scale


def test_uniform_lattice_interior_is_at_rest():
    prm = baseline_params(gravity=(0.0, 0.0, 0.0))
    n = 14
    s = F32(1.0 / 64.0)  # compressed lattice, exact in FP32: p > 0 everywhere inside
    g = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = np.ones((len(g), 4), F32)
    pos[:, :3] = F32(2.0) + g.astype(F32) * s
    sv = O.Solver(prm, pos, None, big_box(0.0, 6.0), precision=64)
    sv.step(1e-4)
    st = O.by_id(sv.state())
    # density is uniform from the 3rd layer in (support 2.56 spacings), so
    # forces are uniform two layers further in
    inner = np.all((g >= 5) & (g <= n - 6), axis=1)
    assert st["p"][inner].min() > 0
    # interior accelerations cancel by symmetry (up to FP64 rounding of the lattice)
    assert np.max(np.abs(st["vel"][inner, :3])) < 1e-9
    assert np.max(np.abs(st["vel"][~inner, :3])) > 1e-4  # the surface layer does move


def boundary_psi_reference(points, centre, h):
    """FP64 Psi of one wall particle: rho0 / sum_j W(|x_c - x_j| / h), all j."""
    r = np.linalg.norm(points - points[centre], axis=1) / h
    w = np.array([O.lib().orc_cubic_w(q, h) for q in r])
    return 1000.0 / w.sum()


def test_boundary_mass_flat_wall():
    prm = baseline_params()
    s = 0.02
    ax = np.arange(-9, 10) * s
    lz = np.arange(3) * s
    P = np.stack(np.meshgrid(ax, ax, lz, indexing="ij"), -1).reshape(-1, 3)
    bpos = np.ones((len(P), 4), F32)
    bpos[:, :3] = P + 2.0  # shift so all coordinates are positive
    fl = np.array([[2.0, 2.0, 2.5, 1.0]], F32)
    sv = O.Solver(prm, fl, bpos, big_box(0.0, 6.0), precision=64)
    b = sv.boundary()
    centre = int(np.argmin(np.linalg.norm(b[:, :3] - (np.array([2.0, 2.0, 2.0 + s])), axis=1)))
    h = float(F32(4.0) * F32(0.01))
    ref = boundary_psi_reference(b[:, :3], centre, h)
    assert b[centre, 3] == pytest.approx(ref, rel=1e-12)
    # a flat 3-layer wall at spacing 2r: Psi ~ rho0 * (2r)^3 within 1%
    assert b[centre, 3] == pytest.approx(1000.0 * s ** 3, rel=0.01)


def test_fp32_and_fp64_oracles_agree_one_step():
    fl, bd, box = small_scene()
    prm = baseline_params()
    a = O.Solver(prm, fl, bd, box, precision=64)
    b = O.Solver(prm, fl, bd, box, precision=32)
    a.step(1e-4)
    b.step(1e-4)
    sa, sb = O.by_id(a.state()), O.by_id(b.state())
    assert np.max(np.abs(sa["rho"] - sb["rho"]) / sa["rho"]) < 1e-6
    np.testing.assert_array_equal(a.permutation(), b.permutation())


def test_reference_preset_runs_surface_tension():
    fl, bd, box = small_scene(scenes.DamBreak(dims=(8, 8, 8), box_max=(1.2, 1.0, 1.0), r=0.02, seed=2))
    prm = reference_params()
    s = O.Solver(prm, fl, bd, box, precision=64)
    s.step(0.01, 3)
    t, last = s.time()
    assert last == pytest.approx(float(F32(5e-4)), rel=1e-6)  # at rest first: dt_max
    st = s.state()
    assert np.all(np.isfinite(st["pos"]))


# ---------------------------------------------------------------- scenes
def test_dam_break_counts_and_jitter():
    d = SMALL
    fl, bd = O.make_dam_break(d.r, d.dims, d.lattice_min(), F32(d.jitter_frac) * F32(d.r), d.seed, d.box_min,
                              d.box_max, d.layers)
    n_in = [round((d.box_max[a] - d.box_min[a]) / (2 * d.r)) for a in range(3)]
    assert len(fl) == np.prod(d.dims)
    assert len(bd) == np.prod([k + 2 * d.layers for k in n_in]) - np.prod(n_in)
    lattice = np.stack(np.meshgrid(*[np.arange(k) for k in d.dims], indexing="ij"), -1).reshape(-1, 3)
    site = np.array(d.lattice_min(), F32) + lattice.astype(F32) * F32(2 * d.r)
    jit = fl[:, :3] - site
    assert np.max(np.abs(jit)) <= 0.1 * d.r * (1 + 1e-5)
    assert np.std(jit) > 0.05 * d.r  # uniform on [-0.1r, 0.1r): std = 0.0577 r
    # walls lie outside the box, on the fluid lattice continuation
    outside = np.any((bd[:, :3] < np.array(d.box_min)) | (bd[:, :3] > np.array(d.box_max)), axis=1)
    assert outside.all()
