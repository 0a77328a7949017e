"""Pins of the oracle's density sum and force pass (no GPU).

Each test drives the oracle's own density and force passes (Solver.eval:
sort, lists, density, forces, no integration) on a configuration whose
answer is known independently:
- pair forces: two or three particles with rho and p set by hand, against
  the published expression of each term, with the kernel gradient taken
  from the textbook kernel's derivative written out here;
- density: brute-force sums over every particle pair (no grid) and a closed
  lattice sum over integer offsets.
Every docstring names the mistakes the test catches.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import baseline_params, reference_params
from wcsph import scenes

F32 = np.float32


def derived(prm):
    """Support radius h and particle mass m formed in float32 as the solver
    forms them (s = 2r, h = 2s, m = rho0 * 0.8 s^3), widened to FP64."""
    r = F32(prm.particle_radius)
    s = F32(2.0) * r
    h = F32(2.0) * s
    m = F32(prm.rest_density) * (s * s * s * F32(0.8))
    return float(h), float(m)


def cubic_w(r, h):
    """Textbook 3-D cubic spline on support [0, h] (Monaghan & Lattanzio 1985
    in the q = r/h form): sigma = 8 / (pi h^3)."""
    q = np.asarray(r, np.float64) / h
    sig = 8.0 / (math.pi * h ** 3)
    return np.where(q <= 0.5, sig * (6 * q ** 3 - 6 * q ** 2 + 1), np.where(q <= 1.0, 2 * sig * (1 - q) ** 3, 0.0))


def cubic_dwdr(r, h):
    """dW/dr of the cubic spline above, differentiated by hand."""
    q = r / h
    sig = 8.0 / (math.pi * h ** 3)
    if q <= 0.5:
        return sig / h * (18 * q * q - 12 * q)
    if q <= 1.0:
        return -6.0 * sig / h * (1 - q) ** 2
    return 0.0


def spiky_dwdr_reference(r, h):
    """dW/dr of the textbook spiky kernel 15/(pi h^6)(h - r)^3 (Mueller 2003)
    times h^-6, the normalisation the reference's spiky gradient carries
    (kernel.cuh:43-51)."""
    return -45.0 / (math.pi * h ** 6) * (h - r) ** 2 * h ** -6


def grad_i(xi, xj, h, kind):
    """Gradient of W(|x_i - x_j|) with respect to x_i."""
    d = np.asarray(xi, np.float64) - np.asarray(xj, np.float64)
    r = float(np.linalg.norm(d))
    dw = cubic_dwdr(r, h) if kind == "cubic" else spiky_dwdr_reference(r, h)
    return dw * d / r


def akinci_cohesion(r, h):
    """Cohesion spline C(r) of Akinci et al. 2013, eq. 2."""
    c = 32.0 / (math.pi * h ** 9)
    if 2 * r > h and r <= h:
        return c * (h - r) ** 3 * r ** 3
    if r > 0 and 2 * r <= h:
        return c * (2 * (h - r) ** 3 * r ** 3 - h ** 6 / 64.0)
    return 0.0


def pair(prm, d_over_h, vi=(0, 0, 0), vj=(0, 0, 0), boundary=False):
    """Fluid particle i at (3, 3, 3) and a partner at distance d along a
    direction off every axis: a second fluid particle j, or (boundary=True)
    one boundary particle b.  Returns the solver and the float32 positions."""
    h, _ = derived(prm)
    u = np.array([1.0, 2.0, 2.0]) / 3.0
    xi = np.array([3.0, 3.0, 3.0, 1.0], F32)
    xj = np.array([*(np.array([3.0, 3.0, 3.0]) + d_over_h * h * u), 1.0], F32)
    box = scenes.make_box((0.0, 0.0, 0.0), (6.0, 6.0, 6.0))
    if boundary:
        fl = xi[None, :]
        vel = np.array([[*vi, 0.0]], F32)
        return O.Solver(prm, fl, xj[None, :], box, fluid_vel=vel, precision=64, parallel=False), xi, xj
    fl = np.stack([xi, xj])
    vel = np.array([[*vi, 0.0], [*vj, 0.0]], F32)
    return O.Solver(prm, fl, None, box, fluid_vel=vel, precision=64, parallel=False), xi, xj


# ------------------------------------------------------------------ pressure
@pytest.mark.parametrize("preset,kind", [("benchmark", "cubic"), ("reference", "spiky")])
@pytest.mark.parametrize("q", [0.3, 0.55, 0.85])
def test_pressure_pair_closed_form(preset, kind, q):
    """a^p_i = -m (p_i/rho_i^2 + p_j/rho_j^2) grad_i W_ij for two fluid
    particles with rho and p set by hand (Becker & Teschner 2007, eq. 6;
    compute_pressure_accel_n_normal.cu:98-124).  Unequal rho and p on the two
    sides catch a p_i/rho_i^2 term doubled or a p_j/rho_j^2 term dropped; the
    exact value catches a missing or squared mass factor; the repulsion check
    catches a flipped sign; q on both branches of each gradient catches a
    wrong polynomial piece; the spiky case catches a wrong h^-9
    normalisation."""
    prm = baseline_params(gravity=(0.0, 0.0, 0.0)) if preset == "benchmark" else \
        reference_params(gravity=(0.0, 0.0, 0.0), surface_tension=0.0)
    h, m = derived(prm)
    sv, xi, xj = pair(prm, q)
    rho = np.array([1100.0, 1300.0])
    p = np.array([50.0, 200.0])
    out = sv.eval(rho=rho, p=p)
    g = grad_i(xi[:3], xj[:3], h, kind)
    expect = -m * (p[0] / rho[0] ** 2 + p[1] / rho[1] ** 2) * g
    np.testing.assert_allclose(out["ap"][0, :3], expect, rtol=1e-12, atol=0)
    np.testing.assert_allclose(out["ap"][1, :3], -expect, rtol=1e-12, atol=0)  # grad_j W = -grad_i W
    # pressure pushes the particles apart
    assert np.dot(out["ap"][0, :3], xi[:3].astype(np.float64) - xj[:3]) > 0


def test_pressure_scales_with_particle_mass():
    """The pressure sum carries m once: doubling rho0 (m = rho0 0.8 s^3)
    with rho and p held fixed doubles a^p.  Catches m dropped or squared
    independently of the closed form."""
    outs = []
    for rho0 in (1000.0, 2000.0):
        prm = baseline_params(gravity=(0.0, 0.0, 0.0), rest_density=rho0)
        sv, _, _ = pair(prm, 0.6)
        outs.append(sv.eval(rho=np.array([1500.0, 1500.0]), p=np.array([100.0, 100.0]))["ap"][0, :3])
    np.testing.assert_allclose(outs[1], 2.0 * outs[0], rtol=1e-12)


# ---------------------------------------------------------------- boundary
@pytest.mark.parametrize("preset,kind", [("benchmark", "cubic"), ("reference", "spiky")])
@pytest.mark.parametrize("q", [0.4, 0.7])
def test_akinci_boundary_term_closed_form(preset, kind, q):
    """One fluid particle and one boundary particle:
    a^p_i = -Psi_b (p_i/rho_i^2 + p_i/rho0^2) grad_i W_ib (Akinci et al.
    2012, eq. 10; compute_pressure_accel_n_normal.cu:337-365), with
    Psi_b = rho0 / W(0) for a lone boundary particle (eq. 5).  Catches the
    boundary using p_b or rho_b instead of the fluid particle's own p and
    rho0, a Psi factor dropped, the mass m used instead of Psi, or a flipped
    sign (the wall must push the particle away)."""
    prm = baseline_params(gravity=(0.0, 0.0, 0.0)) if preset == "benchmark" else \
        reference_params(gravity=(0.0, 0.0, 0.0), surface_tension=0.0)
    h, m = derived(prm)
    sv, xi, xb = pair(prm, q, boundary=True)
    rho0 = float(F32(prm.rest_density))
    psi = sv.boundary()[0, 3]
    assert psi == pytest.approx(rho0 / (8.0 / (math.pi * h ** 3)), rel=1e-14)
    rho_i, p_i = 1180.0, 320.0
    out = sv.eval(rho=np.array([rho_i]), p=np.array([p_i]))
    expect = -psi * (p_i / rho_i ** 2 + p_i / rho0 ** 2) * grad_i(xi[:3], xb[:3], h, kind)
    np.testing.assert_allclose(out["ap"][0, :3], expect, rtol=1e-12, atol=0)
    assert np.dot(out["ap"][0, :3], xi[:3].astype(np.float64) - xb[:3]) > 0


# --------------------------------------------------------------- viscosity
def test_monaghan_viscosity_closing_and_separating_pair():
    """Artificial viscosity (Monaghan 1992, eq. 4.2 with beta = 0):
    a_i = -m Pi_ij grad_i W_ij, Pi_ij = -nu (v_ij . x_ij) / (r^2 + 0.01 h^2)
    for v_ij . x_ij < 0 and 0 otherwise, nu = 2 alpha h c_s / (rho_i + rho_j),
    c_s = sqrt(k gamma / rho0).  Unequal densities catch rho_i + rho_j
    replaced by 2 rho_i; the exact value catches a missing factor 2, h or
    c_s; the separating pair catches the switch dropped; the dissipation
    check catches a flipped sign."""
    alpha = 0.05
    prm = baseline_params(gravity=(0.0, 0.0, 0.0), viscosity=alpha)
    h, m = derived(prm)
    u = np.array([1.0, 2.0, 2.0]) / 3.0
    w = np.array([2.0, -1.0, 0.0]) / math.sqrt(5.0)  # perpendicular to u
    vi = 0.7 * u + 0.3 * w  # i moves towards j (j sits at +u from i)
    vj = -0.4 * u
    rho = np.array([1050.0, 1250.0])
    p = np.zeros(2)
    sv, xi, xj = pair(prm, 0.45, vi=vi, vj=vj)
    out = sv.eval(rho=rho, p=p)
    vi32 = np.array(vi, F32).astype(np.float64)
    vj32 = np.array(vj, F32).astype(np.float64)
    xij = xi[:3].astype(np.float64) - xj[:3]
    vij = vi32 - vj32
    vr = float(np.dot(vij, xij))
    assert vr < 0
    cs = math.sqrt(float(F32(prm.stiffness)) * float(F32(prm.exponent)) / float(F32(prm.rest_density)))
    nu = 2.0 * float(F32(alpha)) * h * cs / (rho[0] + rho[1])
    pi_ij = -nu * vr / (float(np.dot(xij, xij)) + 0.01 * h * h)
    expect = -m * pi_ij * grad_i(xi[:3], xj[:3], h, "cubic")
    np.testing.assert_allclose(out["anp"][0, :3], expect, rtol=1e-12, atol=0)
    np.testing.assert_allclose(out["anp"][1, :3], -expect, rtol=1e-12, atol=0)
    # viscosity removes relative kinetic energy
    assert np.dot(out["anp"][0, :3] - out["anp"][1, :3], vij) < 0
    # the same pair moving apart feels no artificial viscosity
    sv2, _, _ = pair(prm, 0.45, vi=-vi, vj=-vj)
    out2 = sv2.eval(rho=rho, p=p)
    assert np.all(out2["anp"][:, :3] == 0.0)


def test_laplacian_viscosity_closed_form():
    """Laplacian viscosity of the reference preset:
    a_i = 2(d+2) nu sum_j (m/rho_j) (v_ij . x_ij) / (r^2 + 0.01 h^2) grad_i W_ij
    with d = 3, the cubic-spline gradient (Monaghan 2005, eq. 5.17;
    compute_non_pressure_accel.cu:138-144).  Unequal densities catch rho_j
    replaced by rho_i; the exact value catches the factor 10 = 2(d+2), nu or
    m dropped; the dissipation check catches a flipped sign."""
    prm = reference_params(gravity=(0.0, 0.0, 0.0), surface_tension=0.0)
    h, m = derived(prm)
    nu = float(F32(prm.viscosity))
    u = np.array([1.0, 2.0, 2.0]) / 3.0
    vi, vj = 0.8 * u, np.array([0.1, -0.2, 0.3])
    rho = np.array([900.0, 1400.0])
    sv, xi, xj = pair(prm, 0.62, vi=vi, vj=vj)
    out = sv.eval(rho=rho, p=np.zeros(2))
    xij = xi[:3].astype(np.float64) - xj[:3]
    vij = np.array(vi, F32).astype(np.float64) - np.array(vj, F32).astype(np.float64)
    e = float(np.dot(xij, xij)) + 0.01 * h * h
    g = grad_i(xi[:3], xj[:3], h, "cubic")
    expect_i = 10.0 * nu * (m / rho[1]) * float(np.dot(vij, xij)) / e * g
    expect_j = 10.0 * nu * (m / rho[0]) * float(np.dot(-vij, -xij)) / e * (-g)
    np.testing.assert_allclose(out["anp"][0, :3], expect_i, rtol=1e-12, atol=0)
    np.testing.assert_allclose(out["anp"][1, :3], expect_j, rtol=1e-12, atol=0)
    # m_i a_i + m_j a_j: with rho_i != rho_j the pair terms differ, but both
    # oppose the relative motion
    assert np.dot(out["anp"][0, :3] - out["anp"][1, :3], vij) < 0


# --------------------------------------------------------- surface tension
@pytest.mark.parametrize("q", [0.35, 0.75])
def test_surface_tension_pair_akinci_closed_form(q):
    """Surface tension of Akinci et al. 2013 for two fluid particles:
    n_i = h sum_j (m/rho_j) grad_i W_ij (eq. 6, cubic gradient),
    a_i = -gamma sum_j 2 rho0 / (rho_i + rho_j) (m C(r) x_ij/|x_ij| + n_i - n_j)
    (eqs. 1, 4, 7, divided by m_i; compute_non_pressure_accel.cu:113-145).
    Catches the normal's h or m factor, n_i - n_j reversed, the symmetric
    K_ij factor replaced by rho0/rho_i, the cohesion direction flipped, or
    the 2 rho0 factor dropped; q on both cohesion branches catches a wrong
    spline piece."""
    gamma = 0.13
    prm = reference_params(gravity=(0.0, 0.0, 0.0), viscosity=0.0, surface_tension=gamma)
    h, m = derived(prm)
    rho0 = float(F32(prm.rest_density))
    rho = np.array([950.0, 1120.0])
    sv, xi, xj = pair(prm, q)
    out = sv.eval(rho=rho, p=np.zeros(2))
    g = grad_i(xi[:3], xj[:3], h, "cubic")
    n_i = h * (m / rho[1]) * g
    n_j = h * (m / rho[0]) * (-g)
    np.testing.assert_allclose(out["nrm"][0, :3], n_i, rtol=1e-12, atol=0)
    np.testing.assert_allclose(out["nrm"][1, :3], n_j, rtol=1e-12, atol=0)
    xij = xi[:3].astype(np.float64) - xj[:3]
    r = float(np.linalg.norm(xij))
    k_ij = 2.0 * rho0 / (rho[0] + rho[1])
    gm = float(F32(gamma))
    expect_i = -gm * k_ij * (m * akinci_cohesion(r, h) * xij / r + n_i - n_j)
    expect_j = -gm * k_ij * (m * akinci_cohesion(r, h) * (-xij) / r + n_j - n_i)
    np.testing.assert_allclose(out["anp"][0, :3], expect_i, rtol=1e-11, atol=0)
    np.testing.assert_allclose(out["anp"][1, :3], expect_j, rtol=1e-11, atol=0)


# ----------------------------------------------------------------- density
def test_interior_lattice_density_is_the_lattice_sum():
    """A particle deep inside an exact lattice (spacing a = 1/64, exact in
    float32) has rho = m sum_{o in Z^3, |o| a <= h} W(|o| a), the sum taken
    over integer offsets here (no grid, no neighbour list).  Catches the
    self term W(0) dropped, the mass factor missing, a neighbour cell
    skipped (each of the 27 cells holds some of the offsets), or a density
    pass that stops short of the support."""
    prm = baseline_params(gravity=(0.0, 0.0, 0.0))
    h, m = derived(prm)
    n = 9
    a = 1.0 / 64.0
    g = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = np.ones((len(g), 4), F32)
    pos[:, :3] = F32(2.0) + g.astype(F32) * F32(a)
    sv = O.Solver(prm, pos, None, scenes.make_box((0.0, 0.0, 0.0), (6.0, 6.0, 6.0)), precision=64,
                  parallel=False)
    out = sv.eval()
    k = int(math.ceil(h / a))
    o = np.stack(np.meshgrid(*[np.arange(-k, k + 1)] * 3, indexing="ij"), -1).reshape(-1, 3)
    rr = np.linalg.norm(o * a, axis=1)
    lattice_sum = m * float(np.sum(cubic_w(rr[rr <= h], h)))
    centre = int(np.flatnonzero(np.all(g == n // 2, axis=1))[0])
    assert out["rho"][centre] == pytest.approx(lattice_sum, rel=1e-13)
    # more than W(0) alone, and the lattice is dense enough to be compressed
    assert lattice_sum > 1.2 * float(F32(prm.rest_density))


def test_density_with_wall_matches_brute_force():
    """rho_i = m (W(0) + sum_{j != i, |x_ij| <= h} W_ij) + sum_{b, |x_ib| <= h} Psi_b W_ib
    (compute_rho_p.cu:84-101, 291-317; Akinci et al. 2012, eq. 4) by brute
    force over every pair of a jittered block above a 3-layer wall, with
    Psi_b from the oracle's boundary set (pinned by the flat-wall test of
    the boundary mass).  Catches a missing Psi factor, the wall weighted by
    m instead of Psi, boundary neighbours counted twice or missed, the mass
    applied to the wall sum, and a fluid neighbour missed by the grid."""
    prm = baseline_params(gravity=(0.0, 0.0, 0.0))
    h, m = derived(prm)
    s = 0.02
    rng = np.random.default_rng(12)
    g = np.stack(np.meshgrid(np.arange(6), np.arange(6), np.arange(4), indexing="ij"), -1).reshape(-1, 3)
    fl = np.ones((len(g), 4), F32)
    fl[:, :3] = (np.array([2.0, 2.0, 2.0]) + g * s * 0.9 + rng.uniform(-0.002, 0.002, (len(g), 3))).astype(F32)
    wx = np.arange(-3, 9) * s
    wz = 2.0 - s * np.arange(1, 4)
    W3 = np.stack(np.meshgrid(wx, wx, wz, indexing="ij"), -1).reshape(-1, 3)
    bd = np.ones((len(W3), 4), F32)
    bd[:, :3] = (W3 + np.array([2.0, 2.0, 0.0])).astype(F32)
    sv = O.Solver(prm, fl, bd, scenes.make_box((0.0, 0.0, 0.0), (6.0, 6.0, 6.0)), precision=64, parallel=False)
    out = sv.eval()
    B = sv.boundary()
    X = fl[:, :3].astype(np.float64)
    d_ff = np.linalg.norm(X[:, None, :] - X[None, :, :], axis=2)
    w_ff = np.where(d_ff <= h, cubic_w(d_ff, h), 0.0)
    d_fb = np.linalg.norm(X[:, None, :] - B[None, :, :3], axis=2)
    w_fb = np.where(d_fb <= h, cubic_w(d_fb, h), 0.0)
    expect = m * w_ff.sum(axis=1) + w_fb @ B[:, 3]  # w_ff's diagonal is W(0)
    assert (w_fb > 0).any(axis=1).sum() >= 36  # the bottom layer sees the wall
    np.testing.assert_allclose(out["rho"], expect, rtol=1e-12)
