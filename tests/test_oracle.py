"""Pins the CPU oracle against closed forms and invariants.

The reference ships no tests and its core cannot be built here (it includes
boost through io/configuration.h), so the oracle is pinned by what the
mathematics fixes: streaming is a permutation, BGK conserves mass and
momentum and relaxes the non-equilibrium part by (1 - 1/tau), the textbook
equilibrium is a fixed point, half-way bounce-back conserves mass, Couette
flow is linear between a no-slip and a moving wall, uniform equilibrium flow
passes inflow/outflow/pressure/free-slip boundaries unchanged, and the
first-step and last-wins semantics of the reference's object model. The
extensions (equilibrium initialisation, Brinkman porosity with a body force)
are pinned by their closed forms: a uniform porous box follows the scalar
recurrence u' = u (1 - omega alpha) + omega (1 - alpha) tau g exactly.
"""
import numpy as np
import pytest

from oracle import ora
from lbm_amd import scenario
from tests.lbm_math import feq_textbook, moments, second_moment


def lat(q):
    return ora.lattice(q)


def interior(a):
    return a[1:-1, 1:-1, 1:-1]


@pytest.mark.parametrize("q", [15, 19, 27])
def test_stream_is_periodic_permutation(q):
    c, w = lat(q)
    d = ora.OracleDomain(5, 6, 7, 0.8, q=q, periodic=(1, 1, 1))
    d.finalize()
    rng = np.random.default_rng(1)
    f0 = rng.integers(0, 1 << 20, size=d.ghost_shape + (q,)).astype(np.float64)
    d.set_pdf(f0)
    d.stream()
    d.swap()
    f1 = d.get_pdf()
    inner = interior(f0)
    for i in range(q):
        expect = np.roll(inner[..., i], shift=(c[i, 2], c[i, 1], c[i, 0]), axis=(0, 1, 2))
        assert np.array_equal(interior(f1)[..., i], expect), i


@pytest.mark.parametrize("q", [15, 19, 27])
@pytest.mark.parametrize("tau", [0.6, 1.0, 1.7])
def test_bgk_single_cell(q, tau):
    """One collision of one cell (a 1x1x1 periodic box streams onto itself)."""
    c, w = lat(q)
    rng = np.random.default_rng(q)
    rho0, u0 = 1.03, np.array([0.04, -0.02, 0.03])
    f = feq_textbook(c, w, rho0, u0) * (1 + 0.05 * rng.standard_normal(q))
    d = ora.OracleDomain(1, 1, 1, tau, q=q, periodic=(1, 1, 1))
    d.finalize()
    g = np.tile(f, (3, 3, 3, 1))
    d.set_pdf(g)
    d.step(1)
    f1 = d.get_pdf()[1, 1, 1]
    rho, u = moments(c, f)
    rho1, u1 = moments(c, f1)
    # mass: sum_q feq_q = rho (1 - 1.5 u.u (1 - 1/(3 C_S^2))) with the
    # reference's C_S literal, a ~1e-14 relative deficit at |u| ~ 0.05
    assert abs(rho1 / rho - 1) < 1e-13
    assert np.allclose(rho1 * u1, rho * u, rtol=0, atol=1e-11 * rho * np.abs(u).max())  # momentum
    # non-equilibrium relaxation of the second moment
    pi_eq = rho * (np.eye(3) / 3 + np.outer(u, u))
    pi, pi1 = second_moment(c, f), second_moment(c, f1)
    assert np.allclose(pi1 - pi_eq, (1 - 1 / tau) * (pi - pi_eq), rtol=0, atol=1e-11)
    # every population: f' - feq = (1 - 1/tau) (f - feq)
    fe = feq_textbook(c, w, rho, u)
    assert np.allclose(f1 - fe, (1 - 1 / tau) * (f - fe), rtol=0, atol=1e-12)


@pytest.mark.parametrize("q", [15, 19, 27])
def test_equilibrium_is_fixed_point(q):
    c, w = lat(q)
    d = ora.OracleDomain(4, 3, 5, 0.7, q=q, periodic=(1, 1, 1))
    d.finalize()
    rho0, u0 = 0.98, np.array([0.01, 0.03, -0.02])
    d.set_pdf(np.broadcast_to(feq_textbook(c, w, rho0, u0), d.ghost_shape + (q,)).copy())
    d.step(10)
    rho, u = d.moments()
    assert np.allclose(rho, rho0, rtol=0, atol=1e-12)
    assert np.allclose(u, u0, rtol=0, atol=1e-11)


@pytest.mark.parametrize("q", [15, 19, 27])
def test_periodic_conservation(q):
    c, w = lat(q)
    d = ora.OracleDomain(6, 5, 4, 0.8, q=q, periodic=(1, 1, 1))
    d.finalize()
    rng = np.random.default_rng(7)
    rho = 1 + 0.01 * rng.standard_normal((6, 7, 8))
    u = 0.03 * rng.standard_normal((6, 7, 8, 3))
    d.set_pdf(feq_textbook(c, w, rho, u) * (1 + 0.01 * rng.standard_normal((6, 7, 8, q))))
    m0 = d.conservation()
    d.step(20)
    m1 = d.conservation()
    assert abs(m1[0] / m0[0] - 1) < 1e-14
    assert np.allclose(m1[1:], m0[1:], rtol=0, atol=1e-10 * np.abs(m0[1:]).max())


@pytest.mark.parametrize("q", [15, 19, 27])
def test_closed_noslip_box_conserves_mass(q):
    c, w = lat(q)
    xl, yl, zl = 6, 5, 7
    d = ora.OracleDomain(xl, yl, zl, 0.9, q=q)
    for ext in ("z0", "zmax", "x0", "xmax", "y0", "ymax"):
        d.set_boundary_box("noslip", None, *scenario.face_box(ext, xl, yl, zl))
    d.finalize()
    rng = np.random.default_rng(3)
    rho = 1 + 0.01 * rng.standard_normal(d.ghost_shape)
    u = 0.02 * rng.standard_normal(d.ghost_shape + (3,))
    d.set_pdf(feq_textbook(c, w, rho, u))
    # step 1 pulls the (arbitrary) stored wall values; bounce-back starts at step 2
    d.step(1)
    m0 = d.conservation()
    d.step(50)
    m1 = d.conservation()
    assert abs(m1[0] / m0[0] - 1) < 1e-13


@pytest.mark.parametrize("q", [15, 19, 27])
def test_couette_flow_is_linear(q):
    """Steady plane Couette flow between half-way bounce-back walls at
    z = 1/2 (no-slip) and z = zl + 1/2 (moving at U): u_x = U (z - 1/2) / zl."""
    xl, yl, zl, U = 3, 3, 8, 0.01
    d = ora.OracleDomain(xl, yl, zl, 1.0, q=q, periodic=(1, 1, 0))
    d.set_boundary_box("noslip", None, *scenario.face_box("z0", xl, yl, zl))
    d.set_boundary_box("movingwall", [U, 0, 0, 1], *scenario.face_box("zmax", xl, yl, zl))
    d.finalize()
    d.step(4000)
    _, u = d.moments()
    z = np.arange(1, zl + 1)
    expect = U * (z - 0.5) / zl
    assert np.allclose(u[:, :, :, 0], expect[:, None, None], rtol=0, atol=1e-9)
    assert np.abs(u[..., 1:]).max() < 1e-12


def _uniform(q, rho0, u0, shape):
    c, w = lat(q)
    return np.broadcast_to(feq_textbook(c, w, rho0, u0), shape + (q,)).copy()


@pytest.mark.parametrize("q", [15, 19, 27])
def test_uniform_flow_passes_inflow_outflow(q):
    xl, yl, zl = 8, 3, 3
    uin = np.array([0.03, 0.0, 0.0])
    d = ora.OracleDomain(xl, yl, zl, 0.8, q=q, periodic=(0, 1, 1))
    d.set_boundary_box("inflow", [*uin, 1.0], *scenario.face_box("x0", xl, yl, zl))
    d.set_boundary_box("outflow", [0, 0, 0, 1.0], *scenario.face_box("xmax", xl, yl, zl))
    d.finalize()
    d.set_pdf(_uniform(q, 1.0, uin, d.ghost_shape))
    d.step(30)
    rho, u = d.moments()
    assert np.allclose(rho, 1.0, rtol=0, atol=1e-11)
    assert np.allclose(u, uin, rtol=0, atol=1e-11)


@pytest.mark.parametrize("q", [19])
def test_channel_from_rest_reaches_inflow_velocity(q):
    xl, yl, zl = 10, 3, 3
    uin = np.array([0.02, 0.0, 0.0])
    d = ora.OracleDomain(xl, yl, zl, 0.8, q=q, periodic=(0, 1, 1))
    d.set_boundary_box("inflow", [*uin, 1.0], *scenario.face_box("x0", xl, yl, zl))
    d.set_boundary_box("outflow", [0, 0, 0, 1.0], *scenario.face_box("xmax", xl, yl, zl))
    d.finalize()
    d.step(3000)
    rho, u = d.moments()
    assert np.allclose(u, uin, rtol=0, atol=2e-5)


@pytest.mark.parametrize("q", [15, 19, 27])
def test_uniform_state_passes_pressure_boundary(q):
    xl, yl, zl = 3, 3, 6
    d = ora.OracleDomain(xl, yl, zl, 0.8, q=q, periodic=(1, 1, 0))
    d.set_boundary_box("pressure", [0, 0, 0, 1.01], *scenario.face_box("z0", xl, yl, zl))
    d.set_boundary_box("pressure", [0, 0, 0, 1.01], *scenario.face_box("zmax", xl, yl, zl))
    d.finalize()
    d.set_pdf(_uniform(q, 1.01, np.zeros(3), d.ghost_shape))
    d.step(30)
    rho, u = d.moments()
    assert np.allclose(rho, 1.01, rtol=0, atol=1e-11)
    assert np.abs(u).max() < 1e-11


@pytest.mark.parametrize("q", [15, 19, 27])
def test_freeslip_keeps_tangential_flow_and_momentum(q):
    xl, yl, zl = 3, 6, 3
    d = ora.OracleDomain(xl, yl, zl, 0.8, q=q, periodic=(1, 0, 1))
    d.set_boundary_box("freeslip", None, *scenario.face_box("y0", xl, yl, zl))
    d.set_boundary_box("freeslip", None, *scenario.face_box("ymax", xl, yl, zl))
    d.finalize()
    u0 = np.array([0.02, 0.0, 0.01])
    d.set_pdf(_uniform(q, 1.0, u0, d.ghost_shape))
    d.step(40)
    rho, u = d.moments()
    assert np.allclose(u, u0, rtol=0, atol=1e-11)
    # a sheared profile keeps its total tangential momentum (no wall stress)
    c, w = lat(q)
    y = np.arange(d.ghost_shape[1])
    prof = np.zeros(d.ghost_shape + (3,))
    prof[..., 0] = 0.02 * np.cos(np.pi * (y[None, :, None] - 0.5) / yl)
    d2 = ora.OracleDomain(xl, yl, zl, 0.8, q=q, periodic=(1, 0, 1))
    d2.set_boundary_box("freeslip", None, *scenario.face_box("y0", xl, yl, zl))
    d2.set_boundary_box("freeslip", None, *scenario.face_box("ymax", xl, yl, zl))
    d2.finalize()
    d2.set_pdf(feq_textbook(c, w, np.ones(d.ghost_shape), prof))
    m0 = d2.conservation()
    d2.step(100)
    m1 = d2.conservation()
    assert abs(m1[1] - m0[1]) < 1e-12 * abs(m0[1]) + 1e-14


@pytest.mark.parametrize("q", [15, 19, 27])
def test_first_step_semantics_and_moving_wall_rule(q):
    """Step 1 pulls the initial weights from every boundary cell, so the whole
    cavity stays uniform; in step 2 a lid link carries Ladd's moving-wall
    value f_qbar + 2 w_q rho c_q.u_w / c_s^2 with rho of the adjacent fluid
    cell's post-collision pdfs."""
    c, w = lat(q)
    xl, yl, zl, U = 5, 4, 6, 0.05
    d = ora.OracleDomain(xl, yl, zl, 0.6, q=q)
    scenario.cavity(xl, yl, zl, lid=(U, 0, 0)).apply(d)
    d.finalize()
    d.step(1)
    f1 = d.get_pdf()
    inner = interior(f1).reshape(-1, q)
    assert (inner == inner[0]).all()
    d.stream()
    d.swap()
    f2 = d.get_pdf()
    rho1 = 0.0
    for v in inner[0]:
        rho1 += v
    cs2 = ora.cs() ** 2
    x, y, z = 2, 2, zl  # lid-adjacent interior cell away from the side walls
    for i in range(q):
        if c[i, 2] == -1:  # pulled from the lid at z = zl + 1
            expect = inner[0][q - 1 - i] + 2.0 * w[i] * rho1 * (c[i, 0] * U) / cs2
            assert abs(f2[z, y, x, i] - expect) < 1e-16, i
        else:
            assert f2[z, y, x, i] == inner[0][i]


def test_last_wins_box_order():
    """cavity.xml: the side faces come after the lid, so the lid's edges and
    corners are no-slip (include/io/scenario.h:183-186, domain.hpp:189-190)."""
    xl, yl, zl = 5, 4, 6
    d = ora.OracleDomain(xl, yl, zl, 0.6, q=19)
    scenario.cavity(xl, yl, zl).apply(d)
    k = d.kinds()  # (z, y, x)
    assert (k[zl + 1, 1:yl + 1, 1:xl + 1] == ora.MOVINGWALL).all()
    assert (k[zl + 1, 0, :] == ora.NOSLIP).all() and (k[zl + 1, :, 0] == ora.NOSLIP).all()
    assert (k[zl + 1, yl + 1, :] == ora.NOSLIP).all() and (k[zl + 1, :, xl + 1] == ora.NOSLIP).all()
    assert (interior(k) == 0).all()


@pytest.mark.parametrize("rho_ref", [1.0, 1.2])
def test_outflow_divides_by_reference_density(rho_ref):
    """Outflow value = feq_q + feq_qbar - f_qbar of the adjacent cell, with
    u = momentum / rho_ref and feq at rho_ref (include/boundary.hpp:143-147)."""
    q = 19
    c, w = lat(q)
    xl, yl, zl = 4, 2, 2
    d = ora.OracleDomain(xl, yl, zl, 0.8, q=q, periodic=(0, 1, 1))
    d.set_boundary_box("noslip", None, *scenario.face_box("x0", xl, yl, zl))
    d.set_boundary_box("outflow", [0, 0, 0, rho_ref], *scenario.face_box("xmax", xl, yl, zl))
    d.finalize()
    rng = np.random.default_rng(5)
    f = _uniform(q, 1.0, np.array([0.03, 0.01, 0]), d.ghost_shape) * (1 + 0.02 * rng.standard_normal(d.ghost_shape + (q,)))
    d.set_pdf(f)
    d.step(1)
    fa = d.get_pdf()
    n = fa[1, 1, xl]  # fluid cell next to the outflow face
    d.stream()
    d.swap()
    fb = d.get_pdf()
    j = n @ c.astype(float)
    fe = feq_textbook(c, w, rho_ref, j / rho_ref)
    for i in range(q):
        if c[i, 0] == -1:  # pulled from the outflow cell at x = xl + 1
            expect = fe[i] + fe[q - 1 - i] - n[q - 1 - i]
            assert abs(fb[1, 1, xl, i] - expect) < 1e-11, i


def test_unassigned_ghost_rejected():
    d = ora.OracleDomain(4, 4, 4, 0.8, q=19)
    for ext in ("z0", "zmax", "x0", "xmax", "y0"):  # ymax missing
        d.set_boundary_box("noslip", None, *scenario.face_box(ext, 4, 4, 4))
    with pytest.raises(ValueError):
        d.finalize()


@pytest.mark.parametrize("q", [15, 19, 27])
def test_init_equilibrium_is_textbook_feq(q):
    """f = feq(rho, u) per cell (include/collision.hpp:34-51 with the
    reference's C_S literal), so the moments return rho and u."""
    c, w = lat(q)
    rng = np.random.default_rng(5)
    n = 3
    rho = 1 + 0.01 * rng.standard_normal((n, n, n))
    u = 0.03 * rng.standard_normal((n, n, n, 3))
    d = ora.OracleDomain(n, n, n, 0.7, q=q, periodic=(1, 1, 1))
    d.finalize()
    d.init_equilibrium(rho, u)
    f = interior(d.get_pdf())
    assert np.allclose(f, feq_textbook(c, w, rho, u), rtol=0, atol=1e-13)
    r1, u1 = d.moments()
    assert np.allclose(r1, rho, rtol=1e-12, atol=0)
    assert np.allclose(u1, u, rtol=0, atol=1e-12)


@pytest.mark.parametrize("q", [15, 19, 27])
@pytest.mark.parametrize("alpha", [0.0, 0.3, 1.0])
def test_brinkman_uniform_box_recurrence(q, alpha):
    """Uniform porosity and body force on a periodic box at rest: every cell
    stays uniform and u_n obeys u' = u (1 - omega alpha) + omega (1 - alpha)
    tau g, whose fixed point is the Darcy plateau (1 - alpha) tau g / alpha;
    alpha = 0 accelerates freely (u_n = n g) and alpha = 1 only decays."""
    tau, g, n, steps = 0.9, np.array([2e-5, -1e-5, 5e-6]), 2, 60
    d = ora.OracleDomain(n, n, n, tau, q=q, periodic=(1, 1, 1))
    d.set_brinkman(np.full(n ** 3, alpha), g)
    d.finalize()
    d.init_equilibrium(np.ones((n, n, n)), np.zeros((n, n, n, 3)))
    om = 1 / tau
    u = np.zeros(3)
    for _ in range(steps):
        u = u * (1 - om * alpha) + om * (1 - alpha) * tau * g
    d.step(steps)
    rho1, u1 = d.moments()
    assert np.allclose(u1, u, rtol=1e-9, atol=1e-15)
    assert np.allclose(rho1, 1.0, rtol=0, atol=1e-13)
    if alpha == 0.0:
        assert np.allclose(u1, steps * g, rtol=1e-9)
    if 0 < alpha < 1:
        plateau = (1 - alpha) * tau * g / alpha
        assert np.all(np.abs(u) < np.abs(plateau))
