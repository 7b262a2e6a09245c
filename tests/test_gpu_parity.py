"""GPU parity: the HIP path through the C ABI against the CPU oracle.

F64 (the reference's fp64 operation order) must be bit-identical to the
oracle on every compared value. The shifted storage modes are compared with
the oracle's fp64 fields through relative L2 / max norms with the tolerances
written in each test. Non-fluid cells are compared too (the values the
reference's boundary pass leaves in them), except ghost cells of periodic
axes, which neither side reads.
"""
import sys

import numpy as np
import pytest

from lbm_amd import capi, scenario
from oracle import ora
from tests.lbm_math import feq_textbook, tgv_fields

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_pair(q, xl, yl, zl, tau, sc=None, periodic=(0, 0, 0), storage="f64", mask=None):
    g = capi.Domain(xl, yl, zl, tau, q=q, storage=storage, periodic=periodic)
    o = ora.OracleDomain(xl, yl, zl, tau, q=q, periodic=periodic)
    for d in (g, o):
        if mask is not None:
            d.set_solid_mask(mask)
        if sc is not None:
            sc.apply(d)
        d.finalize()
    return g, o


def compared_cells(o):
    """Ghosted-cell mask of what both sides define."""
    m = np.ones(o.ghost_shape, dtype=bool)
    for axis, per in zip((2, 1, 0), o.periodic):
        if per:
            idx = [slice(None)] * 3
            idx[axis] = 0
            m[tuple(idx)] = False
            idx[axis] = -1
            m[tuple(idx)] = False
    return m


def assert_pdf_identical(g, o, label=""):
    fg, fo = g.get_pdf(), o.get_pdf()
    m = compared_cells(o)
    diff = fg[m] != fo[m]
    if diff.any():
        where = np.argwhere((fg != fo).any(-1) & m)[:5]
        raise AssertionError(f"{label}: {diff.sum()} pdf values differ; first cells (z,y,x) {where.tolist()}; "
                             f"max |d| {np.abs(fg[m] - fo[m]).max():.3e}")


def fluid_interior(o):
    return o.kinds()[1:-1, 1:-1, 1:-1] == 0


def rel_err(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)), \
        float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


# ---------------------------------------------------------------- F64 exact --
@pytest.mark.parametrize("q", [15, 19, 27])
@pytest.mark.parametrize("dims", [(8, 8, 8), (6, 7, 9)])
def test_cavity_small_bit_exact(q, dims):
    xl, yl, zl = dims
    say(f"cavity {dims} q={q}")
    g, o = make_pair(q, xl, yl, zl, 0.6, scenario.cavity(xl, yl, zl))
    assert_pdf_identical(g, o, "initial")
    done = 0
    for t in (1, 2, 3, 10, 100):
        g.step(t - done)
        o.step(t - done)
        done = t
        assert_pdf_identical(g, o, f"step {t}")
        rg, ug = g.moments()
        ro, uo = o.moments()
        m = fluid_interior(o)
        assert np.array_equal(rg[m], ro[m]) and np.array_equal(ug[m], uo[m]), t


def test_c1_cavity_64_bit_exact():
    """Lid-driven cavity, D3Q19, 64^3, tau 0.6, lid 0.05 along x on zmax,
    1000 steps from rest; rho and u compared every 100 steps."""
    n = 64
    say("C1 cavity 64^3 q=19 1000 steps")
    g, o = make_pair(19, n, n, n, 0.6, scenario.cavity(n, n, n, lid=(0.05, 0, 0)))
    m = fluid_interior(o)
    for t in range(100, 1001, 100):
        g.step(100)
        o.step(100)
        rg, ug = g.moments()
        ro, uo = o.moments()
        er = rel_err(rg[m], ro[m])
        eu = rel_err(ug[m], uo[m])
        say(f"  t={t} rho rel L2/max {er[0]:.2e}/{er[1]:.2e}  u {eu[0]:.2e}/{eu[1]:.2e}")
        assert np.array_equal(rg[m], ro[m]), t
        assert np.array_equal(ug[m], uo[m]), t
    assert np.abs(uo[m]).max() > 1e-3  # the lid has driven a flow


def pipe_mask(xl, yl, zl, r, centre=None):
    z, y, x = np.meshgrid(np.arange(1, zl + 1), np.arange(1, yl + 1), np.arange(1, xl + 1), indexing="ij")
    cx, cy, cz = centre or (xl / 3, (yl + 1) / 2, (zl + 1) / 2)
    solid = (x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 <= r * r
    return (~solid).astype(np.uint8)


@pytest.mark.parametrize("q", [19, 27])
def test_channel_inflow_outflow_obstacle_bit_exact(q):
    xl, yl, zl = 24, 10, 9
    sc = scenario.Scenario("pipe", xl, yl, zl)
    for ext, kind, p in [("z0", 1, None), ("zmax", 1, None), ("x0", 5, (0.03, 0, 0, 1.0)),
                         ("xmax", 4, (0, 0, 0, 1.0)), ("y0", 1, None), ("ymax", 1, None)]:
        sc.boundaries.append(scenario.Boundary(kind, p or (0, 0, 0, 1.0), scenario.face_box(ext, xl, yl, zl)))
    say(f"channel+sphere q={q}")
    g, o = make_pair(q, xl, yl, zl, 0.7, sc, mask=pipe_mask(xl, yl, zl, 3.0))
    for t in (1, 2, 50, 200):
        g.step(t - (0 if t == 1 else prev))
        o.step(t - (0 if t == 1 else prev))
        prev = t
        assert_pdf_identical(g, o, f"step {t}")


@pytest.mark.parametrize("name", ["step.xml", "shearflow.xml"])
@pytest.mark.parametrize("q", [15, 19, 27])
def test_reference_scenarios_bit_exact(name, q):
    from pathlib import Path
    sc = scenario.parse_scenario_file(Path(__file__).parent / "golden" / "scenarios" / name)
    say(f"{name} q={q}")
    tau = 0.6
    g, o = make_pair(q, sc.xl, sc.yl, sc.zl, tau, sc)
    for t in (1, 2, 30, 100):
        g.step(t - (0 if t == 1 else prev))
        o.step(t - (0 if t == 1 else prev))
        prev = t
        assert_pdf_identical(g, o, f"{name} step {t}")


@pytest.mark.parametrize("q", [15, 19, 27])
def test_periodic_tgv_bit_exact(q):
    n = 24
    say(f"periodic TGV {n}^3 q={q}")
    g, o = make_pair(q, n, n + 2, n - 3, 0.8, periodic=(1, 1, 1))
    rho, u = tgv_fields(n, n + 2, n - 3, u0=0.03)
    c, w = ora.lattice(q)
    f = np.broadcast_to(np.array(w), (n - 1, n + 4, n + 2, q)).copy()
    f[1:-1, 1:-1, 1:-1] = feq_textbook(c, w, rho, u)
    g.set_pdf(f)
    o.set_pdf(f)
    g.step(40)
    o.step(40)
    assert_pdf_identical(g, o, "tgv")


@pytest.mark.parametrize("q", [19, 27])
def test_couette_mixed_periodic_bit_exact(q):
    xl, yl, zl = 7, 5, 11
    sc = scenario.Scenario("couette", xl, yl, zl)
    sc.boundaries.append(scenario.Boundary(1, (0, 0, 0, 1.0), scenario.face_box("z0", xl, yl, zl)))
    sc.boundaries.append(scenario.Boundary(2, (0.04, 0.01, 0, 1.0), scenario.face_box("zmax", xl, yl, zl)))
    say(f"couette q={q}")
    g, o = make_pair(q, xl, yl, zl, 0.9, sc, periodic=(1, 1, 0))
    g.step(150)
    o.step(150)
    assert_pdf_identical(g, o, "couette")


@pytest.mark.parametrize("storage,tol", [("f64s", 1e-10), ("f32", 2e-5)])
@pytest.mark.parametrize("q", [15, 19, 27])
def test_brinkman_channel_against_oracle(q, storage, tol):
    """Porous channel: random porosity in [0, 1] per cell, body force, no-slip
    walls in y and z, periodic in x, 300 steps from rest, against the
    oracle's Brinkman extension (fp64, reference operation order)."""
    xl, yl, zl = 12, 9, 7
    rng = np.random.default_rng(31 + q)
    alpha = rng.uniform(0.0, 1.0, xl * yl * zl)
    g3 = (2e-5, 5e-6, 0.0)
    sc = scenario.Scenario("porous", xl, yl, zl)
    for ext in ("y0", "ymax", "z0", "zmax"):
        sc.boundaries.append(scenario.Boundary(1, (0, 0, 0, 1.0), scenario.face_box(ext, xl, yl, zl)))
    say(f"brinkman channel q={q} {storage}")
    g = capi.Domain(xl, yl, zl, 0.7, q=q, storage=storage, periodic=(1, 0, 0))
    o = ora.OracleDomain(xl, yl, zl, 0.7, q=q, periodic=(1, 0, 0))
    for d in (g, o):
        sc.apply(d)
        d.set_brinkman(alpha, g3)
        d.finalize()
    g.step(300)
    o.step(300)
    rg, ug = g.moments()
    ro, uo = o.moments()
    eu = rel_err(ug, uo)
    er = rel_err(rg - 1, ro - 1)
    say(f"  u rel L2 {eu[0]:.2e} max {eu[1]:.2e}; drho rel L2 {er[0]:.2e}; max |u| {np.abs(uo).max():.2e}")
    assert np.abs(uo).max() > 1e-5  # the force drives a flow
    assert eu[0] < tol and eu[1] < 10 * tol


def test_set_pdf_midrun_bit_exact():
    n = 10
    say("set_pdf mid-run")
    g, o = make_pair(19, n, n, n, 0.6, scenario.cavity(n, n, n))
    g.step(5)
    o.step(5)
    assert_pdf_identical(g, o, "before set")
    f = o.get_pdf()
    rng = np.random.default_rng(11)
    f = f * (1 + 1e-3 * rng.standard_normal(f.shape))
    g.set_pdf(f)
    o.set_pdf(f)
    g.step(7)
    o.step(7)
    assert_pdf_identical(g, o, "after set")


def test_conservation_matches_oracle_and_closed_box():
    n = 16
    say("conservation")
    sc = scenario.Scenario("box", n, n, n)
    for ext in ("z0", "zmax", "x0", "xmax", "y0", "ymax"):
        sc.boundaries.append(scenario.Boundary(1, (0, 0, 0, 1.0), scenario.face_box(ext, n, n, n)))
    g, o = make_pair(19, n, n, n, 0.7, sc)
    rho, u = tgv_fields(n, n, n, u0=0.02)
    g.init_equilibrium(rho, u)
    f = g.get_pdf()
    o.set_pdf(f)
    g.step(1)
    o.step(1)
    m0 = g.conservation()
    assert np.allclose(m0, o.conservation(), rtol=1e-13, atol=1e-13)
    g.step(200)
    m1 = g.conservation()
    assert abs(m1[0] / m0[0] - 1) < 1e-13


def test_geometry_errors():
    say("geometry errors")
    g = capi.Domain(6, 6, 6, 0.7, q=19)
    for ext in ("z0", "zmax", "x0", "xmax", "y0"):
        g.set_boundary_box("noslip", None, *scenario.face_box(ext, 6, 6, 6))
    with pytest.raises(capi.LbmError) as e:
        g.finalize()
    assert e.value.code == -6
    with pytest.raises(capi.LbmError):
        g.set_boundary_box("noslip", None, 0, 8, 0, 0, 0, 0)  # beyond the ghost layer
    h = capi.Domain(4, 4, 4, 0.7, q=19, storage="f64")
    with pytest.raises(capi.LbmError) as e:
        h.set_brinkman(np.zeros(64), (1e-5, 0, 0))
    assert e.value.code == -5
    with pytest.raises(capi.LbmError) as e:
        h.step(1)
    assert e.value.code == -2


# ----------------------------------------------------------- shifted modes --
# Tolerances: relative L2 / max of velocity against the fp64 oracle after the
# stated steps. F64S differs from the reference only by c_s^2 = 1/3 exactly
# and summation order; F32 by fp32 rounding; F16 and BF16 by their storage
# precision (11 and 8 significand bits).
SHIFTED_TOL = {"f64s": 1e-9, "f32": 2e-4, "f16": 5e-3, "bf16": 5e-2}


@pytest.mark.parametrize("storage", ["f64s", "f32", "f16", "bf16"])
@pytest.mark.parametrize("q", [15, 19, 27])
def test_shifted_cavity_error(storage, q):
    n = 24
    say(f"shifted cavity {storage} q={q}")
    g, o = make_pair(q, n, n, n, 0.6, scenario.cavity(n, n, n), storage=storage)
    g.step(300)
    o.step(300)
    m = fluid_interior(o)
    rg, ug = g.moments()
    ro, uo = o.moments()
    eu = rel_err(ug[m], uo[m])
    er = rel_err(rg[m] - 1, ro[m] - 1)
    say(f"  {storage} q={q}: u rel L2 {eu[0]:.2e} max {eu[1]:.2e}; drho rel L2 {er[0]:.2e}")
    assert eu[0] < SHIFTED_TOL[storage]


@pytest.mark.parametrize("storage", ["f64s", "f32", "f16", "bf16"])
def test_shifted_periodic_tgv_error(storage):
    n = 32
    say(f"shifted TGV {storage}")
    q = 19
    g, o = make_pair(q, n, n, n, 0.8, periodic=(1, 1, 1), storage=storage)
    rho, u = tgv_fields(n, n, n, u0=0.02)
    g.init_equilibrium(rho, u)
    c, w = ora.lattice(q)
    f = np.broadcast_to(np.array(w), (n + 2, n + 2, n + 2, q)).copy()
    f[1:-1, 1:-1, 1:-1] = feq_textbook(c, w, rho, u)
    o.set_pdf(f)
    g.step(200)
    o.step(200)
    rg, ug = g.moments()
    ro, uo = o.moments()
    eu = rel_err(ug, uo)
    say(f"  {storage}: u rel L2 {eu[0]:.2e} max {eu[1]:.2e}")
    assert eu[0] < SHIFTED_TOL[storage]


VECS = {"f64": (1, 2), "f64s": (1, 2), "f32": (1, 2, 4), "f16": (2, 4, 8), "bf16": (2, 4, 8)}


@pytest.mark.parametrize("storage", ["f64", "f64s", "f32", "f16", "bf16"])
@pytest.mark.parametrize("q", [15, 19, 27])
def test_results_independent_of_cells_per_thread(storage, q):
    """Kernels round every product and sum as written (no FMA contraction), so
    the kernel form (cell tiles, or row kernels with V cells per thread plus
    the near-boundary kernel) and the block width do not change a bit of the
    result: a walled cavity and a periodic Taylor-Green box with an odd x
    length (partial vectors). Rows of 130 cells and more keep the line-aligned
    layout the vector row kernels need in every storage format. The periodic
    box steps singly here (the two-step pass has its own tests)."""
    say(f"V independence {storage} q={q}")
    runs = {}
    for v in (0,) + VECS[storage]:
        for bx in (0, 64):
            cav = capi.Domain(130, 7, 5, 0.6, q=q, storage=storage)
            scenario.cavity(130, 7, 5).apply(cav)
            cav.finalize()
            per = capi.Domain(131, 6, 4, 0.8, q=q, storage=storage, periodic=(1, 1, 1))
            per.finalize()
            per.set_time_blocking(1)
            rho, u = tgv_fields(131, 6, 4, u0=0.03)
            per.init_equilibrium(rho, u)
            for d in (cav, per):
                d.set_tuning(v, bx)
                d.step(23)
            if v == 0:
                assert cav.info()["cell_tiles"] == 1 and per.info()["cell_tiles"] == 0
            runs[(v, bx)] = (cav.get_pdf(), per.get_pdf())
    ref = runs[(VECS[storage][0], 0)]
    for key, (fc, fp) in runs.items():
        assert np.array_equal(fc.view(np.uint64), ref[0].view(np.uint64)), key
        assert np.array_equal(fp.view(np.uint64), ref[1].view(np.uint64)), key


# --------------------------------------------------- full-size properties --
@pytest.mark.parametrize("q", [15, 19, 27])
def test_tgv_256_decay_and_mass(q):
    """Periodic 256^3 Taylor-Green vortex (u0 = 0.02, k = 2 pi / 256, density
    noise 1e-4), tau = 0.8, fp32 shifted storage, 500 steps: kinetic energy
    decays as exp(-4 nu k^2 t) with nu = (tau - 1/2)/3; mass is conserved."""
    n, tau, steps = 256, 0.8, 500
    say(f"TGV 256^3 q={q}")
    g = capi.Domain(n, n, n, tau, q=q, storage="f32", periodic=(1, 1, 1))
    g.finalize()
    rho, u = tgv_fields(n, n, n, u0=0.02)
    g.init_equilibrium(rho, u)
    m0 = g.conservation()
    _, u0 = g.moments()
    e0 = float((u0 ** 2).sum())
    g.step(steps)
    m1 = g.conservation()
    _, u1 = g.moments()
    e1 = float((u1 ** 2).sum())
    nu = (tau - 0.5) / 3
    k = 2 * np.pi / n
    rate = np.log(e1 / e0) / (-4 * nu * k * k * steps)
    say(f"  decay rate / analytic = {rate:.6f}; mass drift {m1[0] / m0[0] - 1:.2e}")
    assert abs(rate - 1) < 5e-3
    assert abs(m1[0] / m0[0] - 1) < 1e-6


@pytest.mark.parametrize("storage", ["f64", "f32", "f16"])
def test_packed_rows_layout(storage):
    """A domain whose rows are shorter than two 128-byte lines packs them at
    pitch xl + 2 and steps with cell tiles: fp64 matches the oracle bit for
    bit, fp32/fp16 within their storage tolerance; vector row kernels are
    refused for it."""
    n = 20
    say(f"packed rows {storage}")
    g, o = make_pair(19, n, n, n, 0.6, scenario.cavity(n, n, n), storage=storage)
    info = g.info()
    assert info["cell_tiles"] == 1
    g.step(60)
    o.step(60)
    if storage == "f64":
        assert_pdf_identical(g, o, "packed")
    else:
        m = fluid_interior(o)
        eu = rel_err(g.moments()[1][m], o.moments()[1][m])
        assert eu[0] < SHIFTED_TOL[storage]
    if storage in ("f64", "f32"):
        with pytest.raises(capi.LbmError):
            g.set_tuning(2, 0)
