"""Two time steps per pass on domains with boundary cells (k_emit_rules +
k_wall2) against single steps.

With fp32 storage (by default) or fp16/bf16 storage (on request), walls,
solid obstacles and inflow/outflow/pressure faces step two steps per pass: before the pass every boundary link gets
its rule value from the stored values of the fluid cell it serves, step 1
pulls plainly, and a cell next to a non-fluid cell leaves the rule values its
step 2 needs, computed from its own step-1 values, in its own LDS slots. The
arithmetic is the single step's, so every fluid cell below must equal the
same run in single steps bit for bit, and so must the whole field as
lbm_get_pdf reports it (it writes the boundary links' rule values itself):
tiles cut by walls on every face, partial tiles, solid spheres, inflow,
outflow, pressure and moving walls, periodic axes mixed with walls, odd step
counts, several z chunk sizes, the reference's custom-extent scenario and the
256^3 cavity. The rule values cross memory in the storage format in both
forms (the single step rounds them through it too), so 16-bit storage is
bitwise as well. FreeSlip, Brinkman and fp64 walled domains keep single steps.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

from lbm_amd import capi, scenario, workloads

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).parent / "golden" / "scenarios"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def run(shape, q, storage, tau, sc, spp, steps, z_chunk=0, periodic=(0, 0, 0), mask=None):
    d = capi.Domain(*shape, tau, q=q, storage=storage, periodic=periodic)
    if mask is not None:
        d.set_solid_mask(mask)
    sc.apply(d)
    d.finalize()
    d.set_time_blocking(spp, z_chunk)
    info = d.info()
    d.step(steps)
    r, u = d.moments()
    f = d.get_pdf()
    d.close()
    fluid = np.ones(shape[::-1], dtype=bool) if mask is None else mask.astype(bool)
    if sc.fluid_mask is not None:
        fluid &= sc.fluid_mask.astype(bool)
    return info, r[fluid], u[fluid], f


def check(shape, q, storage, tau, sc, steps, periodic=(0, 0, 0), mask=None, chunks=(0,), label=""):
    _, r1, u1, f1 = run(shape, q, storage, tau, sc, 1, steps, periodic=periodic, mask=mask)
    for zc in chunks:
        info, r2, u2, f2 = run(shape, q, storage, tau, sc, 2, steps, z_chunk=zc, periodic=periodic, mask=mask)
        assert info["steps_per_pass"] == 2, (label, info)
        assert np.array_equal(bits(r1), bits(r2)) and np.array_equal(bits(u1), bits(u2)), (label, zc)
        same = bits(f1) == bits(f2)
        assert same.all(), f"{label} z_chunk={zc}: {int((~same).sum())} pdf values differ"
    assert np.abs(u1).max() > 1e-5, label


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("q", [15, 19, 27])
def test_cavity_pass_equals_single_steps(q, storage):
    for shape in ((70, 19, 13), (40, 9, 5), (130, 17, 33)):
        say(f"walled pass: cavity {shape} q={q} {storage}")
        sc = scenario.cavity(*shape, lid=(0.05, 0.02, 0.0))
        check(shape, q, storage, 0.6, sc, 9, chunks=(0, 1, 3), label=f"cavity {shape}")


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("q", [19, 27])
def test_sphere_channel_pass_equals_single_steps(q, storage):
    shape = (96, 40, 36)
    say(f"walled pass: sphere channel {shape} q={q} {storage}")
    sc = workloads.sphere_channel(*shape, u_in=0.04, solid_fraction=0.2, seed=7)
    check(shape, q, storage, 0.6, sc, 31, label="sphere channel")


def test_pressure_and_periodic_axes_pass_equals_single_steps():
    # channel along x: inflow at x0, pressure at xmax, walls in y, periodic z
    shape = (80, 24, 10)
    sc = scenario.Scenario("pressure channel", *shape)
    for ext, kind, p in [("y0", capi.NOSLIP, None), ("ymax", capi.MOVINGWALL, (0.03, 0.0, 0.01, 1.0)),
                         ("x0", capi.INFLOW, (0.03, 0.0, 0.0, 1.0)), ("xmax", capi.PRESSURE, (0.0, 0.0, 0.0, 0.995))]:
        sc.boundaries.append(scenario.Boundary(kind, p or (0.0, 0.0, 0.0, 1.0), scenario.face_box(ext, *shape), "", ext))
    for storage in ("f32",):
        say(f"walled pass: pressure channel, periodic z, {storage}")
        check(shape, 19, storage, 0.7, sc, 24, periodic=(0, 0, 1), chunks=(0, 3), label="pressure channel")
    # outflow + solid slab, periodic y
    sc2 = scenario.Scenario("outflow channel", *shape)
    for ext, kind, p in [("z0", capi.NOSLIP, None), ("zmax", capi.NOSLIP, None),
                         ("x0", capi.INFLOW, (0.04, 0.0, 0.0, 1.0)), ("xmax", capi.OUTFLOW, (0.0, 0.0, 0.0, 1.0))]:
        sc2.boundaries.append(scenario.Boundary(kind, p or (0.0, 0.0, 0.0, 1.0), scenario.face_box(ext, *shape), "", ext))
    mask = np.ones(shape[::-1], dtype=np.uint8)
    mask[2:7, 5:15, 30:38] = 0
    say("walled pass: outflow channel with a solid block, periodic y")
    check(shape, 27, "f32", 0.65, sc2, 17, periodic=(0, 1, 0), mask=mask, label="outflow channel")


@pytest.mark.parametrize("q", [15, 19, 27])
def test_reference_step_scenario_pass_equals_single_steps(q):
    sc = scenario.parse_scenario_file(GOLDEN / "step.xml")
    say(f"walled pass: step.xml q={q}")
    check((sc.xl, sc.yl, sc.zl), q, "f32", 0.6, sc, 40, label="step.xml")


def test_freeslip_brinkman_and_fp64_walled_domains_step_singly():
    sc = scenario.parse_scenario_file(GOLDEN / "shearflow.xml")
    d = capi.Domain(sc.xl, sc.yl, sc.zl, 0.6, q=19, storage="f32")
    sc.apply(d)
    d.finalize()
    assert d.info()["steps_per_pass"] == 1
    d.close()
    shape = (70, 12, 10)  # rows long enough for the row kernel (set_tuning below)
    d = capi.Domain(*shape, 0.6, q=19, storage="f32")
    scenario.cavity(*shape).apply(d)
    d.set_brinkman(np.full(shape[::-1], 0.1))
    d.finalize()
    assert d.info()["steps_per_pass"] == 1
    d.close()
    d = capi.Domain(*shape, 0.6, q=19, storage="f32")
    scenario.cavity(*shape).apply(d)
    d.finalize()
    assert d.info()["steps_per_pass"] == 2
    d.set_tuning(2, 0)
    assert d.info()["steps_per_pass"] == 1
    d.close()
    # D3Q27 fp32: the pass by default too; single steps on request
    d = capi.Domain(*shape, 0.6, q=27, storage="f32")
    scenario.cavity(*shape).apply(d)
    d.finalize()
    assert d.info()["steps_per_pass"] == 2
    d.set_time_blocking(1)
    assert d.info()["steps_per_pass"] == 1
    d.set_time_blocking(0)
    assert d.info()["steps_per_pass"] == 2
    with pytest.raises(capi.LbmError):
        d.set_time_blocking(-1)
    d.close()
    # 16-bit storage: single steps by default (measured faster), the pass on
    # request; fp64 has no walled pass
    for storage, forced in (("f16", 2), ("bf16", 2), ("f64", 1), ("f64s", 1)):
        d = capi.Domain(*shape, 0.6, q=19, storage=storage)
        scenario.cavity(*shape).apply(d)
        d.finalize()
        assert d.info()["steps_per_pass"] == 1, storage
        d.set_time_blocking(2)
        assert d.info()["steps_per_pass"] == forced, storage
        d.close()
    # 16-bit rows of 32 cells or fewer step singly even on request (as
    # bulk-only domains do)
    d = capi.Domain(32, 12, 10, 0.6, q=19, storage="f16")
    scenario.cavity(32, 12, 10).apply(d)
    d.finalize()
    d.set_time_blocking(2)
    assert d.info()["steps_per_pass"] == 1
    d.close()


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_cavity_256_pass_equals_single_steps(storage):
    """The benchmark's walled case (256^3 lid-driven cavity, D3Q19, tau 0.6):
    41 steps (the first single, then 20 passes) equal 41 single steps bit for
    bit."""
    n = 256
    say(f"walled pass: 256^3 cavity {storage}")
    sc = scenario.cavity(n, n, n, lid=(0.05, 0.0, 0.0))
    out = {}
    for spp in (1, 2):
        d = capi.Domain(n, n, n, 0.6, q=19, storage=storage)
        sc.apply(d)
        d.finalize()
        d.set_time_blocking(spp)
        d.step(41)
        out[spp] = d.moments() + (d.conservation(),)
        d.close()
    for a, b in zip(out[1], out[2]):
        assert np.array_equal(bits(a), bits(b))
