"""In-place (AA-pattern) layout: one distribution buffer updated in place.

Against the two-buffer layout on the same inputs: identical fields (every
defined ghosted cell, fp64 and fp32 storage) after odd and even numbers of
steps, for lid-driven cavities, an inflow/outflow channel with an obstacle,
the reference's step scenario, periodic and mixed-periodic boxes and a
Brinkman channel; stepping continues correctly across reads and writes of
the field; fp16 and bf16 agree bit for bit too; the layout holds one buffer;
FreeSlip, the adjoint, batching and row kernels are refused."""
import sys

import numpy as np
import pytest

from lbm_amd import capi, scenario, workloads
from oracle import ora
from tests.lbm_math import tgv_fields
from tests.test_gpu_parity import compared_cells, pipe_mask

pytestmark = pytest.mark.gpu

GOLDEN = __import__("pathlib").Path(__file__).resolve().parent / "golden" / "scenarios"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def pair(q, dims, tau, storage, sc=None, periodic=(0, 0, 0), mask=None, brinkman=None):
    out = []
    for layout in ("ab", "aa"):
        d = capi.Domain(*dims, tau, q=q, storage=storage, periodic=periodic, layout=layout)
        if mask is not None:
            d.set_solid_mask(mask)
        if sc is not None:
            sc.apply(d)
        if brinkman is not None:
            d.set_brinkman(*brinkman)
        d.finalize()
        out.append(d)
    return out


def same_fields(ab, aa, dims, periodic, label):
    fa, fb = ab.get_pdf(), aa.get_pdf()
    o = ora.OracleDomain(*dims, 1.0, q=fa.shape[-1], periodic=periodic)
    m = compared_cells(o)
    o.close()
    diff = fa[m] != fb[m]
    assert not diff.any(), f"{label}: {int(diff.sum())} values differ, max {np.abs(fa[m] - fb[m]).max():.3e}"
    ra, ua = ab.moments()
    rb, ub = aa.moments()
    assert np.array_equal(ra, rb) and np.array_equal(ua, ub), label


@pytest.mark.parametrize("storage", ["f64", "f32"])
@pytest.mark.parametrize("q", [15, 19, 27])
def test_cavity_matches_two_buffer_layout(q, storage):
    dims = (9, 8, 7)
    ab, aa = pair(q, dims, 0.6, storage, scenario.cavity(*dims, lid=(0.05, 0.02, 0.0)))
    assert aa.info()["layout"] == 1 and ab.info()["layout"] == 0
    done = 0
    for n in (1, 1, 5, 30):
        ab.step(n)
        aa.step(n)
        done += n
        same_fields(ab, aa, dims, (0, 0, 0), f"cavity q={q} {storage} after {done} steps")
    say(f"cavity q={q} {storage}: identical after 1, 2, 7, 37 steps")


@pytest.mark.parametrize("storage", ["f64", "f32"])
@pytest.mark.parametrize("q", [19, 27])
def test_channel_with_obstacle(q, storage):
    xl, yl, zl = 14, 9, 9
    sc = scenario.Scenario("pipe", xl, yl, zl)
    for ext, kind, p in [("z0", 1, None), ("zmax", 1, None), ("y0", 1, None), ("ymax", 1, None),
                         ("x0", 5, (0.03, 0.0, 0.0, 1.0)), ("xmax", 4, (0.0, 0.0, 0.0, 1.0))]:
        sc.boundaries.append(scenario.Boundary(kind, p or (0.0, 0.0, 0.0, 1.0), scenario.face_box(ext, xl, yl, zl)))
    mask = pipe_mask(xl, yl, zl, 3.2)
    mask[3:6, 3:6, 6:9] = 0
    ab, aa = pair(q, (xl, yl, zl), 0.7, storage, sc, mask=mask)
    for n in (3, 40):
        ab.step(n)
        aa.step(n)
        same_fields(ab, aa, (xl, yl, zl), (0, 0, 0), f"channel q={q} {storage}")


def test_reference_step_scenario():
    sc = scenario.parse_scenario_file(GOLDEN / "step.xml")
    dims = (sc.xl, sc.yl, sc.zl)
    ab, aa = pair(19, dims, 0.6, "f64", sc)
    ab.step(25)
    aa.step(25)
    same_fields(ab, aa, dims, (0, 0, 0), "step.xml")


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_periodic_and_mixed_boxes(storage):
    dims = (16, 12, 10)
    ab, aa = pair(19, dims, 0.8, storage, periodic=(1, 1, 1))
    rho, u = tgv_fields(*dims, u0=0.03)
    for d in (ab, aa):
        d.init_equilibrium(rho, u)
        d.step(11)
    same_fields(ab, aa, dims, (1, 1, 1), "periodic TGV")
    xl, yl, zl = 7, 5, 11
    sc = scenario.Scenario("couette", xl, yl, zl)
    sc.boundaries.append(scenario.Boundary(1, (0, 0, 0, 1.0), scenario.face_box("z0", xl, yl, zl)))
    sc.boundaries.append(scenario.Boundary(2, (0.04, 0.01, 0, 1.0), scenario.face_box("zmax", xl, yl, zl)))
    ab, aa = pair(27, (xl, yl, zl), 0.9, storage, sc, periodic=(1, 1, 0))
    ab.step(33)
    aa.step(33)
    same_fields(ab, aa, (xl, yl, zl), (1, 1, 0), "couette")


def test_brinkman_channel():
    xl, yl, zl = 12, 9, 7
    alpha = np.random.default_rng(4).uniform(0, 1, xl * yl * zl)
    sc = scenario.Scenario("porous", xl, yl, zl)
    for ext in ("y0", "ymax", "z0", "zmax"):
        sc.boundaries.append(scenario.Boundary(1, (0, 0, 0, 1.0), scenario.face_box(ext, xl, yl, zl)))
    ab, aa = pair(19, (xl, yl, zl), 0.7, "f32", sc, periodic=(1, 0, 0), brinkman=(alpha, (2e-5, 0, 0)))
    ab.step(21)
    aa.step(21)
    same_fields(ab, aa, (xl, yl, zl), (1, 0, 0), "brinkman")


def test_reads_and_writes_between_steps():
    """Reading the field (which brings it back to the natural representation),
    writing it with set_pdf and initialising it between steps give the same
    run as the two-buffer layout."""
    dims = (10, 9, 8)
    ab, aa = pair(19, dims, 0.65, "f64", scenario.cavity(*dims))
    for d in (ab, aa):
        d.step(3)
        d.moments()
        d.step(4)
        d.conservation()
        d.step(1)
    same_fields(ab, aa, dims, (0, 0, 0), "after reads")
    f = ab.get_pdf()
    f[1:-1, 1:-1, 1:-1] *= 1.0 + 1e-3 * np.random.default_rng(2).standard_normal(f[1:-1, 1:-1, 1:-1].shape)
    for d in (ab, aa):
        d.set_pdf(f)
        d.step(5)
    same_fields(ab, aa, dims, (0, 0, 0), "after set_pdf")


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_16bit_matches_two_buffer_layout(storage):
    """Boundary values pass through the storage format in both layouts (as
    the reference's boundary pass stores them in the field), so 16-bit
    storage agrees bit for bit too, after odd and even step counts."""
    dims = (16, 16, 16)
    ab, aa = pair(19, dims, 0.6, storage, scenario.cavity(*dims, lid=(0.05, 0.02, 0.0)))
    done = 0
    for n in (1, 2, 7, 37):
        ab.step(n - done)
        aa.step(n - done)
        done = n
        same_fields(ab, aa, dims, (0, 0, 0), f"cavity {storage} after {done} steps")
    say(f"cavity {storage}: identical after 1, 2, 7, 37 steps")


def test_one_buffer_and_refusals():
    dims = (40, 40, 40)
    ab, aa = pair(19, dims, 0.6, "f32", scenario.cavity(*dims))
    ia, ib = ab.info(), aa.info()
    assert ib["device_bytes"] == ia["device_bytes"] - ia["field_bytes"]
    with pytest.raises(capi.LbmError):
        aa.set_tuning(2, 0)
    with pytest.raises(capi.LbmError):
        capi.Tape(aa, 10)
    with pytest.raises(capi.LbmError):
        capi.Batch([aa])
    fs = capi.Domain(8, 8, 8, 0.6, q=19, storage="f32", layout="aa")
    scenario.cavity(8, 8, 8).apply(fs)
    fs.set_boundary_box(3, (0, 0, 0, 1.0), *scenario.face_box("y0", 8, 8, 8))
    with pytest.raises(capi.LbmError):
        fs.finalize()
