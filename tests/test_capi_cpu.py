"""CPU checks of the drop-in boundary: the shared library loads without a GPU,
exports every entry point declared in include/*.h, and its host-only calls
behave; the scenario/config front-end parses the reference's formats."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from lbm_amd import capi, scenario

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(lbm_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(capi.LIB_PATH))
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    bound = {n for n, _, _ in capi.SIGNATURES}
    missing = [n for n in declared_symbols() if n not in bound]
    assert not missing, missing


def test_host_only_calls():
    lib = capi.load()
    assert lib.lbm_abi_version() == 5
    d = capi.Desc()
    assert lib.lbm_desc_init(ctypes.byref(d)) == 0
    assert d.q == 19 and d.tau == 1.0 and d.spacing[2] == 1.0
    # argument errors are reported before any device work
    h = ctypes.c_void_p()
    d.q = 17
    assert lib.lbm_domain_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert b"q must be" in lib.lbm_last_error()
    d.q, d.xl, d.yl, d.zl, d.tau = 19, 4, 4, 4, 0.5
    assert lib.lbm_domain_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert b"tau" in lib.lbm_last_error()
    assert lib.lbm_lattice(18, None, None, None) == -1


GOLDEN = ROOT / "tests" / "golden" / "scenarios"


def test_scenario_cavity_matches_builder():
    sc = scenario.parse_scenario_file(GOLDEN / "cavity.xml")
    ref = scenario.cavity(40, 40, 40)
    assert (sc.xl, sc.yl, sc.zl) == (40, 40, 40) and sc.name == "Cavity"
    assert [(b.kind, b.params, b.box) for b in sc.boundaries] == \
           [(b.kind, b.params, b.box) for b in ref.boundaries]


def test_scenario_step_custom_extents():
    sc = scenario.parse_scenario_file(GOLDEN / "step.xml")
    assert (sc.xl, sc.yl, sc.zl) == (30, 30, 50)
    b0 = sc.boundaries[0]
    assert b0.kind == scenario.KINDS["inflow"] and b0.box == (17, 31, 0, 31, 0, 0)
    assert b0.params == (0.0, 0.0, 0.05, 1.0)
    assert sc.boundaries[1].box == (0, 16, 0, 31, 0, 25)
    assert sc.boundaries[2].box == (0, 31, 0, 31, 51, 51)


def test_scenario_shearflow_conditions():
    sc = scenario.parse_scenario_file(GOLDEN / "shearflow.xml")
    kinds = [b.condition for b in sc.boundaries]
    assert kinds == ["pressure", "outflow", "freeslip", "freeslip", "noslip", "noslip"]
    assert sc.boundaries[0].params[3] == 1.005


def test_scenario_errors():
    with pytest.raises(ValueError, match="not complete"):
        scenario.face_box("1 2 3", 4, 4, 4)
    bad = GOLDEN / "bad_condition.xml"
    with pytest.raises(ValueError, match="not supported"):
        scenario.parse_scenario_file(bad)


def test_vtk_mask_reader(tmp_path):
    xl, yl, zl = 5, 3, 2
    mask = (np.arange(xl * yl * zl) % 3 != 0).astype(np.uint8)
    p = tmp_path / "m.vtk"
    lines = ["# vtk DataFile Version 2.0", "mask", "ASCII", "", "DATASET STRUCTURED_POINTS",
             f"DIMENSIONS {xl} {yl} {zl}", "", "ORIGIN 1.5 -2 0", "SPACING 0.5 1 2", "",
             f"POINT_DATA {xl * yl * zl}", "SCALARS inputfluidMask unsigned_char",
             "LOOKUP_TABLE default"]
    lines += [" ".join(str(v) for v in mask[i:i + 7]) for i in range(0, mask.size, 7)]
    p.write_text("\n".join(lines) + "\n")
    m, origin, spacing = scenario.read_vtk_mask(p)
    assert m.shape == (zl, yl, xl)
    assert np.array_equal(m.reshape(-1), mask)
    assert origin == (1.5, -2.0, 0.0) and spacing == (0.5, 1.0, 2.0)


def test_config_file(tmp_path):
    p = tmp_path / "config.cfg"
    p.write_text("# comment\ncollision-model = bgk\ntau = 0.6\ntimesteps = 1000\n"
                 "timesteps-per-plot = 1\noutput-dir = vtk\nscenario-file = scenarios/cavity.xml\n")
    cfg = scenario.parse_config_file(p, {"timesteps": "5"})
    assert cfg["tau"] == 0.6 and cfg["timesteps"] == 5 and cfg["omp-threads"] == 1
    p.write_text("tau = 0.6\n")
    with pytest.raises(ValueError, match="required"):
        scenario.parse_config_file(p)


def test_binding_checks_array_sizes():
    """Host arrays handed to the C ABI are checked for their size before the
    call (the library reads exactly that many values from the pointer)."""
    v = capi._f64(np.arange(6).reshape(2, 3), 6, "x")
    assert v.dtype == np.float64 and v.shape == (6,) and v.flags.c_contiguous
    with pytest.raises(ValueError, match="porosity direction: expected 8 values, got 7"):
        capi._f64(np.zeros(7), 8, "porosity direction")
