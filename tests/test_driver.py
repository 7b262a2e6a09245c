"""The drop-in `lbm` executable (lbm_amd/apps/lbm_main.cpp over include/lbm_hip/*.hpp):
config + scenario parsing on the CPU, and on the GPU full runs whose .vts
output is compared with the CPU oracle on the same scenario."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from lbm_amd import scenario

ROOT = Path(__file__).resolve().parents[1]
LBM = ROOT / "lbm_amd" / "lbm"
GOLDEN = ROOT / "tests" / "golden" / "scenarios"


def write_cfg(tmp: Path, scen: Path, steps: int, plot: int, extra: str = "") -> Path:
    p = tmp / "config.cfg"
    p.write_text(f"# test run\ncollision-model = bgk\ntau = 0.6\ntimesteps = {steps}\n"
                 f"timesteps-per-plot = {plot}\noutput-dir = {tmp / 'vtk'}\nscenario-file = {scen}\n{extra}")
    return p


def run(args, cwd):
    return subprocess.run([str(LBM), *map(str, args)], cwd=cwd, capture_output=True, text=True, timeout=600)


@pytest.mark.parametrize("name", ["cavity.xml", "step.xml", "shearflow.xml"])
def test_dry_run_lists_boundaries_in_file_order(tmp_path, name):
    cfg = write_cfg(tmp_path, GOLDEN / name, 5, 0)
    r = run(["--dry-run", cfg], tmp_path)
    assert r.returncode == 0, r.stderr
    got = re.findall(r"^boundary (.+) (\w+)$", r.stdout, flags=re.M)
    sc = scenario.parse_scenario_file(GOLDEN / name)
    assert got == [(b.extent, b.condition) for b in sc.boundaries]


def test_config_errors(tmp_path):
    cfg = write_cfg(tmp_path, GOLDEN / "cavity.xml", 5, 0)
    r = run(["--dry-run", cfg, "--tau", "0.4"], tmp_path)
    assert r.returncode == 1 and "An error occured: tau" in r.stderr
    (tmp_path / "bad.cfg").write_text("tau = 0.6\ntimesteps = 3\n")
    r = run(["--dry-run", tmp_path / "bad.cfg"], tmp_path)
    assert r.returncode == 1 and "required but missing" in r.stderr
    r = run([], tmp_path)
    assert r.returncode == 1 and "Usage" in r.stdout


def read_vts(path: Path):
    text = path.read_text()
    arrays = re.findall(r'<DataArray[^>]*Name="(\w+)"[^>]*>(.*?)</DataArray>', text, flags=re.S)
    out = {name: np.array(body.split(), dtype=np.float64) for name, body in arrays}
    return out["Density"], out["Velocity"].reshape(-1, 3)


def oracle_moments(sc, steps, q=19, tau=0.6):
    from oracle import ora
    o = ora.OracleDomain(sc.xl, sc.yl, sc.zl, tau, q=q)
    sc.apply(o)
    o.finalize()
    o.step(steps)
    rho, u = o.moments()
    fluid = o.kinds()[1:-1, 1:-1, 1:-1] == 0
    return rho.reshape(-1), u.reshape(-1, 3), fluid.reshape(-1)


@pytest.mark.gpu
@pytest.mark.parametrize("name,lattice,q", [("cavity16", "D3Q19", 19), ("step.xml", "D3Q19", 19),
                                            ("shearflow.xml", "D3Q27", 27), ("step.xml", "D3Q15", 15)])
def test_driver_matches_oracle(tmp_path, name, lattice, q):
    print(f"driver {name} {lattice}", flush=True)
    if name == "cavity16":
        scen = tmp_path / "cavity16.xml"
        scen.write_text((GOLDEN / "cavity.xml").read_text().replace('xl="40" yl="40" zl="40"',
                                                                     'xl="16" yl="12" zl="14"'))
    else:
        scen = GOLDEN / name
    steps = 30
    cfg = write_cfg(tmp_path, scen, steps, 15, f"lattice = {lattice}\n")
    r = run([cfg], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "MLUPS:" in r.stdout and "Finished!" in r.stdout
    sc = scenario.parse_scenario_file(scen)
    rho, u = read_vts(tmp_path / "vtk" / f"{sc.name}.{steps}.vts")
    ro, uo, fluid = oracle_moments(sc, steps, q=q)
    assert np.array_equal(rho[fluid], ro[fluid])
    assert np.array_equal(u[fluid], uo[fluid])
    assert (tmp_path / "vtk" / f"{sc.name}.15.vts").exists()


@pytest.mark.gpu
def test_driver_inplace_layout(tmp_path):
    """`layout = aa` (one buffer updated in place) gives the oracle's bits too."""
    print("driver step.xml layout aa", flush=True)
    scen = GOLDEN / "step.xml"
    steps = 17
    cfg = write_cfg(tmp_path, scen, steps, steps, "layout = aa\n")
    r = run([cfg], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    sc = scenario.parse_scenario_file(scen)
    rho, u = read_vts(tmp_path / "vtk" / f"{sc.name}.{steps}.vts")
    ro, uo, fluid = oracle_moments(sc, steps, q=19)
    assert np.array_equal(rho[fluid], ro[fluid])
    assert np.array_equal(u[fluid], uo[fluid])


@pytest.mark.gpu
def test_driver_vtk_mask_scenario(tmp_path):
    """A pipe through a fluid mask (legacy STRUCTURED_POINTS file) with inflow
    and outflow, as build/scenarios/pipe.xml sets up."""
    print("driver mask pipe", flush=True)
    xl, yl, zl = 20, 9, 9
    z, y, x = np.meshgrid(np.arange(zl), np.arange(yl), np.arange(xl), indexing="ij")
    mask = (((y - 4) ** 2 + (z - 4) ** 2) <= 12).astype(np.uint8)
    lines = ["# vtk DataFile Version 2.0", "pipe mask", "ASCII", "DATASET STRUCTURED_POINTS",
             f"DIMENSIONS {xl} {yl} {zl}", "ORIGIN 0 0 0", "SPACING 1 1 1", f"POINT_DATA {mask.size}",
             "SCALARS inputfluidMask unsigned_char", "LOOKUP_TABLE default"]
    lines += [" ".join(map(str, mask.reshape(-1)[i:i + 20])) for i in range(0, mask.size, 20)]
    (tmp_path / "pipe.vtk").write_text("\n".join(lines) + "\n")
    scen = tmp_path / "pipe.xml"
    scen.write_text(f"""<?xml version="1.0" ?>
<scenario name="Pipe flow">
   <domain vtk-file="{tmp_path / 'pipe.vtk'}">
        <boundary extent="z0" condition="noslip" />
        <boundary extent="zmax" condition="noslip" />
        <boundary extent="x0" condition="inflow" vx="0.03" vy="0" vz="0" />
        <boundary extent="xmax" condition="outflow" />
        <boundary extent="y0" condition="noslip" />
        <boundary extent="ymax" condition="noslip" />
    </domain>
</scenario>
""")
    steps = 40
    cfg = write_cfg(tmp_path, scen, steps, steps)
    r = run([cfg], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    sc = scenario.parse_scenario_file(scen)
    rho, u = read_vts(tmp_path / "vtk" / f"Pipe flow.{steps}.vts")
    ro, uo, fluid = oracle_moments(sc, steps)
    assert fluid.sum() < fluid.size  # the mask made solids
    assert np.array_equal(rho[fluid], ro[fluid]) and np.array_equal(u[fluid], uo[fluid])


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["ab", "aa"])
def test_driver_reference_pipe_block(tmp_path, layout):
    """The reference's pipe scenario (build/scenarios/pipe.xml:4-11: noslip
    z0/zmax/y0/ymax, inflow vx = 0.03 at x0, outflow at xmax) on the committed
    block of its own mask build/scenarios/pipe.vtk (x = 80 .. 119 around the
    constriction, tests/golden/pipe_crop.vtk.gz): the driver's .vts output
    equals the fp64 oracle bit for bit on every fluid cell."""
    import gzip
    print(f"driver reference pipe block, layout {layout}", flush=True)
    vtk = tmp_path / "pipe_crop.vtk"
    vtk.write_bytes(gzip.decompress((ROOT / "tests" / "golden" / "pipe_crop.vtk.gz").read_bytes()))
    scen = tmp_path / "pipe.xml"
    scen.write_text(f"""<?xml version="1.0" ?>
<scenario name="Pipe flow">
   <domain vtk-file="{vtk}">
        <boundary extent="z0" condition="noslip" />
        <boundary extent="zmax"    condition="noslip" />
        <boundary extent="x0"   condition="inflow" vx="0.03" vy="0" vz="0" />
        <boundary extent="xmax"  condition="outflow" />
        <boundary extent="y0"  condition="noslip" />
        <boundary extent="ymax"   condition="noslip" />
    </domain>
</scenario>
""")
    steps = 60
    cfg = write_cfg(tmp_path, scen, steps, steps, f"layout = {layout}\n")
    r = run([cfg], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    sc = scenario.parse_scenario_file(scen)
    assert (sc.xl, sc.yl, sc.zl) == (40, 54, 54) and int(sc.fluid_mask.sum()) == 29649
    rho, u = read_vts(tmp_path / "vtk" / f"Pipe flow.{steps}.vts")
    ro, uo, fluid = oracle_moments(sc, steps)
    assert fluid.sum() == 29649
    assert np.abs(uo[fluid, 0]).max() > 0.01  # inflow reached through the constriction's inlet side
    assert np.array_equal(rho[fluid], ro[fluid]) and np.array_equal(u[fluid], uo[fluid])


REF_MAIN = Path("/root/reference/src/main.cpp")


@pytest.mark.skipif(not REF_MAIN.exists(), reason="reference sources not present on this machine")
@pytest.mark.parametrize("lattice", [15, 19, 27])
def test_reference_driver_compiles_against_compat_headers(tmp_path, lattice):
    """The reference's own driver (src/main.cpp), unmodified, compiles and links
    against include/lbm_hip/compat (headers with the reference's names) and
    liblbm_hip.so, for each lattice it selects with -DD3Q; its command-line and
    error paths then behave as the reference's (usage, "An error occured: ...",
    EXIT_FAILURE) before any device work."""
    exe = tmp_path / "lbm_ref_driver"
    cmd = ["g++", "-std=c++17", "-O1", "-fopenmp", f"-DD3Q={lattice}",
           f"-I{ROOT / 'include' / 'lbm_hip' / 'compat'}", f"-I{ROOT / 'include'}", str(REF_MAIN),
           f"-L{ROOT / 'lbm_amd'}", "-llbm_hip", f"-Wl,-rpath,{ROOT / 'lbm_amd'}", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    h = subprocess.run([str(exe), "--help"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert h.returncode == 1 and "Usage" in h.stdout  # the reference exits with 1 after the usage text
    cfg = write_cfg(tmp_path, tmp_path / "missing.xml", 5, 0)
    e = subprocess.run([str(exe), str(cfg)], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert e.returncode == 1
    assert "An error occured: XML file" in e.stderr
