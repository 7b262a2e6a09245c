"""Scenario XML, config.cfg and legacy-VTK mask readers (host side).

Same file formats and the same semantics as the reference's front-end:
  - scenario XML: include/io/scenario.h:21-188 (grammar in
    build/scenarios/README.xml): <scenario name><domain xl yl zl | vtk-file>
    <boundary extent condition .../>*, boundaries applied in file order;
  - config file: include/io/configuration.h:68-145 (`key = value` lines,
    `#` comments, command-line values win);
  - fluid mask: the legacy STRUCTURED_POINTS file read at
    include/io/vtk.hpp:94-157 (mask == 0 -> NoSlip, interior, x fastest).
The parsers are written with the Python standard library (the reference uses
pugixml, boost.program_options and VTK, which are not part of this build).
"""
from __future__ import annotations

import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

KINDS = {"noslip": 1, "movingwall": 2, "freeslip": 3, "outflow": 4, "inflow": 5, "pressure": 6}


@dataclass
class Boundary:
    kind: int
    params: tuple  # vx, vy, vz, rho
    box: tuple     # x0, xE, y0, yE, z0, zE (inclusive, ghosted)
    condition: str = ""
    extent: str = ""


@dataclass
class Scenario:
    name: str
    xl: int
    yl: int
    zl: int
    boundaries: list = field(default_factory=list)
    fluid_mask: np.ndarray | None = None  # (zl, yl, xl) uint8, 1 = fluid
    origin: tuple = (0.0, 0.0, 0.0)
    spacing: tuple = (1.0, 1.0, 1.0)

    def apply(self, dom) -> None:
        """Mask first (the VTK reader runs before the boundary loop,
        include/io/scenario.h:160-186), then boxes in file order."""
        if self.fluid_mask is not None:
            dom.set_solid_mask(self.fluid_mask)
        for b in self.boundaries:
            dom.set_boundary_box(b.kind, b.params, *b.box)


def face_box(extent: str, xl: int, yl: int, zl: int):
    """Face keywords -> full ghosted face boxes (include/io/scenario.h:99-116)."""
    faces = {
        "z0": (0, xl + 1, 0, yl + 1, 0, 0),
        "zmax": (0, xl + 1, 0, yl + 1, zl + 1, zl + 1),
        "x0": (0, 0, 0, yl + 1, 0, zl + 1),
        "xmax": (xl + 1, xl + 1, 0, yl + 1, 0, zl + 1),
        "y0": (0, xl + 1, 0, 0, 0, zl + 1),
        "ymax": (0, xl + 1, yl + 1, yl + 1, 0, zl + 1),
    }
    if extent in faces:
        return faces[extent]
    toks = extent.split()
    if len(toks) < 6:  # include/io/scenario.h:124-125
        raise ValueError(f'Extent "{extent}" is not complete! Must be six values!')
    vals = tuple(int(t) for t in toks[:6])
    if any(v < 0 for v in vals):
        raise ValueError(f'Extent "{extent}" has a negative value')
    return vals


def _attr(node, name: str) -> str:
    v = node.get(name)
    if v is None:  # include/io/scenario.h:21-26
        raise ValueError(f'Missing attribute "{name}" for node "{node.tag}"!')
    return v


def parse_condition(node):
    """Condition attribute -> (kind, params) (include/io/scenario.h:28-88)."""
    cond = _attr(node, "condition")
    if cond in ("noslip", "freeslip"):
        return KINDS[cond], (0.0, 0.0, 0.0, 1.0)
    if cond == "movingwall":
        v = tuple(float(_attr(node, k)) for k in ("vx", "vy", "vz"))
        return KINDS[cond], v + (1.0,)
    if cond == "outflow":
        rho = float(node.get("rho-ref", 1.0))
        return KINDS[cond], (0.0, 0.0, 0.0, rho)
    if cond == "inflow":
        v = tuple(float(_attr(node, k)) for k in ("vx", "vy", "vz"))
        rho = float(node.get("rho-ref", 1.0))
        return KINDS[cond], v + (rho,)
    if cond == "pressure":
        return KINDS[cond], (0.0, 0.0, 0.0, float(_attr(node, "rho-in")))
    raise ValueError(f"{cond} boundary condition not supported!")


def read_vtk_mask(path: str | Path):
    """Legacy ASCII STRUCTURED_POINTS fluid mask (include/io/vtk.hpp:94-157).

    Returns (mask (zl, yl, xl) uint8, origin, spacing). The reference takes
    DIMENSIONS as the interior lengths and reads one value per interior cell
    in x-fastest order.
    """
    text = Path(path).read_text().split("\n")
    dims = origin = spacing = None
    values: list[str] = []
    in_data = False
    for line in text:
        t = line.strip()
        if not t:
            continue
        up = t.upper()
        if in_data:
            values.extend(t.split())
            continue
        if up.startswith("DIMENSIONS"):
            dims = tuple(int(v) for v in t.split()[1:4])
        elif up.startswith("ORIGIN"):
            origin = tuple(float(v) for v in t.split()[1:4])
        elif up.startswith("SPACING") or up.startswith("ASPECT_RATIO"):
            spacing = tuple(float(v) for v in t.split()[1:4])
        elif up.startswith("LOOKUP_TABLE"):
            in_data = True
        elif up.startswith("DATASET") and "STRUCTURED_POINTS" not in up:
            raise ValueError(f"VTK file {path} is not a structured points file")
    if dims is None:
        raise ValueError(f"VTK file {path} has no DIMENSIONS")
    xl, yl, zl = dims
    n = xl * yl * zl
    if len(values) < n:
        raise ValueError(f"VTK file {path}: {len(values)} values for {n} points")
    mask = (np.asarray(values[:n], dtype=np.int64) != 0).astype(np.uint8).reshape(zl, yl, xl)
    return mask, origin or (0.0, 0.0, 0.0), spacing or (1.0, 1.0, 1.0)


def parse_scenario_file(path: str | Path, base_dir: str | Path | None = None) -> Scenario:
    """Scenario XML -> Scenario (include/io/scenario.h:131-188).

    A relative vtk-file path is resolved against base_dir (the reference
    resolves it against the working directory)."""
    try:
        root = ET.parse(str(path)).getroot()
    except (ET.ParseError, OSError) as e:
        raise ValueError(f'XML file "{path}" could not be read properly!') from e
    if root.tag != "scenario":
        raise ValueError("Scenario node missing!")
    name = root.get("name")
    if name is None:
        raise ValueError("Scenario name is missing!")
    dom = root.find("domain")
    if dom is None:
        raise ValueError("Domain node is missing!")
    mask = None
    origin, spacing = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    if dom.get("vtk-file") is not None:
        vp = Path(dom.get("vtk-file"))
        if not vp.is_absolute() and base_dir is not None:
            vp = Path(base_dir) / vp
        mask, origin, spacing = read_vtk_mask(vp)
        zl, yl, xl = mask.shape
    elif all(dom.get(k) is not None for k in ("xl", "yl", "zl")):
        xl, yl, zl = (int(dom.get(k)) for k in ("xl", "yl", "zl"))
        if min(xl, yl, zl) <= 0:
            raise ValueError("domain lengths must be positive")
    else:
        raise ValueError("Neither vtk-file nor xl/yl/zl attribute provided to domain node!")
    sc = Scenario(name, xl, yl, zl, fluid_mask=mask, origin=origin, spacing=spacing)
    for b in dom.findall("boundary"):
        extent = _attr(b, "extent")
        kind, params = parse_condition(b)
        box = face_box(extent, xl, yl, zl)
        sc.boundaries.append(Boundary(kind, params, box, b.get("condition"), extent))
    return sc


def cavity(xl: int, yl: int, zl: int, lid=(0.05, 0.0, 0.0)) -> Scenario:
    """build/scenarios/cavity.xml at any size: noslip z0, movingwall zmax, then
    noslip x0, xmax, y0, ymax (which override the lid's edges)."""
    sc = Scenario("Cavity", xl, yl, zl)
    order = [("z0", 1, (0.0, 0.0, 0.0, 1.0)), ("zmax", 2, tuple(lid) + (1.0,)),
             ("x0", 1, (0.0, 0.0, 0.0, 1.0)), ("xmax", 1, (0.0, 0.0, 0.0, 1.0)),
             ("y0", 1, (0.0, 0.0, 0.0, 1.0)), ("ymax", 1, (0.0, 0.0, 0.0, 1.0))]
    for ext, k, p in order:
        sc.boundaries.append(Boundary(k, p, face_box(ext, xl, yl, zl), "", ext))
    return sc


def parse_config_file(path: str | Path, overrides: dict | None = None) -> dict:
    """config.cfg -> dict (include/io/configuration.h:68-145). Keys of the
    "Configuration options" group; command-line overrides win."""
    allowed = {"collision-model", "tau", "timesteps", "timesteps-per-plot", "scenario-file",
               "omp-threads", "output-dir"}
    cfg: dict = {"collision-model": "bgk", "omp-threads": 1, "output-dir": "vtk"}
    for raw in Path(path).read_text().splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        if "=" not in line:
            raise ValueError(f"bad config line: {raw!r}")
        k, v = (s.strip() for s in line.split("=", 1))
        if k not in allowed:
            raise ValueError(f"unrecognised option '{k}'")
        cfg[k] = v
    for k, v in (overrides or {}).items():
        cfg[k] = v
    for k in ("tau", "timesteps", "timesteps-per-plot", "scenario-file"):
        if k not in cfg:
            raise ValueError(f"the option '--{k}' is required but missing")
    cfg["tau"] = float(cfg["tau"])
    cfg["timesteps"] = int(cfg["timesteps"])
    cfg["timesteps-per-plot"] = int(cfg["timesteps-per-plot"])
    cfg["omp-threads"] = int(cfg["omp-threads"])
    return cfg
