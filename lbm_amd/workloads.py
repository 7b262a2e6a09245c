"""Seeded synthetic workloads of the benchmark and the tests.

- taylor_green: 2-D Taylor-Green vortex in x-y (uniform in z) with a seeded
  density perturbation, on a periodic box.
- sphere_channel: channel along x with inflow at x0, outflow at xmax,
  no-slip walls on the other faces and seeded random spherical obstacles
  (radius 6-12 cells) added until the solid fraction first reaches a target.
- cavity_set: lid-driven cavities with edge lengths log-uniform in [16, 96]
  per axis and Reynolds numbers uniform in [10, 1000], realised through the
  relaxation time tau = 3 U_lid L_x / Re + 1/2 with U_lid = 0.05.
- porosity_pair: the inverse problem's porosity fields, uniform [0, 1] on a
  coarse 32^3 grid (seed s for the control, s + 1 for the target), upsampled
  trilinearly (periodic) to the fine grid.
"""
from __future__ import annotations

import numpy as np

from . import scenario


def taylor_green(xl, yl, zl, u0=0.02, k=None, rho_noise=1e-4, seed=12345):
    """(rho (zl,yl,xl), u (zl,yl,xl,3)); cell x at coordinate x - 1."""
    if k is None:
        k = 2 * np.pi / xl
    x = np.arange(xl)[None, None, :]
    y = np.arange(yl)[None, :, None]
    u = np.zeros((zl, yl, xl, 3))
    u[..., 0] = u0 * np.sin(k * x) * np.cos(k * y)
    u[..., 1] = -u0 * np.cos(k * x) * np.sin(k * y)
    rng = np.random.default_rng(seed)
    rho = 1.0 + rho_noise * rng.standard_normal((zl, yl, xl))
    return rho, u


def sphere_mask(xl, yl, zl, solid_fraction=0.2, rmin=6.0, rmax=12.0, seed=12345, margin=None):
    """Fluid mask (zl, yl, xl), 1 = fluid: spheres with uniform centres and
    radii in [rmin, rmax], added until the solid fraction first reaches
    solid_fraction. `margin` keeps centres away from the inflow/outflow faces."""
    rng = np.random.default_rng(seed)
    solid = np.zeros((zl, yl, xl), dtype=bool)
    ntot = solid.size
    nsolid = 0  # kept equal to solid.sum() as spheres are added
    m = rmax if margin is None else margin
    while nsolid / ntot < solid_fraction:
        cx = rng.uniform(1 + m, xl - m)
        cy, cz = rng.uniform(1, yl), rng.uniform(1, zl)
        r = rng.uniform(rmin, rmax)
        lo = [max(1, int(np.floor(c - r))) for c in (cz, cy, cx)]
        hi = [min(n, int(np.ceil(c + r))) for c, n in ((cz, zl), (cy, yl), (cx, xl))]
        sl = tuple(slice(a - 1, b) for a, b in zip(lo, hi))
        z, y, x = np.meshgrid(np.arange(lo[0], hi[0] + 1), np.arange(lo[1], hi[1] + 1),
                              np.arange(lo[2], hi[2] + 1), indexing="ij")
        before = int(solid[sl].sum())
        solid[sl] |= (x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 <= r * r
        nsolid += int(solid[sl].sum()) - before
    return (~solid).astype(np.uint8)


def sphere_channel(xl, yl, zl, u_in=0.04, solid_fraction=0.2, seed=12345):
    """Scenario + mask of the obstacle channel."""
    sc = scenario.Scenario("Sphere channel", xl, yl, zl)
    for ext, kind, p in [("z0", 1, None), ("zmax", 1, None), ("y0", 1, None), ("ymax", 1, None),
                         ("x0", 5, (u_in, 0.0, 0.0, 1.0)), ("xmax", 4, (0.0, 0.0, 0.0, 1.0))]:
        sc.boundaries.append(scenario.Boundary(kind, p or (0.0, 0.0, 0.0, 1.0),
                                               scenario.face_box(ext, xl, yl, zl), "", ext))
    sc.fluid_mask = sphere_mask(xl, yl, zl, solid_fraction, seed=seed)
    return sc


U_LID = 0.05


def cavity_set(n=512, lmin=16, lmax=96, re_min=10.0, re_max=1000.0, seed=12345):
    """List of (xl, yl, zl, tau, Re): log-uniform edges, uniform Re."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        xl, yl, zl = (int(round(np.exp(rng.uniform(np.log(lmin), np.log(lmax))))) for _ in range(3))
        re = rng.uniform(re_min, re_max)
        tau = 3.0 * U_LID * xl / re + 0.5
        out.append((xl, yl, zl, tau, re))
    return out


def porosity_pair(n=128, nc=32, seed=2024):
    """(coarse control, fine control, fine target) porosity fields."""
    from .autograd import upsample_np
    ac = np.random.default_rng(seed).uniform(0.0, 1.0, (nc, nc, nc))
    at = np.random.default_rng(seed + 1).uniform(0.0, 1.0, (nc, nc, nc))
    return ac, upsample_np(ac, (n, n, n)), upsample_np(at, (n, n, n))
