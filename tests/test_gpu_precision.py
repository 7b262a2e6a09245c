"""Precision of the shifted storage modes against the fp64 path.

Obstacle channel: D3Q27, 384 x 192 x 192 cells, inflow 0.04 along x at x = 0, outflow at
x = max, half-way bounce-back on the four side faces, seeded random spheres
(radius 6-12 cells) up to a solid fraction of 0.2, tau = 0.55, 5000 steps.
Each storage mode (shifted f - w in fp32, fp16 and bf16, all with fp32
arithmetic) is compared with the fp64 path in the reference's operation
order (which matches the CPU oracle bit for bit on the smaller parity cases)
on the same geometry: relative L2 and max error of the velocity and of the
density deviation rho - 1 over the fluid cells after the last step.

Taylor-Green box (the benchmark workload): periodic 256^3 D3Q19, tau = 0.8,
u0 = 0.02, k = 2 pi / 256, density noise 1e-4, 500 steps, same comparison
over all cells.
"""
import sys

import numpy as np
import pytest

from lbm_amd import capi, workloads

pytestmark = pytest.mark.gpu

XL, YL, ZL, TAU, STEPS, U_IN = 384, 192, 192, 0.55, 5000, 0.04
# Regression guards, set above the measured errors: they catch a broken
# mode, they are not the precision target. fp16 on the channel measures
# about 1.4e-3 (its velocities reach 0.12) and misses the 1e-3 target,
# which test_channel_f16_meets_target asserts as an expected failure; bf16
# (8-bit mantissa) gets a loose bound.
U_TOL = {"f32": 1e-5, "f16": 2e-3, "bf16": 3e-2}
U_TARGET_F16 = 1e-3  # velocity relative L2 target for 16-bit storage
U_TOL_TGV = {"f32": 1e-6, "f16": 1e-3, "bf16": 2e-2}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(storage, sc):
    d = capi.Domain(XL, YL, ZL, TAU, q=27, storage=storage)
    sc.apply(d)
    d.finalize()
    d.step(STEPS)
    rho, u = d.moments()
    d.close()
    return rho, u


def errors(a, b):
    diff = a - b
    return (float(np.linalg.norm(diff) / np.linalg.norm(b)),
            float(np.abs(diff).max() / np.abs(b).max()))


@pytest.fixture(scope="module")
def channel():
    sc = workloads.sphere_channel(XL, YL, ZL, u_in=U_IN, solid_fraction=0.2, seed=12345)
    fluid = sc.fluid_mask.astype(bool)
    say(f"C3 channel {XL}x{YL}x{ZL} D3Q27, solid fraction {1 - fluid.mean():.3f}, fp64 reference run")
    rho, u = run("f64", sc)
    assert np.isfinite(u[fluid]).all()
    say(f"  fp64: max |u| {np.abs(u[fluid]).max():.4f}, rho in [{rho[fluid].min():.4f}, {rho[fluid].max():.4f}]")
    return sc, fluid, rho, u


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_channel_precision_vs_fp64(channel, storage):
    sc, fluid, rho64, u64 = channel
    rho, u = run(storage, sc)
    eu = errors(u[fluid], u64[fluid])
    er = errors(rho[fluid] - 1.0, rho64[fluid] - 1.0)
    say(f"  {storage}: velocity rel L2 {eu[0]:.3e} max {eu[1]:.3e}; "
        f"density deviation rel L2 {er[0]:.3e} max {er[1]:.3e}")
    assert np.isfinite(u[fluid]).all()
    assert eu[0] <= U_TOL[storage]


@pytest.mark.xfail(reason="open: fp16 storage on the D3Q27 obstacle channel measures a velocity relative L2 "
                          "of about 1.4e-3 against fp64, above the 1e-3 target (the Taylor-Green box meets it)",
                   strict=False)
def test_channel_f16_meets_target(channel):
    sc, fluid, rho64, u64 = channel
    rho, u = run("f16", sc)
    eu = errors(u[fluid], u64[fluid])
    say(f"  f16 against the 1e-3 target: velocity rel L2 {eu[0]:.3e}")
    assert eu[0] <= U_TARGET_F16


@pytest.fixture(scope="module")
def tgv():
    from tests.lbm_math import tgv_fields
    n = 256
    rho0, u0 = tgv_fields(n, n, n, u0=0.02)
    say("TGV 256^3 D3Q19 fp64 reference run")
    return rho0, u0, tgv_run("f64", rho0, u0)


def tgv_run(storage, rho0, u0, n=256, steps=500):
    d = capi.Domain(n, n, n, 0.8, q=19, storage=storage, periodic=(1, 1, 1))
    d.finalize()
    d.init_equilibrium(rho0, u0)
    d.step(steps)
    out = d.moments()
    d.close()
    return out


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_tgv_256_precision_vs_fp64(tgv, storage):
    rho0, u0, (rho64, u64) = tgv
    rho, u = tgv_run(storage, rho0, u0)
    eu = errors(u, u64)
    er = errors(rho - 1.0, rho64 - 1.0)
    say(f"  TGV {storage}: velocity rel L2 {eu[0]:.3e} max {eu[1]:.3e}; "
        f"density deviation rel L2 {er[0]:.3e} max {er[1]:.3e}")
    assert eu[0] <= U_TOL_TGV[storage]
