"""Edge cases of the domain shape and size.

Degenerate shapes against the CPU oracle, bit for bit in fp64: a single
periodic cell (every direction streams onto itself), a single walled cell
(every link a boundary link), one-cell-thick slabs and boxes with one
periodic axis of length one; in both the two-buffer and the in-place
layout.

Maximum size: a 512^3 D3Q19 fp32 domain whose distribution buffers hold
more than 2^31 elements (11.6 GB each). The flow is a Taylor-Green vortex
uniform in z, so every z slice must stay identical bit for bit; an index
that wraps at 2^31 would corrupt the planes beyond it. Mass is conserved to
fp32 rounding.
"""
import sys

import numpy as np
import pytest

from lbm_amd import capi, scenario
from oracle import ora
from tests.lbm_math import tgv_fields
from tests.test_gpu_parity import compared_cells

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_pair(q, dims, tau, steps, periodic, walls, layout, init=None):
    g = capi.Domain(*dims, tau, q=q, storage="f64", periodic=periodic, layout=layout)
    o = ora.OracleDomain(*dims, tau, q=q, periodic=periodic)
    for d in (g, o):
        if walls:
            scenario.cavity(*dims, lid=(0.04, 0.01, 0.0)).apply(d)
        d.finalize()
    if init is not None:
        c, w = ora.lattice(q)
        f = o.get_pdf()
        rng = np.random.default_rng(q)
        f[1:-1, 1:-1, 1:-1] = np.array(w) * (1 + 0.02 * rng.standard_normal(f[1:-1, 1:-1, 1:-1].shape))
        g.set_pdf(f)
        o.set_pdf(f)
    g.step(steps)
    o.step(steps)
    fg, fo = g.get_pdf(), o.get_pdf()
    m = compared_cells(o)
    return fg[m], fo[m]


@pytest.mark.parametrize("layout", ["ab", "aa"])
@pytest.mark.parametrize("q", [15, 19, 27])
@pytest.mark.parametrize("case", ["cell-periodic", "cell-walled", "slab-walled", "thin-periodic-x"])
def test_degenerate_shapes_bit_exact(case, q, layout):
    dims, periodic, walls = {
        "cell-periodic": ((1, 1, 1), (1, 1, 1), False),
        "cell-walled": ((1, 1, 1), (0, 0, 0), True),
        "slab-walled": ((1, 6, 5), (0, 0, 0), True),
        "thin-periodic-x": ((1, 4, 7), (1, 0, 0), True),
    }[case]
    say(f"{case} q={q} {layout}")
    fg, fo = run_pair(q, dims, 0.7, 9, periodic, walls, layout, init=True)
    assert np.array_equal(fg, fo), f"{int((fg != fo).sum())} values differ"


def test_512_cube_fields_beyond_2_31_elements():
    n = 512
    say("512^3 fp32 TGV (z-uniform)")
    d = capi.Domain(n, n, n, 0.8, q=19, storage="f32", periodic=(1, 1, 1))
    d.finalize()
    info = d.info()
    elems = info["field_bytes"] // info["bytes_per_value"]
    say(f"  elements per buffer {elems} ({info['field_bytes'] / 1e9:.1f} GB)")
    assert elems > 2 ** 31
    rng = np.random.default_rng(0)
    rho2 = 1.0 + 1e-4 * rng.standard_normal((1, n, n))
    rho, u = tgv_fields(n, n, n, u0=0.03, rho_noise=0.0)
    rho = np.broadcast_to(rho2, (n, n, n))
    d.init_equilibrium(np.ascontiguousarray(rho), u)
    m0 = d.conservation()
    d.step(12)
    m1 = d.conservation()
    r, v = d.moments()
    say(f"  mass {m0[0]:.9e} -> {m1[0]:.9e}; max |u| {np.abs(v).max():.4f}")
    assert np.isfinite(v).all()
    assert abs(m1[0] / m0[0] - 1) < 1e-6
    # every z slice equals slice 0 exactly
    for z in (1, 255, 256, 257, 400, n - 1):
        assert np.array_equal(v[z], v[0]), f"slice {z} differs from slice 0"
        assert np.array_equal(r[z], r[0])
    d.close()
