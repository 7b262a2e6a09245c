"""Two time steps per pass (k_bulk2) against single steps.

Where every interior cell is bulk (periodic along all three axes) and the
storage is fp32, fp16 or bf16, lbm_step runs pairs of steps as one pass that
holds the intermediate state on chip. The pass rounds the intermediate to the
storage format and repeats the single step's arithmetic, so every run below
must equal the same run in single steps bit for bit: shapes smaller than
the 64 x 8 tile and not multiples of it, axes of one cell, odd step counts
(pairs plus one single step), Brinkman porosity with a body force, and
several z chunk sizes. The 256^3 case is the benchmark size.
"""
import sys

import numpy as np
import pytest

from lbm_amd import capi, scenario
from oracle import ora
from tests.lbm_math import tgv_fields

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1), (2, 3, 5), (3, 1, 2), (65, 9, 17), (130, 17, 33), (64, 8, 4), (200, 40, 24)]
# the pass is built for every lattice in 4- and 2-byte storage (its LDS slots fit 64 KiB)
BUILT = {(q, s) for q in (15, 19, 27) for s in ("f32", "f16", "bf16")}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(shape, q, storage, tau, brink, f0, steps_per_pass, z_chunk=0):
    d = capi.Domain(*shape, tau, q=q, storage=storage, periodic=(1, 1, 1))
    if brink is not None:
        d.set_brinkman(brink, g=(1e-5, -2e-6, 3e-6))
    d.finalize()
    d.set_time_blocking(steps_per_pass, z_chunk)
    d.set_pdf(f0)
    return d


def initial(shape, q, seed):
    xl, yl, zl = shape
    _, w = ora.lattice(q)
    rng = np.random.default_rng(seed)
    f = np.zeros((zl + 2, yl + 2, xl + 2, q))
    f[1:-1, 1:-1, 1:-1] = np.array(w) * (1 + 0.05 * rng.standard_normal((zl, yl, xl, q)))
    return f


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("q", [15, 19, 27])
def test_two_step_pass_equals_single_steps(storage, q):
    say(f"two-step pass {storage} q={q}")
    for si, shape in enumerate(SHAPES):
        f0 = initial(shape, q, seed=100 * q + si)
        rng = np.random.default_rng(si)
        for brink in (None, rng.uniform(0.0, 1.0, size=shape[::-1])):
            single = make(shape, q, storage, 0.7, brink, f0, 1)
            assert single.info()["steps_per_pass"] == 1
            single.step(7)
            # interior cells: a periodic box's ghost cells are never read, and
            # which buffer's stale ghosts get_pdf shows depends on step parity
            ref = single.get_pdf()[1:-1, 1:-1, 1:-1]
            single.close()
            for zc in (0, 1, 3):
                d = make(shape, q, storage, 0.7, brink, f0, 2, zc)
                info = d.info()
                expect = 2 if (q, storage) in BUILT and (storage == "f32" or shape[0] > 32) else 1
                assert info["steps_per_pass"] == expect, (shape, info)
                d.step(7)
                got = d.get_pdf()[1:-1, 1:-1, 1:-1]
                d.close()
                same = got.view(np.uint64) == ref.view(np.uint64)
                assert same.all(), (f"{shape} brink={brink is not None} z_chunk={zc}: "
                                    f"{int((~same).sum())} values differ")


def test_pass_applies_to_fp32_fp16_bf16_two_buffer_domains():
    say("two-step pass applicability")
    per = capi.Domain(70, 9, 6, 0.8, q=19, storage="f32", periodic=(1, 1, 1))
    per.finalize()
    assert per.info()["steps_per_pass"] == 2
    per.set_time_blocking(1)
    assert per.info()["steps_per_pass"] == 1
    per.set_time_blocking(2, 5)
    assert per.info()["steps_per_pass"] == 2 and per.info()["z_chunk"] == 5
    with pytest.raises(capi.LbmError):
        per.set_time_blocking(3)
    # walled domains take the walled pass (tests/test_gpu_wall2.py)
    walled = capi.Domain(70, 9, 6, 0.8, q=19, storage="f32", periodic=(1, 1, 0))
    scenario.cavity(70, 9, 6).apply(walled)
    walled.finalize()
    assert walled.info()["steps_per_pass"] == 2
    short = capi.Domain(32, 9, 6, 0.8, q=19, storage="f16", periodic=(1, 1, 1))
    short.finalize()
    assert short.info()["steps_per_pass"] == 1  # 16-bit rows of 32 cells: single steps
    for storage in ("f64", "f64s"):
        d = capi.Domain(70, 9, 6, 0.8, q=19, storage=storage, periodic=(1, 1, 1))
        d.finalize()
        assert d.info()["steps_per_pass"] == 1
    aa = capi.Domain(70, 9, 6, 0.8, q=19, storage="f32", periodic=(1, 1, 1), layout="aa")
    aa.finalize()
    assert aa.info()["steps_per_pass"] == 1


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_two_step_pass_256_cube(storage):
    """Benchmark size (D3Q19 256^3, Taylor-Green vortex with density noise,
    tau = 0.8): 201 steps (100 passes and one single step) equal 201 single
    steps bit for bit. A race between the pass's LDS sends and reads would
    show here as a sporadic difference."""
    n = 256
    say(f"two-step pass 256^3 {storage}")
    rho, u = tgv_fields(n, n, n, u0=0.02)
    out = {}
    for spp in (1, 2):
        d = capi.Domain(n, n, n, 0.8, q=19, storage=storage, periodic=(1, 1, 1))
        d.finalize()
        d.set_time_blocking(spp)
        d.init_equilibrium(rho, u)
        d.step(201)
        r, v = d.moments()
        out[spp] = (r, v, d.conservation())
        d.close()
    assert np.array_equal(out[1][0].view(np.uint64), out[2][0].view(np.uint64))
    assert np.array_equal(out[1][1].view(np.uint64), out[2][1].view(np.uint64))
    assert np.array_equal(out[1][2], out[2][2])


def test_time_steps_accounts_every_launch():
    """lbm_time_steps records one event before the first launch and one after
    each launch: the bulk and near intervals add up to the total, a
    bulk-only box (two-step passes plus one single step) has no near time,
    and the row kernel with its near-boundary launch has some."""
    say("time_steps accounting")
    per = capi.Domain(70, 12, 9, 0.8, q=19, storage="f32", periodic=(1, 1, 1))
    per.finalize()
    t = per.time_steps(7)
    per.close()
    assert t["steps"] == 7 and t["near_ms"] == 0.0
    assert t["total_ms"] > 0 and abs(t["bulk_ms"] - t["total_ms"]) <= 1e-3 * t["total_ms"]
    cav = capi.Domain(70, 12, 9, 0.8, q=19, storage="f32")
    scenario.cavity(70, 12, 9).apply(cav)
    cav.finalize()
    cav.set_tuning(2, 0)  # row kernel + near-boundary kernel
    t = cav.time_steps(5)
    cav.close()
    assert t["steps"] == 5 and t["near_ms"] > 0 and t["bulk_ms"] > 0
    assert abs(t["bulk_ms"] + t["near_ms"] - t["total_ms"]) <= 1e-3 * t["total_ms"]
