"""Batched stepping (lbm_batch_*): a batch of uneven lid-driven cavities
with per-domain relaxation times gives, domain by domain, the same bits as
stepping each domain alone; the autotuner leaves the fields untouched; a
subsample is checked against the fp64 CPU oracle."""
import sys

import numpy as np
import pytest

from lbm_amd import capi, scenario, workloads
from oracle import ora

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cavity(xl, yl, zl, tau, storage):
    d = capi.Domain(xl, yl, zl, tau, q=19, storage=storage)
    scenario.cavity(xl, yl, zl, lid=(workloads.U_LID, 0.0, 0.0)).apply(d)
    d.finalize()
    return d


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


KERNELS = {"rows_separate": capi.BATCH_ROWS_SEPARATE, "rows_fused": capi.BATCH_ROWS_FUSED,
           "cells": capi.BATCH_CELLS}


# These cavities are short in x, so their rows are packed and the row kernels
# take one cell per thread, which 16-bit storage does not build (its row
# forms are covered on long rows by test_batch_rows_16bit_equals_single).
FORMS = [(s, c) for s in ("f32", "f16", "f64s") for c in ("rows_separate", "rows_fused", "cells", "tuned")
         if not (s == "f16" and c.startswith("rows"))]


@pytest.mark.parametrize("storage,config", FORMS)
def test_batch_equals_single(storage, config):
    """Every kernel form (row tiles with the near-boundary cells in a second
    launch or in the same pass, 64-cell tiles per wavefront) and the
    autotuner's pick: the same bits as each domain stepped alone."""
    specs = workloads.cavity_set(n=10, lmin=12, lmax=40, seed=7)
    say(f"batch == single {storage} {config}: {[s[:3] for s in specs]}")
    batch_doms = [cavity(*s[:4], storage) for s in specs]
    single_doms = [cavity(*s[:4], storage) for s in specs]
    b = capi.Batch(batch_doms)
    vec = 1
    if config == "tuned":
        say(f"  tune: {b.tune(2):.3f} ms/step, {b.info()}")
    else:
        b.set_config(KERNELS[config], vec, 0, 8)
        assert b.info()["kernel"] == KERNELS[config]
    b.step(37)
    b.sync()
    for d in single_doms:
        d.step(37)
    for db, ds in zip(batch_doms, single_doms):
        rb, ub = db.moments()
        rs, us = ds.moments()
        assert np.array_equal(bits(rb), bits(rs)) and np.array_equal(bits(ub), bits(us))
    # a second call reuses the captured graph and continues from the new state
    b.step(37)
    b.sync()
    for d in single_doms:
        d.step(37)
    for db, ds in zip(batch_doms, single_doms):
        assert np.array_equal(bits(db.get_pdf()), bits(ds.get_pdf()))


@pytest.mark.parametrize("vec", [2, 4])
@pytest.mark.parametrize("config", ["rows_separate", "rows_fused"])
@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_batch_rows_16bit_equals_single(storage, config, vec):
    """The row kernels of 16-bit storage (2 or 4 cells per thread), on
    cavities whose rows are long enough to be line-aligned (126 cells and
    more; shorter rows are packed and take the cell tiles): the same bits as
    each domain stepped alone."""
    specs = [(130, 10, 8, 0.62), (150, 7, 12, 0.55), (200, 5, 6, 0.7), (127, 3, 4, 0.9)]
    say(f"batch rows == single {storage} {config} vec={vec}: {specs}")
    batch_doms = [cavity(*s, storage) for s in specs]
    single_doms = [cavity(*s, storage) for s in specs]
    b = capi.Batch(batch_doms)
    b.set_config(KERNELS[config], vec, 0, 8)
    info = b.info()
    assert info["kernel"] == KERNELS[config] and info["vec"] == vec, info
    b.step(23)
    b.sync()
    for d in single_doms:
        d.step(23)
    for db, ds in zip(batch_doms, single_doms):
        assert np.array_equal(bits(db.get_pdf()), bits(ds.get_pdf()))
        _, u = db.moments()
        assert np.abs(u).max() > 1e-4
    b.close()
    for d in batch_doms + single_doms:
        d.close()


def test_batch_subsample_against_oracle():
    specs = workloads.cavity_set(n=24, lmin=16, lmax=48, re_min=10, re_max=100, seed=11)
    say("batch vs oracle")
    doms = [cavity(*s[:4], "f32") for s in specs]
    b = capi.Batch(doms)
    b.tune(2)
    steps = 200
    b.step(steps)
    b.sync()
    for i in (0, 7, 15, 23):
        xl, yl, zl, tau, re = specs[i]
        o = ora.OracleDomain(xl, yl, zl, tau, q=19)
        scenario.cavity(xl, yl, zl, lid=(workloads.U_LID, 0.0, 0.0)).apply(o)
        o.finalize()
        o.step(steps)
        ro, uo = o.moments()
        rg, ug = doms[i].moments()
        fl = o.kinds()[1:-1, 1:-1, 1:-1] == 0
        e = np.linalg.norm(ug[fl] - uo[fl]) / np.linalg.norm(uo[fl])
        say(f"  cavity {specs[i][:3]} Re {re:.0f}: u rel L2 vs fp64 oracle {e:.2e}")
        assert e < 2e-4


def lid_cavity(xl, yl, zl, tau, storage, lid):
    d = capi.Domain(xl, yl, zl, tau, q=19, storage=storage)
    scenario.cavity(xl, yl, zl, lid=lid).apply(d)
    d.finalize()
    return d


def test_batch_adjoint_matches_single_tapes_and_differences():
    """Batched adjoint of the kinetic energy after 40 steps with respect to
    each cavity's lid velocity: per domain the objective equals a single-domain
    tape's bit for bit, the gradient agrees with it to 1e-12 (fp64 shifted),
    with central differences of the objective on one cavity, and fp32 with
    fp64 to 1e-5."""
    specs = workloads.cavity_set(n=6, lmin=10, lmax=20, re_min=50, re_max=200, seed=3)
    steps, lid = 40, (workloads.U_LID, 0.01, 0.0)
    say(f"batched adjoint: {[s[:3] for s in specs]}")
    grads = {}
    for storage in ("f64s", "f32"):
        doms = [lid_cavity(*s[:4], storage, lid) for s in specs]
        b = capi.Batch(doms)
        t = capi.BatchTape(b, steps, checkpoint_every=7)
        J = t.forward()
        g = t.backward()
        say(f"  {storage}: J {J.round(12).tolist()} timing {t.timing()}")
        grads[storage] = (J, g)
        t.close()
        b.close()
    J64, g64 = grads["f64s"]
    J32, g32 = grads["f32"]
    for i, s in enumerate(specs):
        d = lid_cavity(*s[:4], "f64s", lid)
        tape = capi.Tape(d, steps, capi.OBJ_KINETIC_ENERGY, checkpoint_every=7)
        j1 = tape.forward()
        _, gw = tape.backward(1.0, want_alpha=False)
        tape.close()
        say(f"  cavity {s[:3]}: J batch {J64[i]:.15e} single {j1:.15e}; grad batch {g64[i]} single {gw}")
        assert J64[i] == j1
        assert np.allclose(g64[i], gw, rtol=1e-12, atol=1e-12 * np.abs(gw).max())
        assert np.allclose(g32[i], g64[i], rtol=1e-5, atol=1e-5 * np.abs(g64[i]).max())
    # central differences on cavity 0 (fp64 shifted), both lid components
    xl, yl, zl, tau = specs[0][:4]
    eps = 1e-4
    for comp in (0, 1):
        js = []
        for sgn in (1, -1):
            l = list(lid)
            l[comp] += sgn * eps
            d = lid_cavity(xl, yl, zl, tau, "f64s", tuple(l))
            b = capi.Batch([d])
            t = capi.BatchTape(b, steps)
            js.append(t.forward()[0])
            t.close()
            b.close()
        fd = (js[0] - js[1]) / (2 * eps)
        say(f"  lid component {comp}: adjoint {g64[0][comp]:.10e} fd {fd:.10e}")
        assert abs(g64[0][comp] - fd) <= 1e-6 * abs(fd)


@pytest.mark.parametrize("storage", ["f32", "f64s"])
def test_batch_adjoint_two_levels_equal_one_level(storage):
    """Two checkpoint levels (checkpoints every 12 steps, every 5 inside a
    segment) give the objective and the gradients of one level (every 7
    steps) bit for bit, on 6 cavities over 40 steps."""
    specs = workloads.cavity_set(n=6, lmin=10, lmax=20, re_min=50, re_max=200, seed=3)
    steps, lid = 40, (workloads.U_LID, 0.01, 0.0)
    out = []
    for kw in ({"checkpoint_every": 7}, {"checkpoint_every": 12, "sub_every": 5}):
        doms = [lid_cavity(*s[:4], storage, lid) for s in specs]
        b = capi.Batch(doms)
        t = capi.BatchTape(b, steps, **kw)
        lv = t.levels()
        J = t.forward()
        g = t.backward()
        say(f"  {storage} {kw}: levels {lv}, timing {t.timing()}")
        out.append((lv, J, g))
        t.close()
        b.close()
        for d in doms:
            d.close()
    (l1, J1, g1), (l2, J2, g2) = out
    assert l1 == {"checkpoint_every": 7, "sub_every": 0, "slots_per_domain": 6 + 6 + 2}
    assert l2 == {"checkpoint_every": 12, "sub_every": 5, "slots_per_domain": 4 + 2 + 4 + 2 + 2}
    assert np.array_equal(J1, J2)
    assert np.array_equal(g1.view(np.uint64), g2.view(np.uint64))
