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


@pytest.mark.parametrize("config", ["rows_separate", "rows_fused", "cells", "tuned"])
@pytest.mark.parametrize("storage", ["f32", "f16", "f64s"])
def test_batch_equals_single(storage, config):
    """Every kernel form (row tiles with the near-boundary cells in a second
    launch or in the same pass, 64-cell tiles per wavefront) and the
    autotuner's pick: the same bits as each domain stepped alone."""
    specs = workloads.cavity_set(n=10, lmin=12, lmax=40, seed=7)
    say(f"batch == single {storage} {config}: {[s[:3] for s in specs]}")
    batch_doms = [cavity(*s[:4], storage) for s in specs]
    single_doms = [cavity(*s[:4], storage) for s in specs]
    b = capi.Batch(batch_doms)
    vec = 2 if storage != "f64s" else 1
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
