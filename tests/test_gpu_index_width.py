"""The cell-tile, in-place, batched and batched-adjoint kernels are built
twice: with a uniform direction base plus a 32-bit byte offset (the default
wherever a plane holds less than 4 GiB) and with 64-bit element indices.
lbm_set_index_width forces the second build; both must give the same bits on
a walled cavity (both layouts), a FreeSlip scenario, and a batch with its
batched adjoint."""
import sys
from pathlib import Path

import numpy as np
import pytest

from lbm_amd import capi, scenario, workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def build(dims, storage, sc, width, layout="ab", tau=0.6, q=19):
    d = capi.Domain(*dims, tau, q=q, storage=storage, layout=layout)
    sc.apply(d)
    d.set_index_width(width)
    d.finalize()
    return d


@pytest.mark.parametrize("storage,layout", [("f32", "ab"), ("f32", "aa"), ("f16", "ab"), ("f64", "aa")])
def test_cavity_same_bits_for_both_index_widths(storage, layout):
    dims = (37, 29, 23)
    sc = scenario.cavity(*dims, lid=(0.05, 0.02, 0.0))
    say(f"index width: cavity {dims} {storage} {layout}")
    runs = {}
    for w in (32, 64):
        d = build(dims, storage, sc, w, layout)
        d.step(41)
        runs[w] = d.get_pdf()
        d.close()
    assert np.array_equal(bits(runs[32]), bits(runs[64]))


@pytest.mark.parametrize("q", [19, 27])
def test_freeslip_scenario_same_bits_for_both_index_widths(q):
    sc = scenario.parse_scenario_file(Path(__file__).parent / "golden" / "scenarios" / "shearflow.xml")
    say(f"index width: shearflow.xml q={q}")
    runs = {}
    for w in (32, 64):
        d = build((sc.xl, sc.yl, sc.zl), "f32", sc, w, q=q)
        d.step(60)
        runs[w] = d.get_pdf()
        d.close()
    assert np.array_equal(bits(runs[32]), bits(runs[64]))


@pytest.mark.parametrize("storage", ["f32", "f64s"])
def test_batch_and_batched_adjoint_same_bits_for_both_index_widths(storage):
    specs = workloads.cavity_set(n=5, lmin=10, lmax=24, re_min=50, re_max=200, seed=5)
    say(f"index width: batch of {[s[:3] for s in specs]} {storage}")
    out = {}
    for w in (32, 64):
        doms = []
        for xl, yl, zl, tau, _ in specs:
            doms.append(build((xl, yl, zl), storage, scenario.cavity(xl, yl, zl, lid=(workloads.U_LID, 0.01, 0.0)),
                              w, tau=tau))
        b = capi.Batch(doms)
        b.set_config(capi.BATCH_CELLS, 1, 0, 8)
        b.step(23)
        b.sync()
        fields = [d.get_pdf() for d in doms]
        t = capi.BatchTape(b, 30, checkpoint_every=6)
        J = t.forward()
        g = t.backward()
        t.close()
        b.close()
        for d in doms:
            d.close()
        out[w] = (fields, J, g)
    for f32, f64 in zip(out[32][0], out[64][0]):
        assert np.array_equal(bits(f32), bits(f64))
    assert np.array_equal(bits(out[32][1]), bits(out[64][1]))
    assert np.array_equal(bits(out[32][2]), bits(out[64][2]))
