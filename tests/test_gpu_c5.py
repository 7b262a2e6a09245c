"""The batching configuration at its stated size: 512 lid-driven D3Q19 fp32
cavities with edges log-uniform in [16, 96] (seeded, non-cubic) and Reynolds
numbers uniform in [10, 1000] through per-domain tau = 3 U L / Re + 1/2
(U = 0.05, seed 12345), 2000 steps in the batch's graph-captured launch
after autotuning.

- The per-domain divergence flags (lbm_batch_status) mark exactly the
  cavities whose fluid cells hold NaN or infinite moments at the end.
- A seeded subsample of 16 cavities, including the smallest and the one of
  highest Reynolds number, equals each cavity stepped alone, bit for bit
  (NaN payloads aside: diverged cells compare as NaN in both).
- Three small low-Re cavities agree with the fp64 CPU oracle (the
  reference's step) to 1e-4 in the relative L2 norm of the velocity.
"""
import sys

import numpy as np
import pytest

from lbm_amd import capi, scenario, workloads
from oracle import ora

pytestmark = pytest.mark.gpu

STEPS = 2000


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cavity(xl, yl, zl, tau, storage="f32"):
    d = capi.Domain(xl, yl, zl, tau, q=19, storage=storage)
    scenario.cavity(xl, yl, zl, lid=(workloads.U_LID, 0.0, 0.0)).apply(d)
    d.finalize()
    return d


def same_bits(a, b):
    """Bitwise equal, with every NaN taken as one value."""
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False
    ca = np.ascontiguousarray(np.where(na, 0.0, a)).view(np.uint64)
    cb = np.ascontiguousarray(np.where(nb, 0.0, b)).view(np.uint64)
    return np.array_equal(ca, cb)


@pytest.fixture(scope="module")
def c5():
    specs = workloads.cavity_set(n=512, seed=12345)
    say(f"C5: {len(specs)} cavities, {sum(s[0] * s[1] * s[2] for s in specs)} cells, {STEPS} steps")
    doms = [cavity(*s[:4]) for s in specs]
    b = capi.Batch(doms)
    say(f"  tune: {b.tune(2):.3f} ms/step, {b.info()}")
    assert not b.status().any()
    b.step(STEPS)
    b.sync()
    flags = b.status()
    yield specs, doms, b, flags
    b.close()
    for d in doms:
        d.close()


def test_c5_status_flags_mark_diverged_cavities(c5):
    specs, doms, b, flags = c5
    bad = np.zeros(len(doms), dtype=bool)
    for i, d in enumerate(doms):
        r, u = d.moments()
        bad[i] = not (np.isfinite(r).all() and np.isfinite(u).all())
    taus = np.array([s[3] for s in specs])
    say(f"  diverged: {int(bad.sum())} of {len(doms)}; flagged {int(flags.sum())}; "
        f"largest tau among diverged {taus[bad].max() if bad.any() else None:.5f}")
    assert np.array_equal(flags, bad)
    assert 0 < bad.sum() < len(doms)
    # a reset clears the flags; they come back only where a step reads
    # non-finite values again
    b.status(reset=True)
    assert not b.status().any()


def test_c5_subsample_equals_single_domains(c5):
    specs, doms, b, flags = c5
    cells = [s[0] * s[1] * s[2] for s in specs]
    res = [s[4] for s in specs]
    rng = np.random.default_rng(12345)
    pick = {int(np.argmin(cells)), int(np.argmax(res))}
    while len(pick) < 16:
        pick.add(int(rng.integers(len(specs))))
    for i in sorted(pick):
        s = specs[i]
        d = cavity(*s[:4])
        d.step(STEPS)
        rs, us = d.moments()
        rb, ub = doms[i].moments()
        d.close()
        say(f"  cavity {i} {s[:3]} Re {s[4]:.0f} tau {s[3]:.5f} flagged {bool(flags[i])}: "
            f"batch == single {same_bits(rb, rs) and same_bits(ub, us)}")
        assert same_bits(rb, rs) and same_bits(ub, us)


def test_c5_small_low_re_against_oracle(c5):
    specs, doms, b, flags = c5
    low = [i for i, s in enumerate(specs) if s[4] < 100]
    low.sort(key=lambda i: specs[i][0] * specs[i][1] * specs[i][2])
    assert len(low) >= 3
    for i in low[:3]:
        xl, yl, zl, tau, re = specs[i]
        o = ora.OracleDomain(xl, yl, zl, tau, q=19)
        scenario.cavity(xl, yl, zl, lid=(workloads.U_LID, 0.0, 0.0)).apply(o)
        o.finalize()
        o.step(STEPS)
        ro, uo = o.moments()
        o.close()
        rg, ug = doms[i].moments()
        e = np.linalg.norm(ug - uo) / np.linalg.norm(uo)
        er = np.linalg.norm(rg - ro) / np.linalg.norm(ro - 1.0)
        say(f"  cavity {i} {specs[i][:3]} Re {re:.0f}: u rel L2 {e:.2e}, (rho - 1) rel L2 {er:.2e} vs fp64 oracle")
        assert not flags[i]
        assert e < 1e-4


def test_c5_batched_adjoint_2000_steps():
    """The batched adjoint on the same 512 cavities over the full 2000 steps.
    One level of sqrt(N) checkpoints does not fit in device memory at this
    size, so the automatic choice takes two levels. Gradients are finite on
    every cavity whose objective is finite, and on three small stable
    cavities they agree with a single-domain tape (fp32, 1e-6 relative)."""
    specs = workloads.cavity_set(n=512, seed=12345)
    doms = [cavity(*s[:4]) for s in specs]
    b = capi.Batch(doms)
    t = capi.BatchTape(b, STEPS)
    lv = t.levels()
    J = t.forward()
    g = t.backward()
    tm = t.timing()
    stable = np.isfinite(J)
    say(f"  levels {lv}; timing {tm}; finite objectives {int(stable.sum())} of {len(J)}")
    t.close()
    b.close()
    assert lv["sub_every"] > 0
    assert 0 < stable.sum() < len(J)
    assert np.isfinite(g[stable]).all()
    small = sorted(np.flatnonzero(stable), key=lambda i: specs[i][0] * specs[i][1] * specs[i][2])[:3]
    for d in doms:
        d.close()
    for i in small:
        d = cavity(*specs[i][:4])  # the batch's domains now hold the state after the forward
        tape = capi.Tape(d, STEPS, capi.OBJ_KINETIC_ENERGY)
        j1 = tape.forward()
        _, gw = tape.backward(1.0, want_alpha=False)
        tape.close()
        d.close()
        say(f"  cavity {i} {specs[i][:3]}: J batch {J[i]:.9e} single {j1:.9e}; grad batch {g[i]} single {gw}")
        assert np.allclose(g[i], gw, rtol=1e-6, atol=1e-6 * np.abs(gw).max())
