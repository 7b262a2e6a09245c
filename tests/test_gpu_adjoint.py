"""Adjoint collide-stream step (lbm_adj_*), fp64 forward-mode reference
(lbm_jvp) and central finite differences.

Checks, smallest first: adjoint directional derivatives against the dual-
number JVP in fp64 (agree to rounding), against central differences of the
fp64 forward run, the adjoint boundary rules (no-slip, moving wall) through a
lid-velocity gradient, fp32 against fp64 adjoints, and the full inverse
problem: D3Q19 128^3, Brinkman porosity on a 32^3 grid upsampled
trilinearly, body force 1e-5 along x, 2000 steps, sqrt(N) checkpoints,
misfit of u on the mid-x plane against a second seeded porosity; 16 random
coarse directions checked by central differences (eps = 1e-4) with an
agreement bound of 1e-4 relative, direction by direction.
"""
import sys

import numpy as np
import pytest

from lbm_amd import autograd, capi, scenario

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def periodic_brinkman(n, storage, alpha, g, tau=0.8):
    d = capi.Domain(n, n, n, tau, q=19, storage=storage, periodic=(1, 1, 1))
    d.set_brinkman(alpha, g)
    d.finalize()
    return d


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.fixture(scope="module")
def small_problem():
    n, steps = 10, 30
    rng = np.random.default_rng(3)
    alpha = rng.uniform(0.2, 0.8, (n, n, n))
    target = 1e-3 * rng.standard_normal((n, n, 3))
    return n, steps, alpha, target, rng


def test_adjoint_equals_jvp_fp64(small_problem):
    n, steps, alpha, target, rng = small_problem
    say("adjoint vs jvp f64s")
    d = periodic_brinkman(n, "f64s", alpha, (1e-3, 2e-4, 0))
    tape = capi.Tape(d, steps, capi.OBJ_PLANE_VELOCITY_MISFIT, plane_x=n // 2, target_u=target,
                     checkpoint_every=4)
    J = tape.forward(alpha)
    grad, _ = tape.backward(1.0)
    for k in range(3):
        da = rng.standard_normal(alpha.shape)
        Jd, dJ = tape.jvp(alpha, da)
        adj = float((grad * da).sum())
        say(f"  dir {k}: adjoint {adj:.12e} jvp {dJ:.12e} rel {rel(adj, dJ):.2e}")
        assert rel(Jd, J) < 1e-13
        assert rel(adj, dJ) < 1e-10


def test_adjoint_central_differences_fp64(small_problem):
    n, steps, alpha, target, rng = small_problem
    say("adjoint vs FD f64s")
    d = periodic_brinkman(n, "f64s", alpha, (1e-3, 0, 0))
    tape = capi.Tape(d, steps, capi.OBJ_PLANE_VELOCITY_MISFIT, plane_x=3, target_u=target)
    tape.forward(alpha)
    grad, _ = tape.backward(1.0)
    eps = 1e-4
    for k in range(3):
        da = rng.standard_normal(alpha.shape)
        jp = tape.forward(alpha + eps * da)
        jm = tape.forward(alpha - eps * da)
        fd = (jp - jm) / (2 * eps)
        adj = float((grad * da).sum())
        say(f"  dir {k}: adjoint {adj:.10e} fd {fd:.10e} rel {rel(adj, fd):.2e}")
        assert rel(adj, fd) < 1e-6


def lid_cavity(n, storage, lid, alpha):
    d = capi.Domain(n, n, n, 0.7, q=19, storage=storage)
    scenario.cavity(n, n, n, lid=(lid, 0.0, 0.0)).apply(d)
    d.set_brinkman(alpha, (0.0, 0.0, 0.0))
    d.finalize()
    return d


def test_adjoint_walls_and_lid_velocity_gradient_fp64():
    """Kinetic energy of a lid-driven cavity with a porosity field: the lid
    gradient (adjoint moving-wall and bounce-back rules) against central
    differences of the lid velocity, the porosity gradient against the JVP."""
    n, steps, U = 9, 40, 0.05
    say("cavity lid gradient f64s")
    rng = np.random.default_rng(9)
    alpha = rng.uniform(0.0, 0.3, (n, n, n))
    d = lid_cavity(n, "f64s", U, alpha)
    tape = capi.Tape(d, steps, capi.OBJ_KINETIC_ENERGY, checkpoint_every=7)
    J = tape.forward(alpha)
    grad, gwall = tape.backward(1.0)
    eps = 1e-5
    jp = capi.Tape(lid_cavity(n, "f64s", U + eps, alpha), steps, capi.OBJ_KINETIC_ENERGY).forward(alpha)
    jm = capi.Tape(lid_cavity(n, "f64s", U - eps, alpha), steps, capi.OBJ_KINETIC_ENERGY).forward(alpha)
    fd = (jp - jm) / (2 * eps)
    say(f"  dJ/dU adjoint {gwall[0]:.10e} fd {fd:.10e} rel {rel(gwall[0], fd):.2e}; J {J:.6e}")
    assert rel(gwall[0], fd) < 1e-6
    da = rng.standard_normal(alpha.shape)
    _, dJ = tape.jvp(alpha, da)
    adj = float((grad * da).sum())
    say(f"  porosity dir: adjoint {adj:.10e} jvp {dJ:.10e} rel {rel(adj, dJ):.2e}")
    assert rel(adj, dJ) < 1e-10


def test_fp32_adjoint_close_to_fp64(small_problem):
    n, steps, alpha, target, rng = small_problem
    say("f32 vs f64s adjoint")
    grads = {}
    for st in ("f32", "f64s"):
        d = periodic_brinkman(n, st, alpha, (1e-3, 0, 0))
        tape = capi.Tape(d, steps, capi.OBJ_PLANE_VELOCITY_MISFIT, plane_x=5, target_u=target)
        tape.forward(alpha)
        grads[st], _ = tape.backward(1.0)
    e = np.linalg.norm(grads["f32"] - grads["f64s"]) / np.linalg.norm(grads["f64s"])
    say(f"  rel L2 {e:.2e}")
    assert e < 1e-4


@pytest.mark.parametrize("storage", ["f32", "f64s"])
def test_torch_objective_on_device_equals_host(storage):
    """torch.autograd through LbmObjective and the periodic trilinear map
    (16^3 fine, 4^3 coarse, 24 steps). A coarse porosity on the GPU takes
    the device-pointer entry points and never visits the host: its objective
    and fine gradient equal the host entry points' on the same fine values
    bit for bit, and its coarse gradient (torch's transpose of the map)
    equals upsample_np_adjoint of the fine gradient to rounding, as does the
    gradient of a coarse porosity on the CPU."""
    import torch
    n, nc, steps = 16, 4, 24
    rng = np.random.default_rng(11)
    ac = rng.uniform(0.1, 0.9, (nc, nc, nc))
    target = 1e-3 * rng.standard_normal((n, n, 3))
    d = periodic_brinkman(n, storage, autograd.upsample_np(ac, (n, n, n)), (1e-3, 0, 0))
    tape = capi.Tape(d, steps, capi.OBJ_PLANE_VELOCITY_MISFIT, plane_x=n // 2, target_u=target, checkpoint_every=5)
    obj = autograd.make_objective_function()
    grads = {}
    for dev in ("cuda:0", "cpu"):
        c = torch.tensor(ac, dtype=torch.float64, device=dev, requires_grad=True)
        fine = autograd.upsample_periodic(c, (n, n, n))
        J = obj.apply(fine, tape)
        J.backward()
        grads[dev] = c.grad.detach().cpu().numpy()
        if dev == "cuda:0":
            fine_values = fine.detach().cpu().numpy()
            J_dev = float(J.detach())
            g_dev, _ = tape.backward_device(1.0)
            g_dev = g_dev.cpu().numpy()
        say(f"  {storage} {dev}: J {float(J.detach()):.15e}")
    # the host entry points on the same fine porosity
    J_host = tape.forward(fine_values)
    g_host, _ = tape.backward(1.0)
    assert J_dev == J_host
    assert np.array_equal(g_dev.view(np.uint64), g_host.view(np.uint64))
    ref = autograd.upsample_np_adjoint(g_host, ac.shape)
    for dev, gc in grads.items():
        err = np.abs(gc - ref).max() / np.abs(ref).max()
        say(f"  {storage} {dev}: coarse gradient vs upsample_np_adjoint {err:.2e}")
        assert err < 1e-9
    tape.close()
    d.close()


def test_c4_inverse_problem_gradient():
    """D3Q19 128^3 fp32, porosity from a seeded uniform 32^3 grid (trilinear,
    periodic), g = 1e-5 x, tau 0.8, 2000 steps, sqrt(N) checkpoints; target
    from porosity seed + 1; 16 random coarse directions by central FD
    (eps = 1e-4) of the fp64 forward run; adjoint step <= 3x a forward step.

    eps: at 1e-3 the central difference's own O(eps^2) truncation error
    reached 9.5e-5 of a direction's derivative (measured against the exact
    fp64 JVP), the noise floor of a per-direction 1e-4 bound; at 1e-4 it
    drops a hundredfold, while the rounding of the fp64 objective difference
    stays near 1e-7 of the derivative."""
    n, nc, steps, tau = 128, 32, 2000, 0.8
    g = (1e-5, 0.0, 0.0)
    seed = 2024
    ac = np.random.default_rng(seed).uniform(0.0, 1.0, (nc, nc, nc))
    at = np.random.default_rng(seed + 1).uniform(0.0, 1.0, (nc, nc, nc))
    alpha = autograd.upsample_np(ac, (n, n, n))
    alpha_t = autograd.upsample_np(at, (n, n, n))
    px = n // 2
    say("C4: target run")
    dt = periodic_brinkman(n, "f64s", alpha_t, g, tau)
    dt.step(steps)
    _, ut = dt.moments()
    target = np.ascontiguousarray(ut[:, :, px - 1, :])  # (z, y, 3) at x = px
    dt.close()

    say("C4: fp32 forward + adjoint")
    d32 = periodic_brinkman(n, "f32", alpha, g, tau)
    tape = capi.Tape(d32, steps, capi.OBJ_PLANE_VELOCITY_MISFIT, plane_x=px, target_u=target)
    J32 = tape.forward(alpha)
    grad_f, _ = tape.backward(1.0)
    tm = tape.timing()
    grad_c = autograd.upsample_np_adjoint(grad_f, ac.shape)
    ratio = (tm["recompute_ms"] + tm["adjoint_ms"]) / tm["forward_ms"]
    say(f"  J32 {J32:.6e}; forward {tm['forward_ms']:.1f} ms, recompute {tm['recompute_ms']:.1f} ms, "
        f"adjoint {tm['adjoint_ms']:.1f} ms; (recompute+adjoint)/forward = {ratio:.2f}")
    tape.close()
    d32.close()
    assert ratio <= 3.0

    say("C4: fp64 central differences")
    d64 = periodic_brinkman(n, "f64s", alpha, g, tau)
    t64 = capi.Tape(d64, steps, capi.OBJ_PLANE_VELOCITY_MISFIT, plane_x=px, target_u=target)
    J64 = t64.forward(alpha)
    say(f"  J64 {J64:.6e}")
    rng = np.random.default_rng(seed + 2)
    eps = 1e-4
    adjs, fds, jvps = [], [], []
    for k in range(16):
        dc = rng.standard_normal(ac.shape)
        dfine = autograd.upsample_np(dc, (n, n, n))
        jp = t64.forward(alpha + eps * dfine)
        jm = t64.forward(alpha - eps * dfine)
        fds.append((jp - jm) / (2 * eps))
        adjs.append(float((grad_c * dc).sum()))
        jvps.append(t64.jvp(alpha, dfine)[1])
        say(f"  dir {k:2d}: adjoint(fp32) {adjs[-1]:+.8e} jvp(fp64) {jvps[-1]:+.8e} "
            f"fd(fp64) {fds[-1]:+.8e}  rel vs jvp {rel(adjs[-1], jvps[-1]):.2e}  "
            f"fd-jvp rel {rel(fds[-1], jvps[-1]):.2e}")
    adjs, fds, jvps = map(np.array, (adjs, fds, jvps))
    fd_err = float((np.abs(adjs - fds) / np.abs(fds)).max())
    jvp_err = float((np.abs(adjs - jvps) / np.abs(jvps)).max())
    say(f"  max per-direction |adjoint - fd|/|fd| = {fd_err:.2e}; |adjoint - jvp|/|jvp| = {jvp_err:.2e}")
    # fp32 adjoint against the exact fp64 directional derivative and against
    # central differences (which carry an O(eps^2) truncation error of their
    # own), each direction on its own
    assert jvp_err <= 1e-4
    assert fd_err <= 1e-4


def channel(q, storage, alpha, outlet, shape=(14, 7, 6), tau=0.7, u_in=0.03):
    """Channel along x with a porosity field: inflow at x0, outflow (or a
    pressure face) at xmax, a no-slip wall at y0, a moving wall at ymax
    (velocity along x and z) and periodic z."""
    d = capi.Domain(*shape, tau, q=q, storage=storage, periodic=(0, 0, 1))
    sc = scenario.Scenario("channel", *shape)
    faces = [("y0", capi.NOSLIP, (0.0, 0.0, 0.0, 1.0)), ("ymax", capi.MOVINGWALL, (0.02, 0.0, 0.01, 1.0)),
             ("x0", capi.INFLOW, (u_in, 0.0, 0.0, 1.0)),
             ("xmax", outlet, (0.0, 0.0, 0.0, 1.0 if outlet == capi.OUTFLOW else 0.997))]
    for ext, kind, p in faces:
        sc.boundaries.append(scenario.Boundary(kind, p, scenario.face_box(ext, *shape), "", ext))
    sc.apply(d)
    d.set_brinkman(alpha, (0.0, 0.0, 0.0))
    d.finalize()
    return d


@pytest.mark.parametrize("outlet", ["outflow", "pressure"])
@pytest.mark.parametrize("q", [15, 19, 27])
def test_adjoint_outflow_pressure_against_jvp_and_differences(q, outlet):
    """Adjoint rules of the Outflow and Pressure boundaries (velocity at the
    reference density, equilibrium pair minus the bounced value), with
    Inflow, NoSlip and MovingWall on the same channel, for every lattice:
    porosity directional derivatives agree with the fp64 JVP to 1e-10, and
    with central differences of the objective per direction to 1e-6; the
    moving-wall velocity gradient agrees with central differences."""
    kind = capi.OUTFLOW if outlet == "outflow" else capi.PRESSURE
    shape, steps = (14, 7, 6), 30
    rng = np.random.default_rng(100 + q)
    alpha = rng.uniform(0.0, 0.2, shape[::-1])
    target = 1e-3 * rng.standard_normal((shape[1], shape[2], 3))
    say(f"adjoint {outlet} channel q={q}")
    d = channel(q, "f64s", alpha, kind)
    tape = capi.Tape(d, steps, capi.OBJ_PLANE_VELOCITY_MISFIT, plane_x=shape[0] - 1, target_u=target,
                     checkpoint_every=5)
    J = tape.forward(alpha)
    grad, gwall = tape.backward(1.0)
    eps = 1e-5
    for k in range(3):
        da = rng.standard_normal(alpha.shape)
        Jd, dJ = tape.jvp(alpha, da)
        adj = float((grad * da).sum())
        jp = tape.forward(alpha + eps * da)
        jm = tape.forward(alpha - eps * da)
        fd = (jp - jm) / (2 * eps)
        say(f"  dir {k}: adjoint {adj:.12e} jvp {dJ:.12e} fd {fd:.12e}")
        assert rel(Jd, J) < 1e-13
        assert rel(adj, dJ) < 1e-10
        assert rel(adj, fd) < 1e-6
    # moving-wall velocity, x component, by central differences
    js = []
    for s in (1, -1):
        dd = capi.Domain(*shape, 0.7, q=q, storage="f64s", periodic=(0, 0, 1))
        sc = scenario.Scenario("channel", *shape)
        for ext, kk, p in [("y0", capi.NOSLIP, (0.0, 0.0, 0.0, 1.0)),
                           ("ymax", capi.MOVINGWALL, (0.02 + s * eps, 0.0, 0.01, 1.0)),
                           ("x0", capi.INFLOW, (0.03, 0.0, 0.0, 1.0)),
                           ("xmax", kind, (0.0, 0.0, 0.0, 1.0 if kind == capi.OUTFLOW else 0.997))]:
            sc.boundaries.append(scenario.Boundary(kk, p, scenario.face_box(ext, *shape), "", ext))
        sc.apply(dd)
        dd.set_brinkman(alpha, (0.0, 0.0, 0.0))
        dd.finalize()
        t2 = capi.Tape(dd, steps, capi.OBJ_PLANE_VELOCITY_MISFIT, plane_x=shape[0] - 1, target_u=target)
        js.append(t2.forward(alpha))
        t2.close()
        dd.close()
    fdw = (js[0] - js[1]) / (2 * eps)
    say(f"  dJ/du_wall,x adjoint {gwall[0]:.12e} fd {fdw:.12e}")
    assert rel(gwall[0], fdw) < 1e-6
    tape.close()
    d.close()
