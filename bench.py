"""Benchmark of the collide-stream time step on one MI355X.

Workload: periodic 256^3 D3Q19 box, BGK tau = 0.8, 2-D Taylor-Green vortex
(u0 = 0.02, k = 2 pi / 256) plus a seeded density perturbation 1e-4 N(0,1),
shifted distributions stored as fp32 (the headline) and as fp16 / bf16, all
with fp32 arithmetic. A step is one pull-stream + collide update of all
16.8 M cells; inputs are resident in HBM when timing starts. lbm_step runs
this bulk-only box two steps per pass (k_bulk2: the field is read and
written once per pair of steps, the intermediate state stays on chip); each
mode is also timed in single steps (k_bulk).

Prints one JSON line: MLUPS of the fp32 mode (interior cells x steps / s),
the roofline of the dominant kernel (algorithmic bytes 2*Q*sizeof(storage)
per cell and step, times the cell-steps of one launch, over its
HIP-event-timed average launch; beside it the HBM bytes rocprofv3 measured
per launch), the fp16/bf16 modes and a CPU baseline (the fp64 CPU oracle,
OpenMP on the host cores, bounded sample).
Further legs ride on the same line:
- "walled": the 256^3 lid-driven cavity (D3Q19, tau 0.6) and a 384x192x192
  channel around seeded spheres (D3Q27, tau 0.55), fp32 and fp16, each as
  two steps per pass (k_wall2) and as single steps (k_cells), with the form
  lbm_step picks by default named;
- "report": the periodic box in fp64 (the reference's operation order) and
  in D3Q15 / D3Q27;
- "inplace_aa": the same workload in the in-place (AA) layout, one buffer;
- "adjoint": the inverse problem on a periodic 128^3 D3Q19 fp32 box with a
  Brinkman porosity field (seeded uniform 32^3, trilinear), body force 1e-5 x,
  2000 steps with sqrt(N) checkpoints, plane-velocity misfit objective: wall
  time of the forward pass, of the checkpoint recompute and of the adjoint
  sweep, and their ratio;
- "batch": 512 lid-driven D3Q19 fp32 cavities (edges log-uniform in
  [16, 96], Re uniform in [10, 1000] through per-domain tau) stepped 2000
  times in one graph-captured launch per step after autotuning, then the
  batched adjoint of every cavity's kinetic energy with respect to its lid
  velocity over 2000 steps (two checkpoint levels).
Progress lines (one per finished leg) go to stderr.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-adjoint] [--no-batch]
                    [--modes f32,f16,bf16]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
N = 256
Q = 19
TAU = 0.8
BYTES = {"f32": 4, "f16": 2, "bf16": 2, "f64": 8}


def tgv(n: int, seed: int = 12345):
    k = 2 * np.pi / n
    x = np.arange(n)[None, None, :]
    y = np.arange(n)[None, :, None]
    u = np.zeros((n, n, n, 3))
    u[..., 0] = 0.02 * np.sin(k * x) * np.cos(k * y)
    u[..., 1] = -0.02 * np.cos(k * x) * np.sin(k * y)
    rng = np.random.default_rng(seed)
    rho = 1.0 + 1e-4 * rng.standard_normal((n, n, n))
    return rho, u


def traffic_from_profiles(mode: str, kernel: str = "bulk", workload: str = "d3q19_256"):
    """HBM bytes per launch of a kernel (bulk: k_bulk, pass: k_bulk2, cells:
    k_cells) on a workload, measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, each scaled by the copy-kernel calibration), from the
    committed summary profiles/pmc_<kernel>_<workload>_<mode>.txt; None when
    absent."""
    p = ROOT / "profiles" / f"pmc_{kernel}_{workload}_{mode}.txt"
    if not p.exists():
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", p.read_text(), flags=re.M)
    return float(m.group(1)) if m else None


def run_mode(capi, mode: str, steps: int, warmup: int, rho, u, steps_per_pass: int = 2, q: int = Q):
    d = capi.Domain(N, N, N, TAU, q=q, storage=mode, periodic=(True, True, True))
    d.finalize()
    d.set_time_blocking(steps_per_pass)
    d.init_equilibrium(rho, u)
    d.step(warmup)
    d.sync()
    t0 = time.perf_counter()
    ev = d.time_steps(steps)  # HIP events around every kernel launch, on the domain's stream
    d.sync()
    wall = time.perf_counter() - t0
    cells = N * N * N
    mass = d.conservation()
    info = d.info()
    d.close()
    spp = info["steps_per_pass"]
    # launches in the timed region (an odd count adds one single step)
    launches = steps // spp + steps % spp
    kernel_ms = ev["bulk_ms"] / launches
    # bytes one launch must move: every population read once and written
    # once, whether the launch advances one step or two
    moved = 2 * q * BYTES[mode] * cells
    return {
        "kernel": "k_bulk2" if spp == 2 else "k_bulk",
        "steps_per_pass": spp,
        "z_chunk": info["z_chunk"] if spp == 2 else None,
        "mlups": cells * steps / wall / 1e6,
        "mlups_ghost_inclusive": (N + 2) ** 3 * steps / wall / 1e6,
        "ms_per_step": wall * 1e3 / steps,
        "event_ms_per_step": ev["total_ms"] / steps,
        "kernel_ms": kernel_ms,
        "bytes_per_launch": moved,
        "achieved_gbs": moved / (kernel_ms * 1e-3) / 1e9,
        # step-equivalent rate: 2*Q*s bytes per cell-update over the launch
        # time; above the HBM peak where a launch advances two steps
        "step_equiv_gbs": moved * spp / (kernel_ms * 1e-3) / 1e9,
        "mass": float(mass[0]),
        "vec": info["vec"],
        "block": [info["block_x"], info["block_y"]],
    }


def mode_row(r: dict, mode: str, q: int = Q) -> dict:
    return {"mlups": round(r["mlups"], 1), "kernel": r["kernel"], "steps_per_launch": r["steps_per_pass"],
            "kernel_ms": round(r["kernel_ms"], 5), "bytes_per_launch": r["bytes_per_launch"],
            "achieved_gbs": round(r["achieved_gbs"], 1),
            "roofline_frac": round(r["achieved_gbs"] / HBM_PEAK_GBS, 4),
            "step_equiv_gbs": round(r["step_equiv_gbs"], 1),
            "bytes_per_cell_step": 2 * q * BYTES[mode]}


def walled_case(capi, name: str, q: int, mode: str, dims, tau: float, sc, steps: int, warmup: int = 5):
    """One walled workload, as two steps per pass (k_wall2, forced where the
    domain allows it) and as single steps (64-cell tiles, k_cells), with the
    form lbm_step picks by default named: event-timed kernel time per
    launch, MLUPS over interior and
    over fluid cells, and the roofline on the bytes a launch must move (each
    population of every fluid cell read once and written once: 2*Q*s per
    fluid cell, for one step or two; solid cells are not updated)."""
    out = {}
    for spp in (2, 1):
        d = capi.Domain(*dims, tau, q=q, storage=mode)
        sc.apply(d)
        d.finalize()
        if spp == 2:
            out["default_steps_per_pass"] = d.info()["steps_per_pass"]
        d.set_time_blocking(spp)
        info = d.info()
        d.step(warmup)
        d.sync()
        ev = d.time_steps(steps)
        d.sync()
        d.close()
        k = info["steps_per_pass"]
        launches = steps // k + steps % k
        kernel_ms = ev["bulk_ms"] / launches
        interior, fluid = info["ncells_interior"], info["ncells_fluid"]
        moved = 2 * q * BYTES[mode] * fluid
        gbs = moved / (kernel_ms * 1e-3) / 1e9
        kname = "k_wall2" if k == 2 else ("k_cells" if info["cell_tiles"] else "k_bulk+k_near")
        traffic = traffic_from_profiles(mode, "wall2" if k == 2 else "cells", name)
        row = {"kernel": kname, "steps_per_launch": k, "kernel_ms": round(kernel_ms, 5),
               "ms_per_step": round(ev["total_ms"] / steps, 5),
               "mlups": round(interior / (ev["total_ms"] / steps) / 1e3, 1),
               "mlups_fluid": round(fluid / (ev["total_ms"] / steps) / 1e3, 1),
               "bytes_per_launch": moved, "achieved_gbs": round(gbs, 1),
               "roofline_frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
               "traffic_over_bytes": round(traffic / moved, 4) if traffic else None}
        if k == 2 or spp == 1:
            out["pass" if k == 2 else "single_step"] = row
        if spp == 2 and k == 1:
            out["pass"] = None
        if spp == 1:
            out.update({"lattice": f"D3Q{q}", "storage": mode, "shape": list(dims), "tau": tau, "steps": steps,
                        "interior_cells": interior, "fluid_cells": fluid, "near_cells": info["ncells_near"]})
    head = out["pass"] if out["default_steps_per_pass"] == 2 else out["single_step"]
    out["mlups"] = head["mlups"]
    return out


def walled_leg(capi, steps: int, modes=("f32", "f16")):
    """The scenarios the reference runs are walled: a 256^3 lid-driven cavity
    (D3Q19, tau 0.6, lid 0.05 along x on zmax, the reference's cavity.xml
    geometry) and the 384x192x192 D3Q27 obstacle channel (inflow 0.04 at x0,
    outflow at xmax, no-slip sides, seeded spheres to a solid fraction of 0.2,
    tau 0.55)."""
    from lbm_amd import scenario, workloads
    out = {}
    cav = scenario.cavity(N, N, N, lid=(0.05, 0.0, 0.0))
    for m in modes:
        out[f"cavity256_q19_{m}"] = walled_case(capi, "cavity256_q19", 19, m, (N, N, N), 0.6, cav, steps)
    ch = workloads.sphere_channel(384, 192, 192, u_in=0.04, solid_fraction=0.2, seed=12345)
    for m in modes:
        out[f"channel384_q27_{m}"] = walled_case(capi, "channel384_q27", 27, m, (384, 192, 192), 0.55, ch,
                                                 max(steps // 2, 1))
    return out


def report_leg(capi, steps: int, warmup: int, rho, u):
    """The lattice and precision rows of the measurement table on the same
    periodic 256^3 TGV: D3Q19 fp64 (the reference's operation order), D3Q15
    fp32/fp64, D3Q27 fp32/fp16/bf16/fp64, each in lbm_step's default form."""
    rows = {}
    for q, mode in ((19, "f64"), (15, "f32"), (15, "f64"), (27, "f32"), (27, "f16"), (27, "bf16"), (27, "f64")):
        r = run_mode(capi, mode, steps, warmup, rho, u, q=q)
        row = mode_row(r, mode, q)
        row["mlups_ghost_inclusive"] = round(r["mlups_ghost_inclusive"], 1)
        rows[f"d3q{q}_{mode}"] = row
    return rows


def aa_leg(capi, mode: str, steps: int, rho, u):
    """The same workload in the in-place (AA) layout: one distribution buffer,
    one step per launch; beside it the two-buffer layout in single steps."""
    out = {}
    for layout in ("ab", "aa"):
        d = capi.Domain(N, N, N, TAU, q=Q, storage=mode, periodic=(True, True, True), layout=layout)
        d.finalize()
        d.set_time_blocking(1)
        d.init_equilibrium(rho, u)
        d.step(10)
        d.sync()
        ev = d.time_steps(steps)
        out[layout] = {"kernel_ms": round(ev["bulk_ms"] / steps, 5), "device_gb": round(d.info()["device_bytes"] / 1e9, 3)}
        d.close()
    return {"storage": mode, "steps": steps,
            "mlups": round(N ** 3 / (out["aa"]["kernel_ms"] * 1e3), 1),
            "kernel_ms": out["aa"]["kernel_ms"], "device_gb": out["aa"]["device_gb"],
            "two_buffer_single_step_kernel_ms": out["ab"]["kernel_ms"],
            "two_buffer_device_gb": out["ab"]["device_gb"]}


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def cpu_baseline(rho, u, steps: int = 4):
    """The fp64 CPU oracle (the reference algorithm in its AoS layout) on the
    same 256^3 D3Q19 periodic workload, OpenMP over the host cores. MLUPS are
    given over interior cells and in the reference's own ghost-inclusive
    form (2+xl)(2+yl)(2+zl) * steps / t (src/main.cpp:64-66)."""
    from oracle import ora
    o = ora.OracleDomain(N, N, N, TAU, q=Q, periodic=(True, True, True))
    o.finalize()
    o.init_equilibrium(rho, u)
    o.step(1)  # touch pages
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    o.close()
    cores = ora.num_threads()
    return {"value": N ** 3 * steps / dt / 1e6, "unit": "MLUPS", "cores": cores, "kind": "port",
            "mlups_ghost_inclusive": (N + 2) ** 3 * steps / dt / 1e6,
            "cpu_model": cpu_model(),
            "sample": f"fp64 CPU oracle, D3Q19 256^3 periodic TGV, {steps} steps after 1 untimed, "
                      f"{cores} OpenMP threads ({os.cpu_count()} logical CPUs visible), "
                      f"{dt:.1f} s",
            "layout": "the reference's AoS fp64 cells (Q values per cell) in two ghosted fields, pull "
                      "stream, then the two-pass collide; boundary handlers are a uint16 id per cell "
                      "dispatched by a switch, where the reference holds a Collision pointer per "
                      "cell and calls through a vtable"}


def adjoint_leg(capi, steps: int = 2000):
    """Forward + recompute + adjoint of the 128^3 Brinkman inverse problem."""
    from lbm_amd import workloads
    n, tau, g = 128, 0.8, (1e-5, 0.0, 0.0)
    _, alpha, alpha_t = workloads.porosity_pair(n)
    px = n // 2
    dt = capi.Domain(n, n, n, tau, q=Q, storage="f32", periodic=(True, True, True))
    dt.set_brinkman(alpha_t, g)
    dt.finalize()
    dt.step(steps)
    _, ut = dt.moments()
    target = np.ascontiguousarray(ut[:, :, px - 1, :])
    dt.close()
    d = capi.Domain(n, n, n, tau, q=Q, storage="f32", periodic=(True, True, True))
    d.set_brinkman(alpha, g)
    d.finalize()
    tape = capi.Tape(d, steps, capi.OBJ_PLANE_VELOCITY_MISFIT, plane_x=px, target_u=target)
    tape.forward(alpha)
    tape.backward(1.0)
    tm = tape.timing()
    tape.close()
    d.close()
    fwd = tm["forward_ms"]
    return {"workload": "D3Q19 periodic 128^3 Brinkman, g=1e-5 x, tau=0.8, fp32, plane-misfit objective",
            "steps": steps, "checkpoints": "sqrt(N)",
            "forward_ms": round(fwd, 3), "recompute_ms": round(tm["recompute_ms"], 3),
            "adjoint_ms": round(tm["adjoint_ms"], 3),
            "adjoint_step_over_forward_step": round(tm["adjoint_ms"] / fwd, 3),
            "backward_over_forward": round((tm["recompute_ms"] + tm["adjoint_ms"]) / fwd, 3)}


def batch_leg(capi, steps: int = 2000, n: int = 512, adj_steps: int = 2000):
    """512 uneven lid-driven cavities in one persistent graph-captured launch."""
    from lbm_amd import scenario, workloads
    specs = workloads.cavity_set(n=n)
    doms = []
    for xl, yl, zl, tau, _ in specs:
        d = capi.Domain(xl, yl, zl, tau, q=Q, storage="f32")
        scenario.cavity(xl, yl, zl, lid=(workloads.U_LID, 0.0, 0.0)).apply(d)
        d.finalize()
        doms.append(d)
    b = capi.Batch(doms)
    b.tune(5)
    info = b.info()
    b.step(20)
    b.sync()
    t = b.time_steps(steps)
    # tau = 3 U L / Re + 1/2 reaches 0.5024 for small, high-Re cavities, where
    # BGK is unstable: the device-side flags name the cavities that diverged
    bad = b.status()
    cells = np.array([d.xl * d.yl * d.zl for d in doms])
    finite_cells = int(cells[~bad].sum())
    ms = t["ms"] / steps
    out = {"workload": f"{n} D3Q19 lid-driven cavities, edges log-uniform [16,96], Re U[10,1000], fp32",
           "steps": steps, "cells": t["cells"], "ms_per_step": round(ms, 5),
           "mlups": round(t["mlups"], 1),
           "diverged_domains": int(bad.sum()), "finite_domains": int((~bad).sum()),
           "finite_cells": finite_cells, "mlups_finite_domains": round(finite_cells / (ms * 1e3), 1),
           "bytes_per_step": 2 * Q * 4 * t["cells"],
           "achieved_gbs": round(2 * Q * 4 * t["cells"] / (ms * 1e-3) / 1e9, 1),
           "roofline_frac": round(2 * Q * 4 * t["cells"] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "kernel": info["kernel"], "vec": info["vec"],
           "blocks_per_cu": info["blocks_per_cu"], "tiles": info["ntiles"], "near_cells": info["nnear"],
           "adjoint": None}
    if adj_steps <= 0:
        b.close()
        for d in doms:
            d.close()
        return out
    b.close()
    for d in doms:
        d.close()
    # batched adjoint of each cavity's kinetic energy with respect to its lid
    # velocity, on a fresh copy of the set over adj_steps steps (2000: one
    # level of sqrt(N) checkpoints of all 512 domains would not fit in device
    # memory, so the tape takes two levels)
    doms = []
    for xl, yl, zl, tau, _ in specs:
        d = capi.Domain(xl, yl, zl, tau, q=Q, storage="f32")
        scenario.cavity(xl, yl, zl, lid=(workloads.U_LID, 0.0, 0.0)).apply(d)
        d.finalize()
        doms.append(d)
    b = capi.Batch(doms)
    b.set_config(info["kernel"], info["vec"], info["bx_policy"], info["blocks_per_cu"])  # the tuned launch
    tape = capi.BatchTape(b, adj_steps)
    levels = tape.levels()
    J = tape.forward()
    g = tape.backward()
    tm = tape.timing()
    # cavities that diverge in the forward run have a non-finite objective
    stable = np.isfinite(J)
    tape.close()
    b.close()
    for d in doms:
        d.close()
    out["adjoint"] = {"objective": "kinetic energy per cavity, gradient w.r.t. its lid velocity",
                      "steps": adj_steps, "checkpoints": levels, "forward_ms": round(tm["forward_ms"], 3),
                      "recompute_ms": round(tm["recompute_ms"], 3), "adjoint_ms": round(tm["adjoint_ms"], 3),
                      "adjoint_step_over_forward_step": round(tm["adjoint_ms"] / tm["forward_ms"], 3),
                      "backward_over_forward": round((tm["recompute_ms"] + tm["adjoint_ms"]) / tm["forward_ms"], 3),
                      "stable_cavities": int(stable.sum()),
                      "gradients_finite_on_stable": bool(np.isfinite(g[stable]).all())}
    return out


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--modes", default="f32,f16,bf16")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=50)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-aa", action="store_true", help="skip the in-place layout line")
    ap.add_argument("--no-walled", action="store_true", help="skip the walled (cavity, obstacle channel) leg")
    ap.add_argument("--walled-modes", default="f32,f16")
    ap.add_argument("--no-report", action="store_true", help="skip the lattice / fp64 report rows")
    ap.add_argument("--batch-steps", type=int, default=2000, help="timed steps of the 512-cavity batch")
    ap.add_argument("--batch-adjoint-steps", type=int, default=2000, help="batched adjoint steps (0: skip)")
    ap.add_argument("--calibrate", action="store_true",
                    help="short copy-bandwidth run (3 copies) for rocprofv3 --pmc passes")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs one domain on one GPU; --gpus {a.gpus} is not supported "
              "(multi-GPU slab decomposition is a design note only)", file=sys.stderr)
        sys.exit(2)

    def progress(msg):
        # one line per finished leg on stderr (stdout carries the JSON line)
        print(f"bench: {msg}", file=sys.stderr, flush=True)

    from lbm_amd import capi
    capi.load()
    # plain 1 GiB device copy: achievable-bandwidth reference, and the kernel of
    # known size that scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters
    copy = capi.copy_bandwidth(1 << 30, reps=3 if a.calibrate else 10)
    # the same data movement as a step (Q planes read, Q planes written)
    # without the arithmetic or the pull offsets
    qcopy = capi.copy_planes_bandwidth(Q, N ** 3, reps=3 if a.calibrate else 10)
    rho, u = tgv(N)
    modes = [m for m in a.modes.split(",") if m]
    res = {m: run_mode(capi, m, a.steps, a.warmup, rho, u) for m in modes}
    single = {m: run_mode(capi, m, a.steps, a.warmup, rho, u, steps_per_pass=1) for m in modes}
    progress(f"periodic 256^3 legs done ({', '.join(modes)})")
    head = res.get("f32") or res[modes[0]]
    head_mode = "f32" if "f32" in res else modes[0]
    traffic = traffic_from_profiles(head_mode, "pass" if head["steps_per_pass"] == 2 else "bulk")
    s_head = single[head_mode]
    s_traffic = traffic_from_profiles(head_mode, "bulk")
    line = {
        "metric": "MLUPS + %HBM roofline, D3Q19 256^3 fp32 vs fp16-store",
        "value": round(head["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(head["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Taylor-Green vortex + density noise, seed 12345)",
        "config": {"workload": "D3Q19 periodic 256^3 TGV, BGK tau=0.8, shifted f-w stored fp32",
                   "lattice": "D3Q19", "shape": [N, N, N], "tau": TAU, "storage": head_mode,
                   "steps_per_pass": head["steps_per_pass"], "z_chunk": head["z_chunk"],
                   "mlups_ghost_inclusive": round(head["mlups_ghost_inclusive"], 1)},
        # achieved = bytes one launch must move (each population read once and
        # written once: 2*Q*s per cell, also for a launch that advances two
        # steps) / the event-timed average launch
        "roofline": {"bound": "hbm", "achieved": round(head["achieved_gbs"], 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(head["achieved_gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": head["kernel"], "kernel_ms": round(head["kernel_ms"], 5),
                     "steps_per_launch": head["steps_per_pass"],
                     "bytes_per_launch": head["bytes_per_launch"],
                     "traffic_over_bytes": round(traffic / head["bytes_per_launch"], 4) if traffic else None,
                     # HBM bytes actually moved per launch over its duration
                     "hbm_gbs_measured": (round(traffic / (head["kernel_ms"] * 1e-3) / 1e9, 1)
                                          if traffic else None),
                     # 2*Q*s per cell-update over the launch time: a rate of
                     # step-equivalents, not an HBM rate (a pass moves half of it)
                     "step_equiv_gbs": round(head["step_equiv_gbs"], 1),
                     "copy_gbs": round(copy["gbs"], 1),
                     "q_plane_copy_gbs": round(qcopy["gbs"], 1),
                     "single_step": {"kernel": s_head["kernel"], "kernel_ms": round(s_head["kernel_ms"], 5),
                                     "achieved": round(s_head["achieved_gbs"], 1),
                                     "frac": round(s_head["achieved_gbs"] / HBM_PEAK_GBS, 4),
                                     "traffic": s_traffic,
                                     "bytes_per_launch": s_head["bytes_per_launch"],
                                     "frac_of_copy": round(s_head["achieved_gbs"] / copy["gbs"], 4),
                                     "frac_of_q_plane_copy": round(s_head["achieved_gbs"] / qcopy["gbs"], 4),
                                     "mlups": round(s_head["mlups"], 1),
                                     "vec": s_head["vec"], "block": s_head["block"]}},
        "modes": {m: dict(mode_row(r, m), single_step=mode_row(single[m], m)) for m, r in res.items()},
    }
    if "f16" in res and "f32" in res:
        line["f16_over_f32"] = round(res["f16"]["mlups"] / res["f32"]["mlups"], 3)
        line["f16_over_f32_single_step"] = round(single["f16"]["mlups"] / single["f32"]["mlups"], 3)
    if not a.no_walled:
        line["walled"] = walled_leg(capi, min(a.steps, 200), [m for m in a.walled_modes.split(",") if m])
        progress("walled leg done")
    if not a.no_report:
        line["report"] = report_leg(capi, min(a.steps, 100), a.warmup, rho, u)
        progress("report rows done")
    if not a.no_aa:
        line["inplace_aa"] = aa_leg(capi, head_mode, min(a.steps, 200), rho, u)
        progress("in-place layout leg done")
    line["adjoint"] = None if a.no_adjoint else adjoint_leg(capi)
    progress("adjoint leg done")
    line["batch"] = None if a.no_batch else batch_leg(capi, a.batch_steps, adj_steps=a.batch_adjoint_steps)
    progress("batch leg done")
    line["cpu_baseline"] = None if a.no_cpu else cpu_baseline(rho, u, a.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
