"""Summaries of rocprofv3 result databases for the roofline figures.

    python -m lbm_amd.rocprof_summary stats <kernel-trace results.db>
    python -m lbm_amd.rocprof_summary traffic <FETCH_SIZE results.db> <WRITE_SIZE results.db> \
        --kernel-match k_bulk< --mode-match "Mode)1" --alg-bytes N --copy-bytes B

`stats` lists per-kernel calls and average duration (the --stats view).
`traffic` scales the FETCH_SIZE / WRITE_SIZE counters (KiB per dispatch) of
one kernel by what the same counters read for a plain device copy of known
size in the same run (lbm_copy_bandwidth's k_copy16, --copy-bytes each way), and
prints HBM bytes per launch next to the algorithmic bytes.
"""
from __future__ import annotations

import argparse
import sqlite3
import sys


def kernel_stats(db: str):
    c = sqlite3.connect(db)
    rows = c.execute("select name, total_calls, total_duration, average, percentage "
                     "from top_kernels order by total_duration desc").fetchall()
    return rows


def counter_per_kernel(db: str, counter: str):
    c = sqlite3.connect(db)
    rows = c.execute("select kernel_name, avg(value), count(*) from counters_collection "
                     "where counter_name = ? group by kernel_name", (counter,)).fetchall()
    return {r[0]: (r[1], r[2]) for r in rows}


def _pick(d: dict, *needles):
    hits = [(k, v) for k, v in d.items() if all(n in k for n in needles)]
    if len(hits) != 1:
        raise SystemExit(f"expected one kernel matching {needles}, found {[h[0] for h in hits]}")
    return hits[0]


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("db")
    t = sub.add_parser("traffic")
    t.add_argument("fetch_db")
    t.add_argument("write_db")
    t.add_argument("--kernel-match", default="k_bulk<")
    t.add_argument("--mode-match", default="")
    t.add_argument("--alg-bytes", type=float, required=True)
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--copy-match", default="k_copy16")
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        print(f"{'calls':>6} {'avg_us':>10} {'total_us':>12} {'pct':>6}  kernel")
        for name, calls, tot, avg, pct in kernel_stats(a.db):  # top_kernels durations are in us
            print(f"{calls:6d} {avg:10.2f} {tot:12.1f} {pct:6.2f}  {name}")
        return
    fetch = counter_per_kernel(a.fetch_db, "FETCH_SIZE")
    write = counter_per_kernel(a.write_db, "WRITE_SIZE")
    needles = [a.kernel_match] + ([a.mode_match] if a.mode_match else [])
    kf, (fv, fn) = _pick(fetch, *needles)
    kw, (wv, wn) = _pick(write, *needles)
    cf_name, (cf, _) = _pick(fetch, a.copy_match)
    cw_name, (cw, _) = _pick(write, a.copy_match)
    f_scale = a.copy_bytes / (cf * 1024.0)
    w_scale = a.copy_bytes / (cw * 1024.0)
    fetch_b = fv * 1024.0 * f_scale
    write_b = wv * 1024.0 * w_scale
    print(f"kernel: {kf}")
    print(f"dispatches: fetch {fn}, write {wn}")
    print(f"copy kernel: {cf_name}")
    print(f"copy FETCH_SIZE KiB: {cf:.1f} (known {a.copy_bytes / 1024:.1f}) -> fetch scale {f_scale:.4f}")
    print(f"copy WRITE_SIZE KiB: {cw:.1f} (known {a.copy_bytes / 1024:.1f}) -> write scale {w_scale:.4f}")
    print(f"kernel FETCH_SIZE KiB/dispatch: {fv:.1f} -> {fetch_b:.4e} B")
    print(f"kernel WRITE_SIZE KiB/dispatch: {wv:.1f} -> {write_b:.4e} B")
    print(f"algorithmic_bytes_per_launch: {a.alg_bytes:.6e}")
    print(f"traffic_bytes_per_launch: {fetch_b + write_b:.6e}")
    print(f"traffic_over_algorithmic: {(fetch_b + write_b) / a.alg_bytes:.4f}")


if __name__ == "__main__":
    main(sys.argv[1:])
