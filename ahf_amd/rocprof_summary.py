"""Summaries of rocprofv3 result databases (the SQLite files that
`rocprofv3 ... -d DIR -o run` writes as DIR/run_results.db).

  python -m ahf_amd.rocprof_summary stats DB [--top N]
      per-kernel dispatch count, total / average / min / max duration
  python -m ahf_amd.rocprof_summary counters DB --kernel PREFIX
      counter values per dispatch of the matching kernels
  python -m ahf_amd.rocprof_summary traffic --fetch DB1 --write DB2 --kernel NAME [--kernel NAME ...]
                                            --copy-bytes N --workload TAG [--out JSON]
      HBM bytes per launch of the named kernels from a FETCH_SIZE run and a
      WRITE_SIZE run.  Both counters read KiB per dispatch; each is scaled by
      what it reads for the runtime's device-to-device copy kernel
      (__amd_rocclr_copyBuffer) moving a known N bytes in the same run.
"""
from __future__ import annotations

import argparse
import collections
import json
import sqlite3
import statistics
import sys

COPY_KERNEL = "__amd_rocclr_copyBuffer"


def _short(name):
    return name.split("(")[0]


def kernel_stats(db):
    con = sqlite3.connect(db)
    rows = con.execute("select name, duration from kernels").fetchall()
    agg = collections.OrderedDict()
    for name, dur in rows:
        agg.setdefault(_short(name), []).append(dur)
    out = []
    total = sum(sum(v) for v in agg.values()) or 1
    for k, v in agg.items():
        out.append({"kernel": k, "calls": len(v), "total_ms": sum(v) / 1e6, "avg_us": sum(v) / len(v) / 1e3,
                    "min_us": min(v) / 1e3, "max_us": max(v) / 1e3, "pct": 100.0 * sum(v) / total})
    out.sort(key=lambda r: -r["total_ms"])
    return out


def _matches(name, kernels):
    """kernels: one name prefix, or a list of exact kernel names"""
    if isinstance(kernels, str):
        return name.startswith(kernels)
    return name in kernels


def dispatch_counters(db, kernels):
    """{dispatch_id: (kernel, grid_size, duration_ns, {counter: value})} for
    the matching kernels (see _matches)."""
    con = sqlite3.connect(db)
    rows = con.execute("select dispatch_id, kernel_name, grid_size, counter_name, sum(value), duration "
                       "from counters_collection group by dispatch_id, counter_name").fetchall()
    out = {}
    for did, kn, gs, cn, v, dur in rows:
        if not _matches(_short(kn), kernels):
            continue
        rec = out.setdefault(did, (_short(kn), gs, dur, {}))
        rec[3][cn] = v
    return dict(sorted(out.items()))


def _per_dispatch(db, kernels, counter):
    return [c[counter] for _, (_, _, _, c) in dispatch_counters(db, kernels).items() if counter in c]


def _probe_copies(db, counter):
    """counter values of the copy probe's dispatches: the runtime copy kernel
    launches with the largest grid (the library's own small copies use the
    same kernel with small grids)"""
    recs = [(gs, c[counter]) for _, (_, gs, _, c) in dispatch_counters(db, COPY_KERNEL).items() if counter in c]
    if not recs:
        return []
    g = max(gs for gs, _ in recs)
    return [v for gs, v in recs if gs == g]


def traffic(fetch_db, write_db, kernels, copy_bytes, units=None):
    """HBM bytes per launch of the stage made of `kernels`; with units, per
    unit of work instead of per dispatch (e.g. per grid level when a level's
    stage is one or two dispatches)"""
    f_copy = _probe_copies(fetch_db, "FETCH_SIZE")
    w_copy = _probe_copies(write_db, "WRITE_SIZE")
    if not f_copy or not w_copy:
        raise SystemExit("no copy-kernel dispatches in the databases: cannot calibrate")
    # KiB read for the copy kernel vs the bytes it moves
    f_scale = copy_bytes / (1024.0 * statistics.median(f_copy))
    w_scale = copy_bytes / (1024.0 * statistics.median(w_copy))
    f_k = _per_dispatch(fetch_db, kernels, "FETCH_SIZE")
    w_k = _per_dispatch(write_db, kernels, "WRITE_SIZE")
    if not f_k or not w_k:
        raise SystemExit(f"no dispatches of {kernels} with FETCH_SIZE / WRITE_SIZE")
    fetch = 1024.0 * f_scale * sum(f_k) / (units or len(f_k))
    write = 1024.0 * w_scale * sum(w_k) / (units or len(w_k))
    return {"kernel": kernels if isinstance(kernels, str) else " + ".join(kernels), "dispatches": len(f_k), "fetch_bytes_per_launch": fetch,
            "write_bytes_per_launch": write, "bytes_per_launch": fetch + write,
            "fetch_scale": f_scale, "write_scale": w_scale,
            "copy_fetch_KiB": statistics.median(f_copy), "copy_write_KiB": statistics.median(w_copy),
            "copy_bytes": copy_bytes}


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m ahf_amd.rocprof_summary")
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("db")
    a.add_argument("--top", type=int, default=30)
    b = sub.add_parser("counters")
    b.add_argument("db")
    b.add_argument("--kernel", required=True)
    c = sub.add_parser("traffic")
    c.add_argument("--fetch", required=True)
    c.add_argument("--write", required=True)
    c.add_argument("--kernel", required=True, action="append",
                   help="exact kernel name; repeat for a stage made of several kernels")
    c.add_argument("--copy-bytes", type=int, required=True)
    c.add_argument("--units", type=int, default=None,
                   help="divide by this many launches (the stage's timed regions) instead of the dispatch count")
    c.add_argument("--workload", required=True)
    c.add_argument("--source", default="")
    c.add_argument("--out")
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        print(f"{'kernel':48s} {'calls':>6s} {'total_ms':>10s} {'avg_us':>10s} {'min_us':>10s} {'max_us':>10s} {'%':>6s}")
        for r in kernel_stats(args.db)[:args.top]:
            print(f"{r['kernel'][:48]:48s} {r['calls']:6d} {r['total_ms']:10.3f} {r['avg_us']:10.2f} "
                  f"{r['min_us']:10.2f} {r['max_us']:10.2f} {r['pct']:6.1f}")
    elif args.cmd == "counters":
        for did, (k, gs, dur, cs) in dispatch_counters(args.db, args.kernel).items():
            vals = " ".join(f"{n}={v:.0f}" for n, v in sorted(cs.items()))
            print(f"{did} {k} grid={gs} dur_us={dur / 1e3 if dur else 0:.1f} {vals}")
    else:
        t = traffic(args.fetch, args.write, args.kernel, args.copy_bytes, args.units)
        t["workload"] = args.workload
        t["source"] = args.source
        text = json.dumps(t, indent=1)
        if args.out:
            with open(args.out, "w") as f:
                f.write(text + "\n")
        print(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
