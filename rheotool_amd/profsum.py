"""Summaries of rocprofv3 output for the stress-step kernels.

Two reports, both plain text for a profiles folder:

* ``stats``: the per-kernel table of a ``rocprofv3 --kernel-trace --stats``
  run (``*_kernel_stats.csv``), with kernel names shortened to the names the
  library's own timing uses (``k_kry_xr<f32>``, ``k_assemble<f64>`` ...).
* ``traffic``: HBM bytes per launch from two ``rocprofv3 --pmc`` runs
  (``FETCH_SIZE``, then ``WRITE_SIZE``, KiB per dispatch).  Each counter is
  scaled by the factor that makes it read the true size of a plain copy
  kernel profiled the same way (``k_copy``: known bytes read and written per
  launch).  Writes a JSON map kernel -> bytes per launch (what ``bench.py``
  reports as ``roofline.traffic``) and a Markdown table.

Usage::

    python -m rheotool_amd.profsum stats STATS_CSV --cmd "..." > out.md
    python -m rheotool_amd.profsum traffic --fetch F.csv --write W.csv \\
        --copy-fetch CF.csv --copy-write CW.csv --copy-bytes 2147483648 \\
        --json out.json --cmd "..." > out.md
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import re
import statistics
import sys


def short_name(full: str) -> str:
    """'void rt::k_residual<float, 1>(rt::KryArgs<float>)' -> 'k_residual<f32>'."""
    s = full.strip()
    if s.startswith("void "):
        s = s[5:]
    s = s.replace("rt::", "")
    # cut the argument list: the first '(' outside template brackets
    depth = 0
    for i, ch in enumerate(s):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            s = s[:i]
            break
    m = re.match(r"^([A-Za-z_0-9]+)<([^,>]+)", s)
    if m:
        t = {"float": "f32", "double": "f64"}.get(m.group(2).strip())
        return f"{m.group(1)}<{t}>" if t else m.group(1)
    return s


def read_counter(path):
    """kernel short name -> list of counter values (KiB), dispatch order."""
    d = collections.defaultdict(list)
    with open(path) as fh:
        for r in csv.DictReader(fh):
            d[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return d


def cmd_stats(a):
    rows = []
    with open(a.csv) as fh:
        for r in csv.DictReader(fh):
            rows.append((short_name(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]),
                         float(r["AverageNs"]), float(r["Percentage"]), float(r["MinNs"]),
                         float(r["MaxNs"])))
    # fold template instances that share a short name (e.g. both residual modes)
    agg = collections.OrderedDict()
    for n, calls, tot, _, pct, mn, mx in rows:
        g = agg.setdefault(n, [0, 0.0, 0.0, float("inf"), 0.0])
        g[0] += calls
        g[1] += tot
        g[2] += pct
        g[3] = min(g[3], mn)
        g[4] = max(g[4], mx)
    out = [f"# rocprofv3 kernel statistics: {a.title}", "", "Command:", "", "```", a.cmd, "```", "",
           "| kernel | calls | total ms | avg ms | % | min ms | max ms |",
           "|---|---:|---:|---:|---:|---:|---:|"]
    for n, (calls, tot, pct, mn, mx) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        out.append(f"| {n} | {calls} | {tot / 1e6:.3f} | {tot / calls / 1e6:.4f} | {pct:.2f} | "
                   f"{mn / 1e6:.4f} | {mx / 1e6:.4f} |")
    print("\n".join(out))


def cmd_traffic(a):
    fetch, write = read_counter(a.fetch), read_counter(a.write)
    cf = statistics.median(read_counter(a.copy_fetch)["k_copy"])
    cw = statistics.median(read_counter(a.copy_write)["k_copy"])
    f_fac = a.copy_bytes / (cf * 1024.0)
    w_fac = a.copy_bytes / (cw * 1024.0)
    res = {}
    lines = [f"# HBM traffic per launch: {a.title}", "", "Commands:", "", "```", a.cmd, "```", "",
             f"Calibration: `k_copy` moving {a.copy_bytes} B each way per launch reads "
             f"FETCH_SIZE = {cf:.1f} KiB and WRITE_SIZE = {cw:.1f} KiB, so FETCH_SIZE is scaled by "
             f"{f_fac:.3f} and WRITE_SIZE by {w_fac:.3f}.", "",
             "Bytes per launch are means over the profiled dispatches (launches that find every "
             "component converged do no work and are included, as in the timing).", "",
             "| kernel | dispatches | read GB/launch | write GB/launch | traffic GB/launch |",
             "|---|---:|---:|---:|---:|"]
    for n in sorted(set(fetch) & set(write)):
        if n.startswith("__amd") or n == "k_copy":
            continue
        fr = statistics.mean(fetch[n]) * 1024.0 * f_fac
        wr = statistics.mean(write[n]) * 1024.0 * w_fac
        if fr + wr <= 0:
            continue
        res[n] = round(fr + wr)
        lines.append(f"| {n} | {len(fetch[n])} | {fr / 1e9:.3f} | {wr / 1e9:.3f} | {(fr + wr) / 1e9:.3f} |")
    with open(a.json, "w") as fh:
        json.dump(dict(unit="bytes per launch", fetch_scale=f_fac, write_scale=w_fac, kernels=res), fh,
                  indent=1, sort_keys=True)
    print("\n".join(lines))


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m rheotool_amd.profsum")
    sp = ap.add_subparsers(dest="what", required=True)
    s = sp.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--cmd", required=True)
    s.add_argument("--title", default="")
    t = sp.add_parser("traffic")
    for k in ("fetch", "write", "copy_fetch", "copy_write", "json"):
        t.add_argument("--" + k.replace("_", "-"), required=True)
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--cmd", required=True)
    t.add_argument("--title", default="")
    a = ap.parse_args(argv)
    {"stats": cmd_stats, "traffic": cmd_traffic}[a.what](a)


if __name__ == "__main__":
    sys.exit(main())
