"""Per-kernel durations from a rocprofv3 kernel trace, over all dispatches and over the last ones.

    python -m aerokinetic_amd.profile_summary <out>_kernel_trace.csv --last 10 --match k_sweep

A benchmark that tunes its launches first (several population buffers, tiles and wave caps,
all dispatching the same kernels) and times its launches last leaves the timed dispatches at
the end of each kernel's list: `--last N` averages the final N dispatches of every kernel
name, which is what the benchmark's own event timing of its last N launches per AA form
measures.  Prints a markdown table (durations in microseconds, dispatch order = start time).
"""
from __future__ import annotations

import argparse
import csv
import sys
from collections import defaultdict


def _column(header, *names):
    low = [h.strip().lower() for h in header]
    for n in names:
        if n.lower() in low:
            return low.index(n.lower())
    raise KeyError(f"none of {names} in the trace header {header}")


def durations(path):
    """{kernel name: [duration in us, in start order]} of a rocprofv3 kernel-trace CSV."""
    with open(path, newline="", encoding="utf-8") as fh:
        rows = csv.reader(fh)
        header = next(rows)
        kn = _column(header, "Kernel_Name", "KernelName")
        st = _column(header, "Start_Timestamp", "BeginNs")
        en = _column(header, "End_Timestamp", "EndNs")
        per = defaultdict(list)
        for r in rows:
            if len(r) <= max(kn, st, en):
                continue
            per[r[kn]].append((int(r[st]), (int(r[en]) - int(r[st])) / 1e3))
    return {k: [d for _, d in sorted(v)] for k, v in per.items()}


def short(name):
    """Kernel name without the argument list: `ak::k_sweep<0, true, false, true>`."""
    depth, out = 0, []
    for ch in name:
        if ch == "(" and depth == 0:
            break
        depth += ch == "<"
        depth -= ch == ">"
        out.append(ch)
    s = "".join(out).strip()
    return s[len("void "):] if s.startswith("void ") else s


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("trace")
    ap.add_argument("--last", type=int, default=10, help="average of the final N dispatches of each kernel")
    ap.add_argument("--match", default="", help="only kernels whose name contains this")
    a = ap.parse_args(argv)
    per = durations(a.trace)
    print(f"| kernel | dispatches | mean us (all) | mean us (last {a.last}) | min–max us (last {a.last}) |")
    print("|---|---:|---:|---:|---:|")
    for name, d in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        if a.match not in name:
            continue
        tail = d[-a.last:]
        print(f"| `{short(name)}` | {len(d)} | {sum(d) / len(d):.1f} | {sum(tail) / len(tail):.1f} | "
              f"{min(tail):.1f}–{max(tail):.1f} |")
    return 0


if __name__ == "__main__":
    sys.exit(main())
