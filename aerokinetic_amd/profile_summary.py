"""Per-kernel durations from a rocprofv3 kernel trace, over all dispatches and over the last ones.

    python -m aerokinetic_amd.profile_summary <out>_kernel_trace.csv --last 10 --match k_sweep

A benchmark that tunes its launches first (several population buffers, tiles and wave caps,
all dispatching the same kernels) and times its launches last leaves the timed dispatches at
the end of each kernel's list: `--last N` averages the final N dispatches of every kernel
name, which is what the benchmark's own event timing of its last N launches per AA form
measures.  Prints a markdown table (durations in microseconds, dispatch order = start time).

    python -m aerokinetic_amd.profile_summary <out>_counter_collection.csv --counters --match k_sweep

reads a `rocprofv3 --pmc` counter file instead: per kernel name, the mean duration and the SQ
instruction counters per wave (counter / SQ_WAVES, summed over the chip).  When the pass holds
SQ_BUSY_CYCLES (summed over the shader engines: 8 XCDs x 4 on MI355X) it also prints the clock
that implies (busy cycles / engines / duration) and, from it, the share of the CU's issue slots
the vector and scalar instructions fill: one wave64 VALU instruction per CU cycle (four SIMDs,
four cycles each) and one scalar instruction per CU cycle.
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


def counters(path):
    """{kernel name: [{counter: value, "us": duration}, one per dispatch in start order]} of a
    rocprofv3 counter-collection CSV (one row per dispatch and counter)."""
    per = {}
    with open(path, newline="", encoding="utf-8") as fh:
        for r in csv.DictReader(fh):
            d = per.setdefault(r["Kernel_Name"], {}).setdefault(int(r["Dispatch_Id"]), {})
            d[r["Counter_Name"]] = float(r["Counter_Value"])
            d["us"] = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
            d["start"] = int(r["Start_Timestamp"])
    return {k: [v for _, v in sorted(ds.items(), key=lambda kv: kv[1]["start"])] for k, ds in per.items()}


def issue_rows(path, match="", cus=256, engines=32):
    """Per kernel name: dispatches, mean duration (us), SQ_INSTS_* per wave, and with
    SQ_BUSY_CYCLES the implied clock (GHz) and the VALU / SALU issue shares of the CU's slots."""
    out = []
    for name, ds in counters(path).items():
        if match not in name or not all("SQ_WAVES" in d for d in ds):
            continue
        us = sum(d["us"] for d in ds) / len(ds)
        row = {"kernel": short(name), "dispatches": len(ds), "us": us}
        for c in sorted({c for d in ds for c in d if c.startswith("SQ_INSTS")}):
            row[c] = sum(d[c] / d["SQ_WAVES"] for d in ds) / len(ds)
        if all("SQ_BUSY_CYCLES" in d for d in ds):
            ghz = sum(d["SQ_BUSY_CYCLES"] / engines / (d["us"] * 1e3) for d in ds) / len(ds)
            row["ghz"] = ghz
            for c, key in (("SQ_INSTS_VALU", "valu_issue"), ("SQ_INSTS_SALU", "salu_issue")):
                if c in ds[0]:
                    row[key] = sum(d[c] / cus / (d["us"] * 1e3 * d["SQ_BUSY_CYCLES"] / engines / (d["us"] * 1e3))
                                   for d in ds) / len(ds)
        out.append(row)
    return out


def print_issue(rows):
    cols = [c for c in rows[0] if c not in ("kernel",)] if rows else []
    print("| kernel | " + " | ".join(cols) + " |")
    print("|---|" + "---:|" * len(cols))
    for r in rows:
        print(f"| `{r['kernel']}` | " + " | ".join(
            f"{r.get(c, float('nan')):.1%}" if c.endswith("_issue") else
            f"{r.get(c, float('nan')):.2f}" if c == "ghz" else
            f"{r[c]}" if c == "dispatches" else f"{r.get(c, float('nan')):.1f}" for c in cols) + " |")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("trace")
    ap.add_argument("--last", type=int, default=10, help="average of the final N dispatches of each kernel")
    ap.add_argument("--match", default="", help="only kernels whose name contains this")
    ap.add_argument("--counters", action="store_true", help="the file is a --pmc counter collection")
    ap.add_argument("--cus", type=int, default=256)
    ap.add_argument("--engines", type=int, default=32, help="shader engines SQ_BUSY_CYCLES is summed over")
    a = ap.parse_args(argv)
    if a.counters:
        print_issue(issue_rows(a.trace, a.match, a.cus, a.engines))
        return 0
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
