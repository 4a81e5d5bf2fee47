"""Memory-pattern probe of the in-place AA sweep: what rate the sweep's population traffic
reaches without the collision, per storage precision, layout and AA form.

    python -m aerokinetic_amd.pattern_probe --shape 1536,1024,64 --reps 10
    python -m aerokinetic_amd.pattern_probe --placements 6 --tiles 64,256 --waves 0,4,3

Each line is one launch pattern of `aerokinetic_amd/lib/libaerokinetic_probe.so`
(csrc/ak_probe.hip): 27 loads and 27 stores per cell, even form (read (x, q), write
(x, opp q)) or odd form (read (x - c_q, opp q), write (x + c_q, q)), fully periodic box,
64 x 4 tiles in the sweep's XCD-contiguous block order.  Layouts: `rows` (interleaved
direction rows, the sweep's layout), `planes` (one plane per direction), `xblocks64` (64
cells along x with their 27 directions together), `blocks1024` (1024 consecutive cells with
their 27 directions together); `rows x2` / `x4` are the even form with 2 or 4 cells per lane
(one 4- or 8-byte access per direction).  Rates are algorithmic bytes (2 x 27 x element bytes
per cell) per second; `mean` is the rate of one even plus one odd launch, which is what a
run of the sweep alternates.  Prints a markdown table (and one JSON line with --json).

--placements N times the pattern (rows layout with the sweep's own row pitch) on N buffers
allocated and held at once, for each tile and cap on resident waves per SIMD: the same
launch on buffers of other physical pages, and under fewer resident waves, moves a
different number of bytes per second.

--offsets O1,O2,... times it on --buffers buffers held at once with the box starting at each
byte offset into the buffer: whether a buffer's rate follows where in its pages the data starts.
"""
from __future__ import annotations

import argparse
import ctypes
import json
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libaerokinetic_probe.so"
LAYOUTS = {"rows": 0, "planes": 1, "xblocks64": 2, "blocks1024": 3}
BYTES = {"fp64": 8, "fp32": 4, "fp16": 2}
PEAK_GBS = 8000.0


def _lib():
    if not LIB_PATH.exists():
        raise FileNotFoundError(f"{LIB_PATH} is missing: build it with make -C aerokinetic_amd/csrc")
    L = ctypes.CDLL(str(LIB_PATH))
    L.akp_pattern.argtypes = [ctypes.c_int] * 8 + [ctypes.POINTER(ctypes.c_double)]
    L.akp_pattern.restype = ctypes.c_int
    L.akp_last_error.restype = ctypes.c_char_p
    L.akp_placements.argtypes = [ctypes.c_int] * 8 + [ctypes.POINTER(ctypes.c_double)]
    L.akp_placements.restype = ctypes.c_int
    L.akp_offsets.argtypes = [ctypes.c_int] * 7 + [ctypes.POINTER(ctypes.c_longlong), ctypes.c_int,
                                                   ctypes.POINTER(ctypes.c_double)]
    L.akp_offsets.restype = ctypes.c_int
    L.akp_vmm.argtypes = [ctypes.c_int] * 6 + [ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
                                               ctypes.POINTER(ctypes.c_double)]
    L.akp_vmm.restype = ctypes.c_int
    L.akp_screen.argtypes = [ctypes.c_int] * 7 + [ctypes.c_longlong, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.akp_screen.restype = ctypes.c_int
    L.akp_orders.argtypes = [ctypes.c_int] * 7 + [ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                                  ctypes.POINTER(ctypes.c_double)]
    L.akp_orders.restype = ctypes.c_int
    L.akp_affinity.argtypes = [ctypes.c_longlong, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.POINTER(ctypes.c_double)]
    L.akp_affinity.restype = ctypes.c_int
    return L


def main_affinity(a, L):
    """Read GB/s of each XCD alone from each region of one buffer."""
    n = a.affinity
    out = (ctypes.c_double * (8 * n))()
    got = L.akp_affinity(a.piece_bytes, n, a.xcd_blocks, a.reps, out)
    if got < 0:
        raise RuntimeError(f"affinity probe: {L.akp_last_error().decode()}")
    print("| region | " + " | ".join(f"XCD {k}" for k in range(8)) + " |", flush=True)
    print("|---:|" + "---:|" * 8, flush=True)
    rows = []
    for r in range(got):
        v = [out[r * 8 + k] for k in range(8)]
        rows.append(v)
        print(f"| {r} | " + " | ".join(f"{x:.0f}" for x in v) + " |", flush=True)
    if a.json:
        print(json.dumps({"region_bytes": a.piece_bytes, "affinity_gbs": rows}))
    return 0


def main_orders(a, L, shape):
    """Table of the pattern on several buffers under several block orders (0: one contiguous slab per
    XCD, -1: hardware order, C: chunks of C tiles dealt to the XCDs in turn)."""
    orders = [int(v) for v in a.orders.split(",")]
    print("| storage | buffer | order | even TB/s | odd TB/s | mean % of 8 TB/s |", flush=True)
    print("|---|---:|---:|---:|---:|---:|", flush=True)
    rows = []
    for st in [s for s in a.storages.split(",") if s]:
        n = len(orders)
        arr = (ctypes.c_int * n)(*orders)
        out = (ctypes.c_double * (2 * n * a.buffers))()
        got = L.akp_orders(BYTES[st], *shape, a.tiles_bx, a.buffers, n, arr, a.reps, out)
        if got < 0:
            raise RuntimeError(f"order probe {st}: {L.akp_last_error().decode()}")
        nbytes = 2.0 * 27 * BYTES[st] * shape[0] * shape[1] * shape[2]
        for k in range(got):
            for j, o in enumerate(orders):
                ge = nbytes / (out[(k * n + j) * 2] * 1e-3) / 1e9
                go = nbytes / (out[(k * n + j) * 2 + 1] * 1e-3) / 1e9
                mean = 2 * ge * go / (ge + go)
                rows.append({"storage": st, "buffer": k, "order": o, "even_gbs": ge, "odd_gbs": go})
                print(f"| {st} | {k} | {o} | {ge / 1e3:.2f} | {go / 1e3:.2f} | {mean / PEAK_GBS * 100:.1f} |", flush=True)
    if a.json:
        print(json.dumps({"shape": shape, "reps": a.reps, "orders": rows}))
    return 0


def main_screen(a, L, shape):
    """Per-piece rates of --screen physical pieces, and buffers made of the fastest / slowest pieces."""
    rows = {}
    for st in [s for s in a.storages.split(",") if s]:
        n = a.screen
        out = (ctypes.c_double * (2 * n + 4))()
        got = L.akp_screen(BYTES[st], shape[0], shape[1], a.nz_piece, shape[2], a.tiles_bx, n, a.piece_bytes, a.reps,
                           out)
        if got < 0:
            raise RuntimeError(f"screen probe {st}: {L.akp_last_error().decode()}")
        pb = 2.0 * 27 * BYTES[st] * shape[0] * shape[1] * a.nz_piece
        fb = 2.0 * 27 * BYTES[st] * shape[0] * shape[1] * shape[2]
        mean = lambda nbytes, te, to: 2 * nbytes / ((te + to) * 1e-3) / 1e9 / PEAK_GBS * 100
        pieces = [mean(pb, out[2 * i], out[2 * i + 1]) for i in range(got)]
        fast = mean(fb, out[2 * n], out[2 * n + 1]) if out[2 * n] > 0 else None
        slow = mean(fb, out[2 * n + 2], out[2 * n + 3]) if out[2 * n + 2] > 0 else None
        srt = sorted(pieces)
        print(f"{st}: {got} pieces of {a.piece_bytes} bytes; piece rates % of 8 TB/s: min {srt[0]:.1f} "
              f"p25 {srt[len(srt) // 4]:.1f} median {srt[len(srt) // 2]:.1f} p75 {srt[3 * len(srt) // 4]:.1f} "
              f"max {srt[-1]:.1f}", flush=True)
        print(f"{st}: in allocation order: " + " ".join(f"{v:.0f}" for v in pieces), flush=True)
        print(f"{st}: buffer of the fastest pieces {fast if fast is None else round(fast, 1)} %, of the slowest "
              f"{slow if slow is None else round(slow, 1)} %", flush=True)
        rows[st] = {"pieces": pieces, "fastest_buffer_pct": fast, "slowest_buffer_pct": slow}
    if a.json:
        print(json.dumps({"shape": shape, "piece_bytes": a.piece_bytes, "nz_piece": a.nz_piece, "screen": rows}))
    return 0


def main_vmm(a, L, shape):
    """Table of the pattern on buffers built from physical pieces of --vmm-chunk bytes, each reached
    through two address ranges mapped onto the same physical memory."""
    print("| storage | piece | buffer | range | even TB/s | odd TB/s | mean % of 8 TB/s |", flush=True)
    print("|---|---:|---:|---:|---:|---:|---:|", flush=True)
    rows = []
    for st in [s for s in a.storages.split(",") if s]:
        for chunk in [int(v) for v in a.vmm.split(",")]:
            out = (ctypes.c_double * (4 * a.buffers + 1))()
            got = L.akp_vmm(BYTES[st], *shape, a.tiles_bx, a.buffers, chunk, a.ranges, a.reps, out)
            if got < 0:
                raise RuntimeError(f"vmm probe {st}: {L.akp_last_error().decode()}")
            if got < a.buffers:
                print(f"(piece {chunk}: {got} of {a.buffers} buffers made: {L.akp_last_error().decode()})", flush=True)
            nbytes = 2.0 * 27 * BYTES[st] * shape[0] * shape[1] * shape[2]
            for k in range(got):
                for r in range(a.ranges):
                    ge = nbytes / (out[4 * k + 2 * r] * 1e-3) / 1e9
                    go = nbytes / (out[4 * k + 2 * r + 1] * 1e-3) / 1e9
                    mean = 2 * ge * go / (ge + go)
                    rows.append({"storage": st, "chunk": chunk, "buffer": k, "range": r, "even_gbs": ge, "odd_gbs": go,
                                 "granularity": out[4 * a.buffers]})
                    print(f"| {st} | {chunk or 'whole'} | {k} | {r} | {ge / 1e3:.2f} | {go / 1e3:.2f} | "
                          f"{mean / PEAK_GBS * 100:.1f} |", flush=True)
    if a.json:
        print(json.dumps({"shape": shape, "reps": a.reps, "vmm": rows}))
    return 0


def offsets(L, storage, shape, bx, nbuf, offs, reps):
    """[[(even GB/s, odd GB/s) per offset] per buffer]: the pattern on `nbuf` buffers held at
    once, the box starting `offs[j]` bytes into each."""
    n = len(offs)
    arr = (ctypes.c_longlong * n)(*offs)
    out = (ctypes.c_double * (2 * n * nbuf))()
    got = L.akp_offsets(BYTES[storage], *shape, bx, nbuf, n, arr, reps, out)
    if got < 0:
        raise RuntimeError(f"offset probe {storage}: {L.akp_last_error().decode()}")
    nbytes = 2.0 * 27 * BYTES[storage] * shape[0] * shape[1] * shape[2]
    return [[(nbytes / (out[(k * n + j) * 2] * 1e-3) / 1e9, nbytes / (out[(k * n + j) * 2 + 1] * 1e-3) / 1e9)
             for j in range(n)] for k in range(got)]


def main_offsets(a, L, shape):
    """Table of the sweep pattern's rate on several buffers, each with the box at several offsets."""
    offs = [int(v) for v in a.offsets.split(",")]
    print("| storage | buffer | offset | even TB/s | odd TB/s | mean % of 8 TB/s |", flush=True)
    print("|---|---:|---:|---:|---:|---:|", flush=True)
    rows = []
    for st in [s for s in a.storages.split(",") if s]:
        for k, per in enumerate(offsets(L, st, shape, a.tiles_bx, a.buffers, offs, a.reps)):
            for off, (ge, go) in zip(offs, per):
                mean = 2 * ge * go / (ge + go)
                rows.append({"storage": st, "buffer": k, "offset": off, "even_gbs": ge, "odd_gbs": go})
                print(f"| {st} | {k} | {off} | {ge / 1e3:.2f} | {go / 1e3:.2f} | {mean / PEAK_GBS * 100:.1f} |",
                      flush=True)
    if a.json:
        print(json.dumps({"shape": shape, "reps": a.reps, "offsets": rows}))
    return 0


def placements(L, storage, shape, bx, waves, nbuf, reps):
    """[(even GB/s, odd GB/s)] of `nbuf` population buffers held at once (rows layout with the
    sweep's row pitch, bx x 256/bx tiles, resident-wave cap `waves`, 0 = none)."""
    out = (ctypes.c_double * (2 * nbuf + 1))()
    got = L.akp_placements(BYTES[storage], *shape, bx, waves, nbuf, reps, out)
    if got < 0:
        raise RuntimeError(f"placement probe {storage}: {L.akp_last_error().decode()}")
    nbytes = out[2 * nbuf]
    return [(nbytes / (out[2 * k] * 1e-3) / 1e9, nbytes / (out[2 * k + 1] * 1e-3) / 1e9) for k in range(got)]


def main_placements(a, L, shape):
    """Table of the sweep pattern's rate on several buffers held at once, per tile and wave cap."""
    print("| storage | tile | waves/SIMD cap | buffer | even TB/s | odd TB/s | mean TB/s | mean % of 8 TB/s |", flush=True)
    print("|---|---|---:|---:|---:|---:|---:|---:|", flush=True)
    rows = []
    for st in [s for s in a.storages.split(",") if s]:
        for bx in [int(v) for v in a.tiles.split(",")]:
            for w in [int(v) for v in a.waves.split(",")]:
                for k, (ge, go) in enumerate(placements(L, st, shape, bx, w, a.placements, a.reps)):
                    mean = 2 / (1 / ge + 1 / go)
                    rows.append({"storage": st, "tile": f"{bx}x{256 // bx}", "waves": w, "buffer": k,
                                 "even_gbs": ge, "odd_gbs": go, "mean_gbs": mean})
                    print(f"| {st} | {bx}x{256 // bx} | {w or '-'} | {k} | {ge / 1e3:.2f} | {go / 1e3:.2f} | "
                          f"{mean / 1e3:.2f} | {mean / PEAK_GBS * 100:.1f} |", flush=True)
    if a.json:
        print(json.dumps({"shape": shape, "reps": a.reps, "placement_rows": rows}))
    return 0


def measure(L, layout, storage, odd, shape, reps, cells_per_lane=1):
    """(ms per launch, GB/s) of one pattern."""
    out = (ctypes.c_double * 2)()
    rc = L.akp_pattern(LAYOUTS[layout], BYTES[storage], 1 if odd else 0, cells_per_lane, *shape, reps, out)
    if rc:
        raise RuntimeError(f"probe {layout} {storage} odd={odd}: {L.akp_last_error().decode()}")
    return out[0], out[1] / (out[0] * 1e-3) / 1e9


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--shape", default="1536,1024,64")
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--storages", default="fp64,fp32,fp16")
    ap.add_argument("--json", action="store_true")
    ap.add_argument("--placements", type=int, default=0,
                    help="instead of the layout table: this many buffers held at once, each timed (rows layout, "
                         "the sweep's row pitch) for every tile of --tiles and wave cap of --waves")
    ap.add_argument("--tiles", default="64,256", help="--placements: tile widths bx (tiles bx x 256/bx)")
    ap.add_argument("--waves", default="0,4,3", help="--placements: resident-wave caps per SIMD (0: none)")
    ap.add_argument("--offsets", default="", help="instead of the layout table: byte offsets (multiples of 256) at "
                    "which the box starts in each of --buffers buffers held at once (tiles --tile-bx x 256/bx)")
    ap.add_argument("--buffers", type=int, default=4)
    ap.add_argument("--affinity", type=int, default=0, help="instead of the layout table: read GB/s of each XCD "
                    "alone from each of this many --piece-bytes regions of one buffer")
    ap.add_argument("--xcd-blocks", type=int, default=1024, help="--affinity: blocks per XCD")
    ap.add_argument("--orders", default="", help="instead of the layout table: block orders (0 one slab per XCD, -1 "
                    "hardware order, C chunks of C tiles) timed on --buffers buffers held at once")
    ap.add_argument("--screen", type=int, default=0, help="instead of the layout table: this many physical pieces "
                    "timed one by one (--nz-piece planes each), then buffers of the fastest and the slowest pieces")
    ap.add_argument("--piece-bytes", type=int, default=1 << 30)
    ap.add_argument("--nz-piece", type=int, default=3)
    ap.add_argument("--ranges", type=int, default=2, help="--vmm: address ranges mapped onto each buffer (1 or 2)")
    ap.add_argument("--vmm", default="", help="instead of the layout table: piece sizes in bytes (0: one piece) of "
                    "--buffers buffers built with the virtual-memory interface, each mapped twice")
    ap.add_argument("--tile-bx", dest="tiles_bx", type=int, default=256)
    a = ap.parse_args(argv)
    shape = tuple(int(v) for v in a.shape.split(","))
    L = _lib()
    if a.offsets:
        return main_offsets(a, L, shape)
    if a.vmm:
        return main_vmm(a, L, shape)
    if a.screen:
        return main_screen(a, L, shape)
    if a.orders:
        return main_orders(a, L, shape)
    if a.affinity:
        return main_affinity(a, L)
    if a.placements:
        return main_placements(a, L, shape)
    rows = []
    print("| storage | layout | even TB/s | odd TB/s | mean TB/s | mean % of 8 TB/s |", flush=True)
    print("|---|---|---:|---:|---:|---:|", flush=True)
    for st in [s for s in a.storages.split(",") if s]:
        for lay in LAYOUTS:
            te, ge = measure(L, lay, st, False, shape, a.reps)
            to, go = measure(L, lay, st, True, shape, a.reps)
            mean = 2 * ge * te / (te + to)  # two launches' bytes over their summed time
            rows.append({"storage": st, "layout": lay, "even_gbs": ge, "odd_gbs": go, "mean_gbs": mean})
            print(f"| {st} | {lay} | {ge / 1e3:.2f} | {go / 1e3:.2f} | {mean / 1e3:.2f} | {mean / PEAK_GBS * 100:.1f} |",
                  flush=True)
        for k in (2, 4):
            if BYTES[st] * k > 8:
                continue
            te, ge = measure(L, "rows", st, False, shape, a.reps, cells_per_lane=k)
            rows.append({"storage": st, "layout": f"rows x{k}", "even_gbs": ge})
            print(f"| {st} | rows x{k} (even only) | {ge / 1e3:.2f} | — | — | — |", flush=True)
    if a.json:
        print(json.dumps({"shape": shape, "reps": a.reps, "rows": rows}))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
