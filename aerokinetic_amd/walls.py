"""Solid masks and the compacted fluid-to-solid link list of the Bouzidi walls.

Semantics of the reference (src/aerokinetic/solver.py:93-122, walls.py:32-49,
geometry.py:58-85): a cell is solid when its centre lies inside a contour (even-odd
test, the geometry is extruded over the periodic span); a link (x, d) joins a
fluid interior cell x to a solid cell x + c_d of one element; its Bouzidi
fraction q is the first polygon crossing along c_d from the fluid cell centre,
computed once per (x, y, c_x, c_y) since it is independent of z and c_z; q < 1/2
with a solid second node x - c_d falls back to halfway bounce-back (q = -1).

The list is kept per xy link (the device kernel extrudes it over z), sorted by
element, then x, y, direction.
"""
from __future__ import annotations

import numpy as np

from .d3q27 import CX, CY


def labels(contours, shape_xy, origin, dx, ghost=1):
    """Element label (0 fluid, k+1 for element k) of every cell centre, with ghosts."""
    nx, ny = shape_xy
    gx = np.arange(-ghost, nx + ghost)
    gy = np.arange(-ghost, ny + ghost)
    xx = (gx + 0.5) * dx + origin[0]
    yy = (gy + 0.5) * dx + origin[1]
    lab = np.zeros((len(gx), len(gy)), dtype=np.int32)
    for k, c in enumerate(contours):
        inside = c.contains(xx[:, None], yy[None, :])
        lab[inside] = k + 1
    return lab


def wall_links(contours, shape_xy, origin, dx):
    """Link arrays x, y, d, q, element (per xy link) and the fallback count per z layer."""
    nx, ny = shape_xy
    lab = labels(contours, shape_xy, origin, dx)
    fluid = lab[1:-1, 1:-1] == 0
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    out = {k: [] for k in ("x", "y", "d", "q", "element")}
    fallback = 0
    for e, contour in enumerate(contours):
        for cxy in sorted({(int(CX[d]), int(CY[d])) for d in range(27)} - {(0, 0)}):
            cx, cy = cxy
            nb = lab[1 + cx:1 + cx + nx, 1 + cy:1 + cy + ny]
            sel = fluid & (nb == e + 1)
            if not sel.any():
                continue
            xs, ys = ix[sel], iy[sel]
            pts = np.column_stack(((xs + 0.5) * dx + origin[0], (ys + 0.5) * dx + origin[1]))
            off = np.tile(np.array([cx * dx, cy * dx]), (len(xs), 1))
            q = contour.link_fraction(pts, off)
            other = pts - off
            blocked = np.zeros(len(xs), dtype=bool)
            for c2 in contours:
                blocked |= c2.contains(other[:, 0], other[:, 1])
            fb = (q < 0.5) & blocked
            q = np.where(fb, -1.0, q)
            for d in range(27):
                if CX[d] == cx and CY[d] == cy:
                    out["x"].append(xs); out["y"].append(ys)
                    out["d"].append(np.full(len(xs), d)); out["q"].append(q)
                    out["element"].append(np.full(len(xs), e))
                    fallback += int(fb.sum())
    if not out["x"]:
        return {k: np.zeros(0, dtype=np.float64 if k == "q" else np.int32) for k in out}, 0
    res = {k: np.concatenate(v) for k, v in out.items()}
    order = np.lexsort((res["d"], res["y"], res["x"], res["element"]))
    res = {k: v[order] for k, v in res.items()}
    for k in ("x", "y", "d", "element"):
        res[k] = res[k].astype(np.int32)
    res["q"] = res["q"].astype(np.float64)
    return res, fallback


def extrude(links, nz):
    """Per-(x, y, z) link arrays (z fastest), for comparisons with z-resolved lists."""
    n = len(links["x"])
    z = np.tile(np.arange(nz), n)
    return {**{k: np.repeat(v, nz) for k, v in links.items()}, "z": z}
