"""Airfoil contours, sections and the NACA 4-digit generator.

Behaviour follows the reference case model (`src/aerokinetic/geometry.py`):

* `Contour` cleans a closed polygon, rejects degenerate or self-intersecting
  input (geometry.py:11-49),
* `Contour.contains` is the even-odd ray test used for solid masks
  (geometry.py:58-68),
* `Contour.link_fraction` returns the first wall intersection q of a
  fluid-to-solid link, clipped to [1e-10, 1] (geometry.py:70-85),
* `Section.normalized` rotates by -alpha about the moment reference and divides
  by the chord (geometry.py:143-148),
* `naca4` is the cosine-spaced 4-digit section with the open (-0.1015) or closed
  (-0.1036) trailing-edge coefficient (geometry.py:155-185).

`naca4_profile` additionally accepts fractional camber/thickness values for
sampled polar sweeps; `naca4` keeps the reference's integer-code contract.
"""
from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np


def _cross(a, b):
    return a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]


def segments_intersect(a, b, c, d) -> bool:
    """Closed-segment intersection with a relative collinearity tolerance."""
    a, b, c, d = (np.asarray(v, dtype=float) for v in (a, b, c, d))
    ab, cd = b - a, d - c
    den = _cross(ab, cd)
    tol = 1e-12 * max(float(np.linalg.norm(ab) * np.linalg.norm(cd)), 1e-30)
    if abs(den) <= tol:
        if abs(_cross(c - a, ab)) > tol:
            return False
        k = int(np.argmax(np.abs(ab)))
        lo = max(min(a[k], b[k]), min(c[k], d[k]))
        hi = min(max(a[k], b[k]), max(c[k], d[k]))
        return lo <= hi + 1e-12
    t = _cross(c - a, cd) / den
    s = _cross(c - a, ab) / den
    return (-1e-12 <= t <= 1 + 1e-12) and (-1e-12 <= s <= 1 + 1e-12)


def _segment_hits(a, b, C, D):
    """segments_intersect(a, b, c, d) elementwise over broadcast segment arrays (same tolerances)."""
    ab, cd, ca = b - a, D - C, C - a
    den = _cross(ab, cd)
    tol = 1e-12 * np.maximum(np.linalg.norm(ab, axis=-1) * np.linalg.norm(cd, axis=-1), 1e-30)
    par = np.abs(den) <= tol
    with np.errstate(divide="ignore", invalid="ignore"):
        t = _cross(ca, cd) / den
        s = _cross(ca, ab) / den
    hit = ~par & (t >= -1e-12) & (t <= 1 + 1e-12) & (s >= -1e-12) & (s <= 1 + 1e-12)
    col = par & (np.abs(_cross(ca, ab)) <= tol)
    if np.any(col):
        shape = np.broadcast_shapes(np.shape(a), np.shape(b), np.shape(C), np.shape(D))
        a_, b_, c_, d_ = (np.broadcast_to(v, shape) for v in (a, b, C, D))
        k = np.argmax(np.abs(np.broadcast_to(ab, shape)), axis=-1)[..., None]
        pick = lambda v: np.take_along_axis(v, k, axis=-1)[..., 0]
        lo = np.maximum(np.minimum(pick(a_), pick(b_)), np.minimum(pick(c_), pick(d_)))
        hi = np.minimum(np.maximum(pick(a_), pick(b_)), np.maximum(pick(c_), pick(d_)))
        hit = hit | (col & (lo <= hi + 1e-12))
    return hit


@dataclass(frozen=True)
class Contour:
    """A closed polygon in a common frame (any length unit)."""

    xy: np.ndarray
    name: str = "element"

    def __post_init__(self):
        pts = np.array(self.xy, dtype=float, copy=True)
        if pts.ndim != 2 or pts.shape[1] != 2 or len(pts) < 3 or not np.isfinite(pts).all():
            raise ValueError("A contour requires at least three finite (x,y) points")
        step = np.linalg.norm(np.diff(pts, axis=0), axis=1)
        pts = pts[np.concatenate(([True], step > 1e-14))]
        if len(pts) > 1 and np.allclose(pts[0], pts[-1], rtol=0, atol=1e-14):
            pts = pts[:-1]
        if len(pts) < 3:
            raise ValueError("Degenerate contour")
        nxt = np.roll(pts, -1, axis=0)
        area = abs(float(np.sum(_cross(pts, nxt)))) / 2
        if area <= 1e-12 * float(np.ptp(pts, axis=0).max()) ** 2:
            raise ValueError("Contour has zero enclosed area")
        n = len(pts)
        ii, jj = np.triu_indices(n, k=2)  # non-adjacent edge pairs (first and last edges meet)
        keep = ~((ii == 0) & (jj == n - 1))
        ii, jj = ii[keep], jj[keep]
        for s0 in range(0, len(ii), 1 << 20):
            i, j = ii[s0:s0 + (1 << 20)], jj[s0:s0 + (1 << 20)]
            if _segment_hits(pts[i], nxt[i], pts[j], nxt[j]).any():
                raise ValueError(f"Self-intersection in contour {self.name}")
        pts.setflags(write=False)
        object.__setattr__(self, "xy", pts)

    @classmethod
    def from_dat(cls, path):
        """Read one Selig-style polygon (lines with exactly two numbers)."""
        rows = []
        for line in Path(path).read_text(encoding="utf-8-sig").splitlines():
            parts = line.replace(",", " ").split()
            if len(parts) != 2:
                continue
            try:
                rows.append((float(parts[0]), float(parts[1])))
            except ValueError:
                continue
        return cls(np.asarray(rows), Path(path).stem)

    @classmethod
    def from_aero2d(cls, airfoil):
        return cls(np.column_stack((airfoil.x, airfoil.y)), getattr(airfoil, "name", "element"))

    def contains(self, x, y):
        """Even-odd crossing test, vectorised over the query points."""
        x, y = np.broadcast_arrays(np.asarray(x, dtype=float), np.asarray(y, dtype=float))
        inside = np.zeros(x.shape, dtype=bool)
        p = self.xy
        q = np.roll(p, -1, axis=0)
        for (ax, ay), (bx, by) in zip(p, q):
            if by == ay:
                continue
            crosses = (ay > y) != (by > y)
            xs = (bx - ax) * (y - ay) / (by - ay) + ax
            inside ^= crosses & (x < xs)
        return inside

    def link_fraction(self, points, offsets):
        """Smallest q in [0,1] with points + q*offsets on the polygon boundary."""
        points = np.asarray(points, dtype=float)
        offsets = np.asarray(offsets, dtype=float)
        best = np.full(len(points), np.inf)
        a = self.xy
        edge = np.roll(a, -1, axis=0) - a
        step = max(1, (1 << 22) // max(len(a), 1))  # points per block of the (point, edge) table
        for i in range(0, len(points), step):
            pt, off = points[i:i + step, None, :], offsets[i:i + step, None, :]
            den = _cross(off, edge[None])
            ok = np.abs(den) > 1e-14
            rel = a[None] - pt
            with np.errstate(divide="ignore", invalid="ignore"):
                t = np.where(ok, _cross(rel, edge[None]) / den, np.inf)
                s = np.where(ok, _cross(rel, off) / den, np.inf)
            hit = ok & (t >= -1e-10) & (t <= 1 + 1e-10) & (s >= -1e-10) & (s <= 1 + 1e-10)
            best[i:i + step] = np.where(hit, t, np.inf).min(axis=1)
        if not np.isfinite(best).all():
            raise ValueError("A fluid-to-solid link has no geometric intersection")
        return np.clip(best, 1e-10, 1.0)


@dataclass(frozen=True)
class Section:
    """One or more non-touching contours with a reference chord and moment point."""

    elements: tuple
    chord: float | None = None
    moment_reference: tuple | None = None

    def __post_init__(self):
        elements = tuple(self.elements)
        if not elements:
            raise ValueError("At least one contour is required")
        if not all(isinstance(e, Contour) for e in elements):
            raise TypeError("Section elements must be Contour objects")
        object.__setattr__(self, "elements", elements)
        pts = np.concatenate([e.xy for e in elements])
        le = pts[np.argmin(pts[:, 0])]
        aft = pts[np.isclose(pts[:, 0], pts[:, 0].max(), rtol=0, atol=1e-12)]
        te = aft.mean(axis=0)
        chord = float(np.linalg.norm(te - le)) if self.chord is None else float(self.chord)
        if not np.isfinite(chord) or chord <= 0:
            raise ValueError("Reference chord must be positive")
        if self.moment_reference is None:
            ref = le + 0.25 * (te - le)
        else:
            ref = np.asarray(self.moment_reference, dtype=float)
        if np.shape(ref) != (2,) or not np.isfinite(ref).all():
            raise ValueError("Moment reference requires two finite coordinates")
        object.__setattr__(self, "chord", chord)
        object.__setattr__(self, "moment_reference", tuple(float(v) for v in ref))
        for i, e in enumerate(elements):
            for f in elements[i + 1:]:
                touching = bool(e.contains(*f.xy[0])) or bool(f.contains(*e.xy[0]))
                if not touching:
                    ea, eb = e.xy, np.roll(e.xy, -1, axis=0)
                    fa, fb = f.xy, np.roll(f.xy, -1, axis=0)
                    touching = bool(_segment_hits(ea[:, None], eb[:, None], fa[None], fb[None]).any())
                if touching:
                    raise ValueError(f"Elements touch or overlap: {e.name}, {f.name}")

    @classmethod
    def from_aero2d(cls, airfoils, **kwargs):
        if hasattr(airfoils, "x"):
            airfoils = [airfoils]
        return cls(tuple(Contour.from_aero2d(a) for a in airfoils), **kwargs)

    def rotation(self, alpha_deg):
        """Matrix that turns the section by -alpha (positive alpha raises the nose)."""
        ang = -np.deg2rad(alpha_deg)
        c, s = np.cos(ang), np.sin(ang)
        return np.array([[c, -s], [s, c]])

    def normalized(self, alpha_deg):
        """Contours in chord units with the moment reference at the origin."""
        rot = self.rotation(alpha_deg)
        ref = np.asarray(self.moment_reference)
        return tuple(Contour(((e.xy - ref) / self.chord) @ rot.T, e.name) for e in self.elements)

    def leading_edge_normalized(self, alpha_deg):
        """Foremost input point mapped into the normalized frame (seed anchor)."""
        pts = np.concatenate([e.xy for e in self.elements])
        le = pts[np.argmin(pts[:, 0])]
        return self.rotation(alpha_deg) @ ((le - np.asarray(self.moment_reference)) / self.chord)


def naca4_profile(camber, position, thickness, n_side=161, closed_te=False, name=None):
    """Cosine-spaced NACA 4-digit polygon for real-valued m, p, t (chord 1)."""
    if thickness <= 0 or (camber > 0 and position <= 0):
        raise ValueError("NACA thickness must be positive; cambered profiles require a nonzero second digit")
    if not isinstance(n_side, int):
        raise ValueError("n_side must be an integer")
    if n_side < 12:
        raise ValueError("Use at least 12 points per side")
    x = (1 - np.cos(np.linspace(0, np.pi, n_side))) / 2
    a4 = -0.1036 if closed_te else -0.1015
    half = 5 * thickness * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x**2 + 0.2843 * x**3 + a4 * x**4)
    yc = np.zeros_like(x)
    dyc = np.zeros_like(x)
    if camber > 0:
        m, p = camber, position
        fore = x < p
        yc[fore] = m / p**2 * (2 * p * x[fore] - x[fore] ** 2)
        dyc[fore] = 2 * m / p**2 * (p - x[fore])
        aft = ~fore
        yc[aft] = m / (1 - p) ** 2 * (1 - 2 * p + 2 * p * x[aft] - x[aft] ** 2)
        dyc[aft] = 2 * m / (1 - p) ** 2 * (p - x[aft])
    th = np.arctan(dyc)
    upper = np.column_stack((x - half * np.sin(th), yc + half * np.cos(th)))
    lower = np.column_stack((x + half * np.sin(th), yc - half * np.cos(th)))
    return Contour(np.concatenate((upper[::-1], lower[1:])), name or "NACA")


def naca4(code="0012", n_side=161, closed_te=False):
    """Four-digit NACA section from its integer code (e.g. "2412")."""
    code = str(code).strip()
    if len(code) != 4 or not code.isascii() or not code.isdigit():
        raise ValueError("NACA code must contain four digits, e.g. 0012 or 2412")
    m, p, t = int(code[0]) / 100, int(code[1]) / 10, int(code[2:]) / 100
    return naca4_profile(m, p, t, n_side, closed_te, f"NACA{code}")


def naca0012(n_side=161, closed_te=False):
    return naca4("0012", n_side, closed_te)
