"""FP64 oracle of the two-level 2:1 airfoil step (test infrastructure only).

Restates the reference's refined path with its own data flow: two collide_stream
lattices (C oracle primitives ako_collide / ako_boundaries / ako_stream), the
NumPy transfer algebra and the coarse/fine interface orchestration.  Followed:

* transfer: moments, regularised reconstruction, stress rescaling, 5-point
  conservative prolongation, 2x2x2 restriction (src/aerokinetic/refinement.py:78-254,
  parallel.py:15-56 for the fused forms);
* face flux of pull streaming through a Cartesian face, split over axis orders
  (interface.py:44-104);
* coarse collide -> coarse face flux -> coarse boundaries + stream -> two fine
  substeps with time-interpolated donors -> reflux -> restriction of the fine
  state (interface.py:297-352 with the overrides of refined_solver.py:183-203,
  271-284);
* fine-wall loads averaged over the two substeps (refined_solver.py:311-345);
* initial field (refined_solver.py:224-269);
* the conservation budget: conserved integrals of the owned cells of both levels
  (refined_solver.py:291-299), the mass and momentum entering through the coarse
  domain's outer faces (the outer face-flux samplers over the ghost-padded coarse array
  after its boundaries, refined_solver.py:215-222, 271-280; both arrays start filled with
  the free stream, ghosts included, refined_solver.py:228-230), the wall mass exchange and
  impulse, `last_budget` and the cumulative sums (refined_solver.py:311-361).
"""
from __future__ import annotations

from itertools import combinations
import math

import numpy as np

from aerokinetic_amd.d3q27 import C as STENCIL
from aerokinetic_amd.refined_plan import RefinedPlan, virtual_core
from oracle import oracle as O

CS2 = 1.0 / 3.0
SECOND = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))
WEIGHTS = np.array([{0: 8 / 27, 1: 2 / 27, 2: 1 / 54, 3: 1 / 216}[int(n)] for n in (STENCIL**2).sum(1)])
_SECOND = np.stack([STENCIL[:, a] * STENCIL[:, b] for a, b in SECOND], axis=1).astype(float)
_HERM = _SECOND.copy()
_HERM[:, :3] -= CS2
_HERM[:, 3:] *= 2
_RECON = WEIGHTS[:, None] * _HERM * 4.5


class TransferError(ValueError):
    pass


def moments(pdf):
    """[rho, j, Pi_neq(6)] of zero-centred pre-collision populations."""
    delta = pdf.sum(-1)
    rho = delta + 1
    if np.any(rho <= 0):
        raise TransferError("Transfer encountered nonpositive density")
    if np.any(pdf + WEIGHTS < 0):
        raise TransferError("Donor has negative physical populations")
    j = pdf @ STENCIL.astype(float)
    if np.any(np.linalg.norm(j, axis=-1) / rho / math.sqrt(CS2) > 0.3):
        raise TransferError("Transfer supports local lattice Mach <= 0.3 only")
    pi = pdf @ _SECOND
    pi[..., :3] -= delta[..., None] * CS2
    for k, (a, b) in enumerate(SECOND):
        pi[..., k] -= j[..., a] * j[..., b] / rho
    return np.concatenate((rho[..., None], j, pi), axis=-1)


def populations(m):
    rho, j = m[..., 0], m[..., 1:4]
    if np.any(rho <= 0):
        raise TransferError("Prediction produced nonpositive density")
    u = j / rho[..., None]
    eq = np.ones((*rho.shape, 27))
    for a in range(3):
        v = u[..., a, None]
        c = STENCIL[:, a]
        eq *= np.where(c == 0, 2 / 3 - v**2, (CS2 + v**2 + c * v) / 2)
    pdf = eq * rho[..., None] + m[..., 4:] @ _RECON.T
    if not np.isfinite(pdf).all() or np.any(pdf < 0):
        raise TransferError("Prediction produced invalid or negative physical populations")
    return pdf - WEIGHTS


def rescale_stress(fields, source, target):
    out = fields.copy()
    pi = out[..., 4:]
    iso = pi[..., :3].mean(axis=-1)
    shear = target.tau_shear * target.dt / (source.tau_shear * source.dt)
    bulk = target.tau_bulk * target.dt / (source.tau_bulk * source.dt)
    pi *= shear
    pi[..., :3] += (bulk - shear) * iso[..., None]
    return out


def prolong_cell_averages(values):
    for axis in range(3):
        idx = np.arange(2, values.shape[axis] - 2)
        centre = np.take(values, idx, axis=axis)
        off = (11 / 64 * (np.take(values, idx + 1, axis=axis) - np.take(values, idx - 1, axis=axis))
               - 3 / 128 * (np.take(values, idx + 2, axis=axis) - np.take(values, idx - 2, axis=axis)))
        values = np.stack((centre - off, centre + off), axis=axis + 1)
        shape = list(centre.shape)
        shape[axis] *= 2
        values = values.reshape(shape)
    return values


def restrict_cell_averages(v):
    nx, ny, nz = v.shape[:3]
    return v.reshape(nx // 2, 2, ny // 2, 2, nz // 2, 2, *v.shape[3:]).mean(axis=(1, 3, 5))


class FaceFlux:
    """Moment flux [rho, j] of pull streaming through the upper face of cell `index` on `axis`."""

    def __init__(self, shape, axis, index, ranges, periodic):
        self.shape, self.axis = tuple(shape), axis
        others = [a for a in range(3) if a != axis]
        self.out_shape = tuple(hi - lo for lo, hi in ranges)
        self.terms = []
        for d, c in enumerate(STENCIL):
            nrm = c[axis]
            if nrm == 0:
                continue
            moving = [a for a in others if c[a] != 0]
            m = len(moving)
            for k in range(m + 1):
                w = math.factorial(k) * math.factorial(m - k) / math.factorial(m + 1)
                for subset in combinations(moving, k):
                    coords = [None] * 3
                    coords[axis] = np.array([index + (1 if nrm < 0 else 0)])
                    for a, (lo, hi) in zip(others, ranges):
                        coords[a] = np.arange(lo, hi) - (c[a] if a in subset else 0)
                    for a in range(3):
                        if periodic:
                            coords[a] = coords[a] % self.shape[a]
                        elif coords[a].min() < 0 or coords[a].max() >= self.shape[a]:
                            raise ValueError("Insufficient donor padding for face flux")
                    self.terms.append((d, np.ix_(*coords), w * nrm * np.r_[1.0, c]))

    def __call__(self, post):
        out = np.zeros((*self.out_shape, 4))
        for d, ix, coef in self.terms:
            out += post[..., d][ix].squeeze(axis=self.axis)[..., None] * coef
        return out


class RefinedOracle:
    pad = 2
    outer_flux_before_boundaries = False
    track_budget = True  # the conserved integrals around every step (last_budget)

    def __init__(self, project, settings, plan: RefinedPlan | None = None):
        self.project, self.settings = project, settings
        self.section, self.flow, self.grid = project.section, project.flow, project.grid
        self.plan = plan or RefinedPlan.create(project, settings)
        p = self.plan
        self.shape = p.shape
        self.box = p.box
        self.coarse_level, self.fine_level = p.coarse_level, p.fine_level
        self.fine_shape = tuple(2 * n for n in self.box.shape)
        self.padded = tuple(n + 2 * self.pad for n in self.fine_shape)
        self.fine_slices = tuple(slice(self.pad, self.pad + n) for n in self.fine_shape)
        U = self.flow.lattice_speed
        self.contours = self.section.normalized(self.flow.alpha_deg)
        # coarse lattice: legacy farfield, z periodic, no walls
        self.coarse = O.Oracle(*self.shape, 1 / self.coarse_level.tau_shear, 1 / self.coarse_level.tau_bulk)
        self.coarse.set_load_frame(0.0, 0.0, 1.0, U)
        # fine padded lattice: fully periodic sync, walls
        self.fine = O.Oracle(*self.padded, 1 / self.fine_level.tau_shear, 1 / self.fine_level.tau_bulk)
        self.fine_origin_padded = (p.origin[0] + self.box.lower[0] * self.coarse_level.dx - self.pad * self.fine_level.dx,
                                   p.origin[1] + self.box.lower[1] * self.coarse_level.dx - self.pad * self.fine_level.dx)
        self.fine.set_geometry(self.contours, self.fine_origin_padded, self.fine_level.dx)
        self.fine.set_load_frame(0.0, 0.0, 1.0, U)
        self.free_fields = np.zeros(10)
        self.free_fields[0], self.free_fields[1] = 1.0, U
        self.free_pdf = populations(self.free_fields)
        self.core = virtual_core(self.box.shape)
        self.fine_core = np.repeat(np.repeat(np.repeat(self.core, 2, 0), 2, 1), 2, 2)
        self.full_periodic_axes = [a for a in range(3) if self.box.lower[a] == 0 and self.box.upper[a] == self.shape[a]]
        # outer face-flux samplers of the coarse domain, on the ghost-padded array
        padded_coarse = tuple(n + 2 for n in self.shape)
        self.outer = []
        for axis in range(3):
            ranges = [(1, self.shape[a] + 1) for a in range(3) if a != axis]
            for side, index in ((-1, 0), (1, self.shape[axis])):
                self.outer.append((side, FaceFlux(padded_coarse, axis, index, ranges, False)))
        self.coarse_owned = np.ones(self.shape, bool)
        self.coarse_owned[self.box.slices] = False
        self.faces, self.support = [], np.zeros(self.shape, bool)
        for axis in range(3):
            if axis in self.full_periodic_axes:
                continue
            others = [a for a in range(3) if a != axis]
            cr = [(self.box.lower[a], self.box.upper[a]) for a in others]
            fr = [(self.pad, self.pad + self.fine_shape[a]) for a in others]
            for side in (-1, 1):
                ci = self.box.lower[axis] - 1 if side < 0 else self.box.upper[axis] - 1
                fi = self.pad - 1 if side < 0 else self.pad + self.fine_shape[axis] - 1
                outside = list(self.box.slices)
                outside[axis] = self.box.lower[axis] - 1 if side < 0 else self.box.upper[axis]
                outside = tuple(outside)
                self.support[outside] = True
                self.faces.append((axis, side, outside, FaceFlux(self.shape, axis, ci, cr, True),
                                   FaceFlux(self.padded, axis, fi, fr, False)))
        self.initialize()

    # -------------------------------------------------------------- state
    def _fine_pdf(self):
        return self.fine.pdf()

    def _restrict_fine(self, fine_pdf):
        safe = fine_pdf.copy()
        safe[self.fine_core] = self.free_pdf
        fields = restrict_cell_averages(moments(safe))
        result = populations(rescale_stress(fields, self.fine_level, self.coarse_level))
        result[self.core] = self.free_pdf
        return result

    def _donor_block(self, coarse_pdf):
        safe = coarse_pdf.copy()
        covered = safe[self.box.slices]
        covered[self.core] = self.free_pdf
        safe[self.box.slices] = covered
        out = safe
        for axis, (lo, hi) in enumerate(zip(self.box.lower, self.box.upper)):
            out = np.take(out, np.arange(lo - 3, hi + 3) % self.shape[axis], axis=axis)
        return out

    def _fine_donors(self, before, after, fraction):
        a, b = moments(before), moments(after)
        fields = prolong_cell_averages((1 - fraction) * a + fraction * b)
        fields = rescale_stress(fields, self.coarse_level, self.fine_level)
        # the owned cells get the free stream here and their own state right after; the
        # reconstruction is elementwise, so only the donor cells are rebuilt
        out = np.broadcast_to(self.free_pdf, (*fields.shape[:3], 27)).copy()
        donor = np.ones(fields.shape[:3], bool)
        donor[self.fine_slices] = False
        out[donor] = populations(fields[donor])
        return out

    def _wrap_z(self, pdf):
        for axis in self.full_periodic_axes:
            n = self.fine_shape[axis]
            for at, src in ((slice(0, self.pad), slice(n, n + self.pad)),
                            (slice(n + self.pad, n + 2 * self.pad), slice(self.pad, 2 * self.pad))):
                t, s = [slice(None)] * 4, [slice(None)] * 4
                t[axis], s[axis] = at, src
                pdf[tuple(t)] = pdf[tuple(s)]

    def _initial_fields(self, shape, dx, origin):
        p, U = self.plan, self.flow.lattice_speed
        rng = np.random.default_rng(self.flow.seed)
        phy, phz = rng.uniform(0, 2 * np.pi, 2)
        ky = 2 * np.pi / (self.shape[1] * self.coarse_level.dx)
        kz = 2 * np.pi / p.span
        pts = np.concatenate([e.xy for e in self.section.elements])
        le = pts[np.argmin(pts[:, 0])]
        ang = -np.deg2rad(self.flow.alpha_deg)
        rot = np.array([[np.cos(ang), -np.sin(ang)], [np.sin(ang), np.cos(ang)]])
        le = rot @ ((le - np.array(self.section.moment_reference)) / self.section.chord)
        x, y, z = np.meshgrid(*[(np.arange(n) + 0.5) * dx + o for n, o in zip(shape, origin)], indexing="ij", sparse=True)
        xx, yy = x + p.origin[0], y + p.origin[1]
        rx, ry = xx - le[0], yy - le[1]
        j = np.zeros((*shape, 3))
        j[..., 0] = U
        amp = self.flow.perturbation * U
        env = np.exp(-((rx - 0.5) / 0.75) ** 2)
        j[..., 1] = amp * np.sin(ky * ry + phy) * np.cos(kz * z + phz) * env
        j[..., 2] = -amp * ky / kz * np.cos(ky * ry + phy) * np.sin(kz * z + phz) * env
        solid = np.zeros(shape[:2], bool)
        for c in self.contours:
            solid |= c.contains(xx[:, :, 0], yy[:, :, 0])
        j[np.broadcast_to(solid[:, :, None], shape)] = 0.0
        fields = np.zeros((*shape, 10))
        fields[..., 0] = 1.0
        fields[..., 1:4] = j
        return fields

    def initialize(self):
        dc, df = self.coarse_level.dx, self.fine_level.dx
        coarse = populations(self._initial_fields(self.shape, dc, (0, 0, 0)))
        origin = tuple(lo * dc - self.pad * df for lo in self.box.lower)
        fine = populations(self._initial_fields(self.padded, df, origin))
        coarse[self.box.slices] = self._restrict_fine(fine[self.fine_slices])
        self.coarse.fill_all(self.free_pdf)
        self.fine.fill_all(self.free_pdf)
        self.coarse.set_pdf(coarse)
        self.fine.set_pdf(fine)
        self.coarse.set_farfield("legacy")
        self.fine.set_farfield("periodic")
        self.steps = 0
        self.fine_fluid = np.broadcast_to(~self.fine.solid()[self.fine_slices[:2]][:, :, None], self.fine_shape)
        self.cumulative_wall_mass_exchange = 0.0
        self.cumulative_wall_impulse = np.zeros(3)
        self.cumulative_external_in = np.zeros(4)
        self.last_budget = {}
        self.initial_integrals = self.conserved_integrals()

    def conserved_integrals(self):
        """Mass and momentum (common units) of the owned cells of both levels."""
        mass, impulse = 0.0, np.zeros(3)
        fine = self.fine.pdf()[self.fine_slices]
        for pdf, owned, level in ((self.coarse.pdf(), self.coarse_owned, self.coarse_level),
                                  (fine, self.fine_fluid, self.fine_level)):
            rho = pdf.sum(axis=-1) + 1
            j = pdf @ STENCIL.astype(float)
            mass += rho[owned].sum() * level.dx**3
            impulse += j[owned].sum(axis=0) * level.dx**3 * level.lattice_velocity_unit
        return {"mass": float(mass), "momentum": impulse.tolist()}

    # -------------------------------------------------------------- step
    def step(self):
        start = self.conserved_integrals() if self.track_budget else None
        L = self.fine.L
        dc, df = self.coarse_level.dx, self.fine_level.dx
        old = self.coarse.pdf()
        fine_state = self.fine.pdf()[self.fine_slices].copy()
        L.ako_collide(self.coarse.h)
        post = self.coarse.pdf()
        coarse_flux = [f[3](post) * dc**3 for f in self.faces]
        if self.outer_flux_before_boundaries:  # a planted mistake for the tests only
            padded = self.coarse.padded_pdf()
        L.ako_boundaries(self.coarse.h)
        if not self.outer_flux_before_boundaries:
            padded = self.coarse.padded_pdf()
        external = np.zeros(4)
        for side, sampler in self.outer:
            external -= side * sampler(padded).sum(axis=(0, 1)) * dc**3
        L.ako_stream(self.coarse.h)
        prediction = self.coarse.pdf()
        before, after = self._donor_block(old), self._donor_block(prediction)
        fine_flux = [np.zeros_like(f) for f in coarse_flux]
        walls = []
        lo = np.array([self.pad] * 3, dtype=np.int32)
        hi = np.array([self.pad + n for n in self.fine_shape], dtype=np.int32)
        for sub in range(2):
            padded = self._fine_donors(before, after, sub / 2)
            padded[self.fine_slices] = fine_state
            self._wrap_z(padded)
            self.fine.set_pdf(padded)
            L.ako_collide(self.fine.h)
            post = self.fine.pdf()
            for i, f in enumerate(self.faces):
                flux = f[4](post)
                a, b = flux.shape[:2]
                fine_flux[i] += flux.reshape(a // 2, 2, b // 2, 2, 4).sum(axis=(1, 3)) * df**3
            L.ako_boundaries(self.fine.h)
            out = np.zeros(6)
            L.ako_measure(self.fine.h, O._ip(lo), O._ip(hi), df, self.fine_origin_padded[0],
                          self.fine_origin_padded[1], O._dp(out))
            walls.append(out)
            L.ako_stream(self.fine.h)
            fine_state = self.fine.pdf()[self.fine_slices].copy()
        correction = np.zeros((*self.shape, 4))
        for (axis, side, outside, _, _), fc, ff in zip(self.faces, coarse_flux, fine_flux):
            correction[outside] += side * (ff - fc) / dc**3
        sup = self.support
        old_m = moments(prediction[sup])
        old_m[..., 4:] = 0
        new_m = old_m.copy()
        new_m[..., :4] += correction[sup]
        prediction[sup] += populations(new_m) - populations(old_m)
        moments(prediction[sup])
        prediction[self.box.slices] = self._restrict_fine(fine_state)
        self.coarse.set_pdf(prediction)
        self.steps += 1
        vol, vel = df**3, self.fine_level.lattice_velocity_unit
        impulse = (walls[0][:3] + walls[1][:3]) * vol * vel
        moment = (walls[0][3] + walls[1][3]) * vol * vel
        wall_mass = (walls[0][4] + walls[1][4]) * vol
        q_area = 0.5 * self.plan.span
        force = impulse / self.coarse_level.dt
        mz = moment / self.coarse_level.dt
        external[1:] *= vel
        self.cumulative_wall_mass_exchange += wall_mass
        self.cumulative_wall_impulse += impulse
        self.cumulative_external_in += external
        record = {"step": self.steps, "t_star": (self.steps - 0.5) * self.coarse_level.dt,
                  "cl": float(force[1] / q_area), "cd": float(force[0] / q_area), "cm": float(-mz / q_area),
                  "wall_mass_exchange": float(wall_mass)}
        if not self.track_budget:
            return record
        # budget of this coarse step
        end = self.conserved_integrals()
        mass_residual = end["mass"] - start["mass"] - external[0] - wall_mass
        momentum_residual = np.array(end["momentum"]) - start["momentum"] - external[1:] + impulse
        volume = (np.count_nonzero(self.coarse_owned) * dc**3 + np.count_nonzero(self.fine_fluid) * df**3)
        self.last_budget = {"mass_residual_per_volume": float(mass_residual / volume),
                            "momentum_residual_per_volume": (momentum_residual / volume).tolist(),
                            "wall_mass_exchange": float(wall_mass), "external_mass_in": float(external[0]),
                            "wall_impulse": impulse.tolist(), "external_impulse_in": external[1:].tolist()}
        return record

    def diagnostics(self):
        """Guards over the leaf cells and the budget (refined_solver.py:347-361)."""
        mach, derr = 0.0, 0.0
        for pdf, owned in ((self.coarse.pdf(), self.coarse_owned), (self.fine.pdf()[self.fine_slices], self.fine_fluid)):
            rho = pdf.sum(axis=-1)[owned] + 1
            u = (pdf @ STENCIL.astype(float))[owned] / rho[:, None]
            mach = max(mach, float(np.linalg.norm(u, axis=-1).max() * math.sqrt(3)))
            derr = max(derr, float(np.max(abs(rho - 1))))
        return {"local_mach_max": mach, "density_error_max": derr,
                "coarse_boundary_calls": self.steps, "fine_boundary_calls": 2 * self.steps,
                "cumulative_wall_mass_exchange": self.cumulative_wall_mass_exchange,
                "cumulative_wall_mass_exchange_over_initial_mass":
                    self.cumulative_wall_mass_exchange / self.initial_integrals["mass"],
                "last_budget": self.last_budget}
