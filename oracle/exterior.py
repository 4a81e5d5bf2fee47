"""FP64 restatement of the `exterior_regularized_fd_v1` farfield (test infrastructure only).

TEST INFRASTRUCTURE ONLY.  Only tests, the smoke check and the CPU-baseline leg of the
benchmark use this module; the product path never imports it.

What the reference does (src/aerokinetic/exterior.py:18-200, solver.py:36-41, 81-90,
157-202, 252-279):

* the physical lattice is padded by one cell in x and y ("storage"); the pad is the
  exterior ring, whose populations are rewritten every step; the storage edge's ghost
  layer is a velocity bounce-back reservoir (exterior.py:29-34);
* step order collide -> boundaries -> wall loads -> stream (collide_stream, pre-collision
  state between steps);
* before the collision, an SSP-RK3 predictor advances the ring's macroscopic state
  (ln rho, u) one step with characteristic rates built from one-sided / central finite
  differences (exterior.py:73-133); inflow and glancing cells are re-imposed after every
  stage (exterior.py:86-110);
* after streaming the ring populations are replaced by a regularized reconstruction
  of the predicted state: product-form equilibrium plus a first-order non-equilibrium
  stress from the velocity gradients (exterior.py:135-151); the change is the
  "exterior source" of mass and momentum (exterior.py:187-200).

Data flow of this restatement (not the reference's full-field arrays): the macroscopic
state is kept only on the *band*, the storage cells within two cells of the storage
edge in x or y, ring cells first; fixed neighbour tables of the ring cells into the band
drive the finite differences.  The lattice itself (collision, streaming, walls,
reservoir) is the C oracle (`ak_oracle.c`) on the storage shape.

Pins: the reference's recorded exterior histories (`tests/golden/exterior_*`); the
reconstruction reproduces rho and rho u to rounding (closed-form check in the tests).
"""
from __future__ import annotations

import math

import numpy as np

from .oracle import Oracle, seed_parameters

CS = math.sqrt(1.0 / 3.0)

# D3Q27 in the oracle's order i = 9 (cx+1) + 3 (cy+1) + (cz+1)
_C = np.array([[i // 9 - 1, (i // 3) % 3 - 1, i % 3 - 1] for i in range(27)], dtype=np.float64)
_W = np.array([{0: 8 / 27, 1: 2 / 27, 2: 1 / 54, 3: 1 / 216}[int(np.abs(c).sum())] for c in _C])


class ExteriorOracle:
    """Uniform airfoil case with the exterior farfield, advanced in the reference's order."""

    def __init__(self, section, flow, grid, plan):
        nx, ny, nz = plan.shape
        if min(nx, ny, nz) < 4:
            raise ValueError("exterior farfield needs a physical lattice of at least 4 cells per axis")
        X, Y, Z = nx + 2, ny + 2, nz
        self.storage_shape = (X, Y, Z)
        self.storage_origin = (plan.origin[0] - plan.dx, plan.origin[1] - plan.dx)
        self.plan, self.flow, self.grid = plan, flow, grid
        self.U = flow.lattice_speed
        self.omega = plan.relaxation_rate
        o = Oracle(X, Y, Z, plan.relaxation_rate, 1.0)
        self.contours = section.normalized(flow.alpha_deg)
        o.set_geometry(self.contours, self.storage_origin, plan.dx)
        o.set_load_frame(self.storage_origin[0] / plan.dx, self.storage_origin[1] / plan.dx,
                         grid.cells_per_chord, self.U)
        # seeded field on the storage cells (the physical wave numbers), then the ring at
        # the free stream (solver.py:157-202)
        o.init_seeded(seed_parameters(section, flow, plan), self.storage_origin, plan.dx)
        self.o = o
        self._tables()
        feq = np.array([self._phi(c[0], self.U) * self._phi(c[1], 0.0) * self._phi(c[2], 0.0) for c in _C]) - _W
        o.set_cells(self.ring_xyz, np.broadcast_to(feq, (len(self.ring_xyz), 27)))
        o.set_farfield("reservoir")
        self.rows = None
        self.steps = 0
        self.source_total = np.zeros(4)
        self.wall_mass = 0.0
        self.wall_total = 0.0
        self.last_report = None

    @staticmethod
    def _phi(c, q):
        return 2.0 / 3.0 - q * q + (1.5 * q * q - 0.5) * c * c + c * q / 2.0

    # ------------------------------------------------------------------ tables
    def _tables(self):
        X, Y, Z = self.storage_shape
        xs, ys = np.meshgrid(np.arange(X), np.arange(Y), indexing="ij")
        depth = np.minimum(np.minimum(xs, X - 1 - xs), np.minimum(ys, Y - 1 - ys))
        ring = np.argwhere(depth == 0)
        band = np.concatenate([ring, np.argwhere((depth >= 1) & (depth <= 2))])
        index = -np.ones((X, Y), dtype=np.int64)
        index[band[:, 0], band[:, 1]] = np.arange(len(band))
        R = len(ring)
        self.R, self.band = R, band
        nv = np.zeros((R, 3))
        nv[:, 0] = np.where(ring[:, 0] == 0, -1, np.where(ring[:, 0] == X - 1, 1, 0))
        nv[:, 1] = np.where(ring[:, 1] == 0, -1, np.where(ring[:, 1] == Y - 1, 1, 0))
        self.ell = np.linalg.norm(nv, axis=-1)
        self.n = nv / self.ell[:, None]
        base = np.array([self.U, 0.0, 0.0])
        self.uninf = self.n @ base
        self.utinf = base - self.uninf[:, None] * self.n
        # three-point differences along x and y: one-sided on the storage edge, central along it
        self.fd = []
        for a, size in ((0, X), (1, Y)):
            edge = np.where(ring[:, a] == 0, -1, np.where(ring[:, a] == size - 1, 1, 0))
            one, two = ring.copy(), ring.copy()
            one[:, a] += np.where(edge == -1, 1, -1)
            two[:, a] += np.where(edge == 0, 1, -2 * edge)
            c0 = 1.5 * edge
            c1 = np.where(edge == 0, -0.5, -2.0 * edge)
            c2 = np.where(edge == 0, 0.5, 0.5 * edge)
            i1, i2 = index[one[:, 0], one[:, 1]], index[two[:, 0], two[:, 1]]
            assert (i1 >= 0).all() and (i2 >= 0).all()
            self.fd.append((i1, i2, c0, c1, c2))
        self.inner = index[(ring[:, 0] - nv[:, 0]).astype(int), (ring[:, 1] - nv[:, 1]).astype(int)]
        assert (self.inner >= 0).all()
        zz = np.arange(Z)
        self.band_xyz = np.column_stack([np.repeat(band, Z, axis=0), np.tile(zz, len(band))]).astype(np.int32)
        self.ring_xyz = self.band_xyz[: R * Z]
        # glancing cells: the two y rows (normal (0, -+1, 0)), index of their row and x
        self.glancing = np.flatnonzero(self.uninf == 0)
        self.g_side = (ring[self.glancing, 1] == Y - 1).astype(int)
        self.g_x = ring[self.glancing, 0]
        self.g_row_index = [index[np.arange(X), 0], index[np.arange(X), Y - 1]]

    # ------------------------------------------------------------------ macroscopic band state
    def band_state(self):
        """(ln rho, u) of every band cell: (band, z, 4)."""
        g = self.o.cells(self.band_xyz)
        rho = 1.0 + g.sum(axis=1)
        u = (g @ _C) / rho[:, None]
        return np.concatenate([np.log(rho)[:, None], u], axis=1).reshape(len(self.band), -1, 4)

    def _grad(self, F):
        """d/dx_a of the band field F at the ring cells: (ring, z, axis, 4)."""
        R = self.R
        c = F[:R]
        out = np.empty((R, F.shape[1], 3, 4))
        for a in (0, 1):
            i1, i2, c0, c1, c2 = self.fd[a]
            out[:, :, a] = c0[:, None, None] * c + c1[:, None, None] * F[i1] + c2[:, None, None] * F[i2]
        out[:, :, 2] = (np.roll(c, -1, axis=1) - np.roll(c, 1, axis=1)) / 2
        return out

    # ------------------------------------------------------------------ predictor
    def _rhs(self, y, cur):
        """Characteristic rates of the ring state y (exterior.py:112-125)."""
        R = self.R
        stage = cur.copy()
        stage[:R] = y
        g = self._grad(stage)
        G, dq = g[..., 1:], g[..., 0]  # G[r, z, a, b] = d_a u_b
        n = self.n[:, None, :]
        un = np.sum(y[..., 1:] * n, axis=-1)
        rp = un + CS * y[..., 0]
        ins = stage[self.inner]
        uin = ins[..., 1:]
        unin = np.sum(uin * n, axis=-1)
        rpin = unin + CS * ins[..., 0]
        div = np.trace(G, axis1=2, axis2=3) - np.sum(n[..., :, None] * G * n[..., None, :], axis=(2, 3))
        gradun = np.sum(G * n[:, :, None, :], axis=-1)
        ut, uinf, ell = self.utinf[:, None, :], self.uninf[:, None], self.ell[:, None]
        rminus = uinf * div - np.sum(ut * (gradun - CS * dq), axis=-1)
        rplus = -(uinf + CS) * (rp - rpin) / ell - np.sum(ut * (gradun + CS * dq), axis=-1) - CS * div
        tangent = y[..., 1:] - un[..., None] * n
        tangentin = uin - unin[..., None] * n
        adv = np.sum(ut[..., :, None] * G, axis=-2)
        adv = adv - np.sum(adv * n, axis=-1)[..., None] * n
        pressure = dq - np.sum(dq * n, axis=-1)[..., None] * n
        dt = -uinf[..., None] * (tangent - tangentin) / ell[..., None] - adv - CS ** 2 * pressure
        out = np.empty_like(y)
        out[..., 0] = (rplus - rminus) / (2 * CS)
        out[..., 1:] = n * ((rplus + rminus) / 2)[..., None] + dt
        return out

    def _traced(self, t):
        """Incoming invariant of the glancing cells: the initial R- of their row, traced back
        along the free stream by U t (exterior.py:86-99)."""
        X = self.storage_shape[0]
        out = np.zeros((self.R, self.storage_shape[2]))
        target = self.g_x - self.U * t
        lower = np.floor(target).astype(int)
        f = target - lower
        acc = np.zeros((len(self.glancing), self.storage_shape[2]))
        for bit, w in ((0, 1.0 - f), (1, f)):
            q = lower + bit
            ok = (q >= 0) & (q < X)
            v = self.rows[self.g_side, np.clip(q, 0, X - 1)]
            acc = acc + w[:, None] * np.where(ok[:, None], v, 0.0)
        out[self.glancing] = acc
        return out

    def _enforce(self, y, t):
        """Impose the incoming invariant on inflow and glancing cells (exterior.py:101-110)."""
        y = y.copy()
        n = self.n[:, None, :]
        un = np.sum(y[..., 1:] * n, axis=-1)
        rp = un + CS * y[..., 0]
        inflow, glancing = self.uninf < 0, self.uninf == 0
        sel = inflow | glancing
        rm = np.where(inflow[:, None], self.uninf[:, None], self._traced(t))
        newun = (rp + rm) / 2
        tangent = y[..., 1:] - un[..., None] * n
        tangent[inflow] = self.utinf[inflow][:, None, :]
        y[sel, :, 0] = ((rp - rm) / (2 * CS))[sel]
        y[sel, :, 1:] = (newun[..., None] * n + tangent)[sel]
        return y

    def predict(self, cur, t):
        """SSP-RK3 step of the ring state from the band state `cur` (exterior.py:127-133)."""
        y0 = cur[: self.R]
        y1 = self._enforce(y0 + self._rhs(y0, cur), t + 1)
        y2 = self._enforce(0.75 * y0 + 0.25 * (y1 + self._rhs(y1, cur)), t + 0.5)
        y3 = self._enforce((y0 + 2 * (y2 + self._rhs(y2, cur))) / 3, t + 1)
        if not np.isfinite(y3).all():
            raise RuntimeError("Non-finite exterior macro predictor")
        return y3

    # ------------------------------------------------------------------ reconstruction
    def regularize(self, V):
        """Ring populations (ring, z, 27) from the band state V whose ring holds the
        prediction, and the largest mass / momentum mismatch (exterior.py:135-151)."""
        R = self.R
        y = V[:R]
        rho = np.exp(y[..., 0])
        u = y[..., 1:]
        G = self._grad(V)[..., 1:]
        div = np.trace(G, axis1=2, axis2=3)
        w = self.omega
        S = -rho[..., None, None] / 3 * ((G + np.swapaxes(G, 2, 3)) / w
                                         + np.eye(3) * (2 / 3 * div * (1 - 1 / w))[..., None, None])
        c = _C[None, None]
        v = u[:, :, None, :]
        phi = 2 / 3 - v * v + (1.5 * v * v - 0.5) * c * c + c * v / 2
        first = -2 * v + 3 * v * c * c + c / 2
        second = 1.5 * c * c - 1
        pdf = rho[..., None] * (phi[..., 0] * phi[..., 1] * phi[..., 2]) - _W
        for a, (j, k) in ((0, (1, 2)), (1, (0, 2)), (2, (0, 1))):
            pdf = pdf + S[..., a, a, None] * second[..., a] * (phi[..., j] * phi[..., k])
        for a, b in ((0, 1), (0, 2), (1, 2)):
            pdf = pdf + S[..., a, b, None] * first[..., a] * first[..., b] * phi[..., 3 - a - b]
        mass = float(np.max(np.abs(1 + pdf.sum(axis=-1) - rho)))
        mom = float(np.max(np.abs(pdf @ _C - rho[..., None] * u)))
        return pdf, max(mass, mom)

    # ------------------------------------------------------------------ step
    def step(self, sample=False):
        """One time step in the reference's order; with sample=True the record of the
        history row (solver.py:252-279, exterior.py:173-200)."""
        cur = self.band_state()
        if self.rows is None:
            # R- = u.n - cs ln(rho) of the two y rows at the first step (n = (0, -+1, 0))
            self.rows = np.stack([-cur[self.g_row_index[0], :, 2] - CS * cur[self.g_row_index[0], :, 0],
                                  cur[self.g_row_index[1], :, 2] - CS * cur[self.g_row_index[1], :, 0]])
        pred = self.predict(cur, self.steps)
        self.o.collide()
        self.o.boundaries()
        loads, mass = self.o.wall_loads()
        self.o.stream()
        nxt = self.band_state()
        nxt[: self.R] = pred
        rebuilt, err = self.regularize(nxt)
        new = rebuilt.reshape(-1, 27)
        old = self.o.cells(self.ring_xyz)
        self.o.set_cells(self.ring_xyz, new)
        self.steps += 1
        dm = new - old
        source = np.r_[dm.sum(), dm.sum(axis=0) @ _C]
        self.source_total += source
        self.wall_mass = mass
        self.wall_total += mass
        self.last_report = {"step": self.steps, "reconstruction_moment_error": err,
                            "exterior_source_mass_momentum": source.tolist()}
        if not sample:
            return None
        tot = loads.sum(axis=0)
        qa = 0.5 * self.U ** 2 * self.grid.cells_per_chord * self.storage_shape[2]
        return {"step": self.steps, "t_star": (self.steps - 0.5) * self.plan.dt_convective,
                "cl": float(tot[1] / qa), "cd": float(tot[0] / qa),
                "cm": float(-tot[3] / (qa * self.grid.cells_per_chord)),
                "element_lattice_loads": loads,
                "wall_mass_exchange_lattice": mass, "wall_mass_exchange_accumulated": self.wall_total,
                "exterior_source_mass_lattice": float(source[0]),
                "exterior_source_mass_accumulated": float(self.source_total[0])}

    def diagnostics(self):
        """Local Mach and |rho - 1| maxima over the storage fluid cells (ring included)."""
        return self.o.diagnostics()

    def populations(self):
        """Pre-collision storage populations (X, Y, Z, 27)."""
        return self.o.pdf()

    def collided(self):
        """The storage populations after the next collision (the GPU's stored form)."""
        return self.o.collided()
