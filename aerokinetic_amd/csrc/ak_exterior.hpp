// ak_exterior.hpp -- gfx950 kernels of the exterior farfield (exterior_regularized_fd_v1).
//
// The lattice is the physical box padded by one cell in x and y; the pad is the exterior
// ring (mask value 3), whose populations are rebuilt every step from a predicted
// macroscopic state (src/aerokinetic/exterior.py:18-200).  The storage edge's ghost layer
// is a velocity bounce-back reservoir (the legacy bounce-back kernel on every side).
//
// Per step, after the boundary kernels and the AA sweep (which skips ring cells):
//   k_ext_band   (ln rho, u) of the band cells one and two cells inside the ring, from the
//                new post-collision state (collision conserves rho and j);
//   k_ext_ring   per ring cell: gather its streamed populations exactly where a sweep of
//                this parity would, replace them by the regularized reconstruction of the
//                predicted state (exterior.py:135-151), record the source (new - old) as
//                block partials, store the ring's macroscopic state for the next
//                prediction, collide and scatter like the sweep;
//   k_ext_stage  the three SSP-RK3 stages of the ring predictor (exterior.py:101-133);
//   k_ext_end    fixed-order sums of the wall mass exchange and the ring source into the
//                accumulators, then the end of the step (on sampled steps the history row).
// The state between steps is post-collision like the other farfields; the reference's
// pre-collision state has the same rho and u.
#pragma once

#include "../../include/aerokinetic.h"
#include "ak_kernels.hpp"

namespace ak {

struct ExtAcc {
  double wall_mass, wall_total;  // this step's wall mass exchange sum(in - out), accumulated
  double src[4], src_total[4];   // this step's exterior source (mass, momentum), accumulated
  double moment_err;             // largest |mass| / |momentum| mismatch of this step's rebuild
  long long steps;               // exterior steps taken
};

struct ExtDev {
  const int32_t *band_xy;  // x | y << 16, ring cells first
  const int32_t *map;      // x * Y + y -> band index, -1 outside the band
  int R2, B2, X, Y, Z;
  double U, omega;
  double *M;               // band state (ln rho, ux, uy, uz) at (b * Z + z) * 4
  double *Y1, *Y2, *PRED;  // ring stage values at (r * Z + z) * 4
  const double *rows;      // initial R- of the two y rows at (side * Z + z) * X + x
  double *part;            // ring kernel block partials, 5 per block
  ExtAcc *acc;
  int *err;                // set to 1 by a non-finite prediction
};

// History columns written by the exterior farfield: wall mass accumulated, source mass,
// source mass accumulated (after the per-element loads of up to MAX_ELEM elements).
constexpr int EXT_COL = AK_HIST_EXT_WALL_MASS_TOTAL;
static_assert(AK_HIST_EXT_WALL_MASS_TOTAL == AK_HIST_FIXED + 4 * MAX_ELEM, "history layout");
constexpr double EXT_CS = 0.57735026918962576451;      // sqrt(1/3)

// State of band cell (x, y, z): ring cells from `ring` (indexed like the band), others from M.
__device__ __forceinline__ void ext_at(const ExtDev &E, const double *ring, int x, int y, int z, double v[4]) {
  const int b = E.map[x * E.Y + y];
  const double *p = (b < E.R2 ? ring : E.M) + ((long long)b * E.Z + z) * 4;
  v[0] = p[0]; v[1] = p[1]; v[2] = p[2]; v[3] = p[3];
}

// Outward normal of a ring cell: integer normal, its length and the free-stream split.
struct RingFrame {
  int ex, ey;
  double n[3], ell, uninf, ut[3];
};
__device__ __forceinline__ RingFrame ring_frame(const ExtDev &E, int x, int y) {
  RingFrame f;
  f.ex = x == 0 ? -1 : (x == E.X - 1 ? 1 : 0);
  f.ey = y == 0 ? -1 : (y == E.Y - 1 ? 1 : 0);
  f.ell = sqrt((double)(f.ex * f.ex + f.ey * f.ey));
  f.n[0] = f.ex / f.ell; f.n[1] = f.ey / f.ell; f.n[2] = 0.0;
  f.uninf = f.n[0] * E.U;
  f.ut[0] = E.U - f.uninf * f.n[0]; f.ut[1] = -f.uninf * f.n[1]; f.ut[2] = -f.uninf * f.n[2];
  return f;
}

// Gradient d_a of the 4 state components at ring cell (x, y, z): three-point one-sided
// differences across the storage edge, central along it, periodic in z (exterior.py:73-84).
__device__ __forceinline__ void ext_grad(const ExtDev &E, const double *ring, int x, int y, int z,
                                         const double c[4], double g[3][4]) {
  double a[4], b[4];
#pragma unroll
  for (int ax = 0; ax < 2; ++ax) {
    const int p = ax == 0 ? x : y, size = ax == 0 ? E.X : E.Y;
    const int e = p == 0 ? -1 : (p == size - 1 ? 1 : 0);
    const int o1 = e == -1 ? 1 : -1, o2 = e == 0 ? 1 : -2 * e;
    const double c0 = 1.5 * e, c1 = e == 0 ? -0.5 : -2.0 * e, c2 = e == 0 ? 0.5 : 0.5 * e;
    if (ax == 0) { ext_at(E, ring, x + o1, y, z, a); ext_at(E, ring, x + o2, y, z, b); }
    else { ext_at(E, ring, x, y + o1, z, a); ext_at(E, ring, x, y + o2, z, b); }
#pragma unroll
    for (int k = 0; k < 4; ++k) g[ax][k] = (c0 * c[k] + c1 * a[k]) + c2 * b[k];
  }
  ext_at(E, ring, x, y, z + 1 == E.Z ? 0 : z + 1, a);
  ext_at(E, ring, x, y, z == 0 ? E.Z - 1 : z - 1, b);
#pragma unroll
  for (int k = 0; k < 4; ++k) g[2][k] = (a[k] - b[k]) / 2;
}

// ------------------------------------------------------------------ band state
// (ln rho, u) of band cells [b0, b1) x z from the stored populations in form PE.
template <int S, bool PE>
__global__ __launch_bounds__(128) void k_ext_band(const typename St<S>::T *__restrict__ A, Desc L, ExtDev E, int b0,
                                                  int b1, float inv_scale) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)(b1 - b0) * E.Z) return;
  const int b = b0 + (int)(t / E.Z), z = (int)(t % E.Z);
  const int xy = E.band_xy[b], x = xy & 0xffff, y = xy >> 16;
  double r = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
#pragma unroll
  for (int j = 0; j < NQ; ++j) {
    const double v = (double)ld<S>(A + loc_out<PE>(L, x, y, z, j), inv_scale);
    r += v;
    if (cxq(j)) jx += cxq(j) * v;
    if (cyq(j)) jy += cyq(j) * v;
    if (czq(j)) jz += czq(j) * v;
  }
  const double rho = 1.0 + r;
  double *m = E.M + ((long long)b * E.Z + z) * 4;
  m[0] = log(rho); m[1] = jx / rho; m[2] = jy / rho; m[3] = jz / rho;
}

// ------------------------------------------------------------------ predictor
// Incoming invariant of a glancing cell: the initial R- of its row traced back along the
// free stream by U t, linear in x, zero outside the storage (exterior.py:86-99).
__device__ __forceinline__ double ext_traced(const ExtDev &E, int x, int y, int z, double t) {
  const int side = y == 0 ? 0 : 1;
  const double target = (double)x - E.U * t;
  const double lo = floor(target);
  const double f = target - lo;
  const int q0 = (int)lo, q1 = q0 + 1;
  const double *row = E.rows + ((long long)side * E.Z + z) * E.X;
  const double v0 = (q0 >= 0 && q0 < E.X) ? row[q0] : 0.0;
  const double v1 = (q1 >= 0 && q1 < E.X) ? row[q1] : 0.0;
  return (1.0 - f) * v0 + f * v1;
}

// Impose the incoming invariant on inflow and glancing cells (exterior.py:101-110).
__device__ __forceinline__ void ext_enforce(const ExtDev &E, const RingFrame &F, int x, int y, int z, double t,
                                            double v[4]) {
  const bool inflow = F.uninf < 0.0, glancing = F.uninf == 0.0;
  if (!(inflow || glancing)) return;
  const double un = v[1] * F.n[0] + v[2] * F.n[1] + v[3] * F.n[2];
  const double rp = un + EXT_CS * v[0];
  const double rm = inflow ? F.uninf : ext_traced(E, x, y, z, t);
  const double nu = (rp + rm) / 2;
  double tg[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) tg[a] = inflow ? F.ut[a] : v[1 + a] - un * F.n[a];
  v[0] = (rp - rm) / (2 * EXT_CS);
#pragma unroll
  for (int a = 0; a < 3; ++a) v[1 + a] = nu * F.n[a] + tg[a];
}

// Characteristic rates of the ring state y at (x, y, z), the ring cells of the stage taken
// from `ring` and the rest of the band from M (exterior.py:112-125).
__device__ __forceinline__ void ext_rhs(const ExtDev &E, const RingFrame &F, const double *ring, int x, int y, int z,
                                        const double yv[4], double out[4]) {
  double g[3][4];
  ext_grad(E, ring, x, y, z, yv, g);
  double ins[4];
  ext_at(E, E.M, x - F.ex, y - F.ey, z, ins);  // inner neighbour: never a ring cell
  const double *n = F.n;
  const double un = yv[1] * n[0] + yv[2] * n[1] + yv[3] * n[2];
  const double rp = un + EXT_CS * yv[0];
  const double unin = ins[1] * n[0] + ins[2] * n[1] + ins[3] * n[2];
  const double rpin = unin + EXT_CS * ins[0];
  // G[a][b] = d_a u_b = g[a][1 + b], dq[a] = d_a ln rho = g[a][0]
  double nGn = 0.0, tr = 0.0, gradun[3], adv[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    tr += g[a][1 + a];
    gradun[a] = g[a][1] * n[0] + g[a][2] * n[1] + g[a][3] * n[2];
    nGn += n[a] * gradun[a];
  }
  const double div = tr - nGn;
  double sm = 0.0, sp = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    sm += F.ut[a] * (gradun[a] - EXT_CS * g[a][0]);
    sp += F.ut[a] * (gradun[a] + EXT_CS * g[a][0]);
  }
  const double rminus = F.uninf * div - sm;
  const double rplus = -(F.uninf + EXT_CS) * (rp - rpin) / F.ell - sp - EXT_CS * div;
#pragma unroll
  for (int b = 0; b < 3; ++b) adv[b] = F.ut[0] * g[0][1 + b] + F.ut[1] * g[1][1 + b] + F.ut[2] * g[2][1 + b];
  const double an = adv[0] * n[0] + adv[1] * n[1] + adv[2] * n[2];
  const double dn = g[0][0] * n[0] + g[1][0] * n[1] + g[2][0] * n[2];
  out[0] = (rplus - rminus) / (2 * EXT_CS);
  const double half = (rplus + rminus) / 2;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double tangent = yv[1 + a] - un * n[a], tangentin = ins[1 + a] - unin * n[a];
    const double dt = -F.uninf * (tangent - tangentin) / F.ell - (adv[a] - an * n[a]) -
                      EXT_CS * EXT_CS * (g[a][0] - dn * n[a]);
    out[1 + a] = n[a] * half + dt;
  }
}

// Stage STAGE (1..3) of the SSP-RK3 predictor at time t = the number of completed steps
// (from the step cursor, or t0 without one).  y0 = M (ring part).
template <int STAGE>
__global__ __launch_bounds__(128) void k_ext_stage(ExtDev E, const StepCtl *__restrict__ ctl, double t0) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)E.R2 * E.Z) return;
  const int r = (int)(t / E.Z), z = (int)(t % E.Z);
  const int xy = E.band_xy[r], x = xy & 0xffff, y = xy >> 16;
  const double time = ctl ? (double)ctl->step : t0;
  const double *src = STAGE == 1 ? E.M : (STAGE == 2 ? E.Y1 : E.Y2);
  double *dst = STAGE == 1 ? E.Y1 : (STAGE == 2 ? E.Y2 : E.PRED);
  const RingFrame F = ring_frame(E, x, y);
  double y0[4], yv[4], k[4], v[4];
  const long long o = t * 4;
#pragma unroll
  for (int c = 0; c < 4; ++c) { y0[c] = E.M[o + c]; yv[c] = src[o + c]; }
  ext_rhs(E, F, src, x, y, z, yv, k);
  double tt;
  if constexpr (STAGE == 1) {
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = y0[c] + k[c];
    tt = time + 1.0;
  } else if constexpr (STAGE == 2) {
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = 0.75 * y0[c] + 0.25 * (yv[c] + k[c]);
    tt = time + 0.5;
  } else {
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = (y0[c] + 2 * (yv[c] + k[c])) / 3;
    tt = time + 1.0;
  }
  ext_enforce(E, F, x, y, z, tt, v);
#pragma unroll
  for (int c = 0; c < 4; ++c) dst[o + c] = v[c];
  if constexpr (STAGE == 3) {
    if (!(isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]) && isfinite(v[3]))) atomicMax(E.err, 1);
  }
}

// ------------------------------------------------------------------ ring rebuild
// Product-form factors of one axis: phi(c, v) = 2/3 - v^2 + (3/2 v^2 - 1/2) c^2 + c v / 2,
// its v-derivative and half second derivative (exterior.py:140-142), c = -1, 0, 1.
__device__ __forceinline__ void ext_factors(double v, double phi[3], double first[3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double c = (double)(k - 1);
    phi[k] = 2.0 / 3.0 - v * v + (1.5 * v * v - 0.5) * c * c + c * v / 2.0;
    first[k] = -2.0 * v + 3.0 * v * c * c + c / 2.0;
  }
}

// Regularized populations (shifted by the weights) of the state v = (ln rho, u) with
// velocity gradient G[a][b] = d_a u_b: rho prod(phi) - w plus the non-equilibrium stress
// S = -rho/3 [(G + G^T)/omega + 2/3 tr(G) (1 - 1/omega) I] through the second-order
// Taylor factors (exterior.py:135-148).
__device__ __forceinline__ void ext_regularized(const double v[4], const double G[3][3], double omega,
                                                double pdf[NQ]) {
  const double rho = exp(v[0]);
  const double div = G[0][0] + G[1][1] + G[2][2];
  const double iso = 2.0 / 3.0 * div * (1.0 - 1.0 / omega);
  double S[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) S[a][b] = -rho / 3 * ((G[a][b] + G[b][a]) / omega + (a == b ? iso : 0.0));
  double ph[3][3], fi[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) ext_factors(v[1 + a], ph[a], fi[a]);
  const double se[3] = {0.5, -1.0, 0.5};  // 3/2 c^2 - 1
#pragma unroll
  for (int i = 0; i < NQ; ++i) {
    const int kx = cxq(i) + 1, ky = cyq(i) + 1, kz = czq(i) + 1;
    const double px = ph[0][kx], py = ph[1][ky], pz = ph[2][kz];
    double f = rho * (px * py * pz) - wq(i);
    f = f + S[0][0] * se[kx] * (py * pz);
    f = f + S[1][1] * se[ky] * (px * pz);
    f = f + S[2][2] * se[kz] * (px * py);
    f = f + S[0][1] * fi[0][kx] * fi[1][ky] * pz;
    f = f + S[0][2] * fi[0][kx] * fi[2][kz] * py;
    f = f + S[1][2] * fi[1][ky] * fi[2][kz] * px;
    pdf[i] = f;
  }
}

// One thread per ring cell x z.  ODD is the form of the sweep that ran this step.
template <int S, bool ODD, bool DIAG>
__global__ __launch_bounds__(128) void k_ext_ring(typename St<S>::T *__restrict__ A, Desc L, ExtDev E,
                                                  typename St<S>::C ws, typename St<S>::C wb, float scale,
                                                  float inv_scale, DiagAcc *__restrict__ diag) {
  using C = typename St<S>::C;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = t < (long long)E.R2 * E.Z;
  double part[5] = {0, 0, 0, 0, 0};  // source mass, momentum x/y/z; moment mismatch (max)
  double du2 = 0.0, ddr = 0.0, drho_min = 1e300;
  int bad = 0;
  if (active) {
    const int r = (int)(t / E.Z), z = (int)(t % E.Z);
    const int xy = E.band_xy[r], x = xy & 0xffff, y = xy >> 16;
    const long long cell = L.base + cell_of(L, x, y, z);
    // streamed (pre-collision) populations, read where this step's sweep would have
    C g[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      long long src;
      if constexpr (ODD) src = cell + (long long)opp(q) * L.QS + nb_off<false>(L, x, y, z, -cxq(q), -cyq(q), -czq(q));
      else src = cell + (long long)q * L.QS;
      g[q] = ld<S>(A + src, inv_scale);
    }
    // reconstruction of the predicted state; gradients over the prediction (ring) and the
    // band's post-stream state
    double v[4], gr[3][4], G[3][3];
    const double *pv = E.PRED + t * 4;
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = pv[c];
    ext_grad(E, E.PRED, x, y, z, v, gr);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) G[a][b] = gr[a][1 + b];
    double pdf[NQ];
    ext_regularized(v, G, E.omega, pdf);
    double dr = 0.0, jx = 0.0, jy = 0.0, jz = 0.0, sm = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      const double f = pdf[i], d = f - (double)g[i];
      dr += f; sm += d;
      if (cxq(i)) { jx += cxq(i) * f; sx += cxq(i) * d; }
      if (cyq(i)) { jy += cyq(i) * f; sy += cyq(i) * d; }
      if (czq(i)) { jz += czq(i) * f; sz += czq(i) * d; }
      g[i] = (C)f;
    }
    const double rho = 1.0 + dr, rv = exp(v[0]);
    part[0] = sm; part[1] = sx; part[2] = sy; part[3] = sz;
    double err = fabs(1.0 + dr - rv);
    err = fmax(err, fabs(jx - rv * v[1]));
    err = fmax(err, fabs(jy - rv * v[2]));
    err = fmax(err, fabs(jz - rv * v[3]));
    part[4] = err;
    // the ring's state for the next prediction: rho and u of the rebuilt populations
    const double ux = jx / rho, uy = jy / rho, uz = jz / rho;
    double *m = E.M + t * 4;
    m[0] = log(rho); m[1] = ux; m[2] = uy; m[3] = uz;
    if constexpr (DIAG) {
      const double u2 = ux * ux + uy * uy + uz * uz;
      if (!(isfinite(u2) && isfinite(dr))) bad = 1;
      else { du2 = u2; ddr = fabs(dr); drho_min = rho; }
    }
    C drho_c, u2_c;
    collide<C>(g, ws, wb, drho_c, u2_c);
    conserve_fp16<S>(g, scale, inv_scale);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      long long dst;
      if constexpr (ODD) dst = cell + (long long)q * L.QS + nb_off<false>(L, x, y, z, cxq(q), cyq(q), czq(q));
      else dst = cell + (long long)opp(q) * L.QS;
      stv<S>(A + dst, g[q], scale);
    }
  }
  // fixed-order block partials of the source and the mismatch
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1)
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      const double w = __shfl_xor(part[c], o, 64);
      part[c] = c == 4 ? fmax(part[c], w) : part[c] + w;
    }
  __shared__ double red[2][5];
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0)
    for (int c = 0; c < 5; ++c) red[w][c] = part[c];
  if constexpr (DIAG) {
    du2 = wave_max(du2); ddr = wave_max(ddr); drho_min = wave_min(drho_min);
    bad = __any(bad) ? 1 : 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double *p = E.part + (long long)blockIdx.x * 5;
    for (int c = 0; c < 5; ++c) p[c] = c == 4 ? fmax(red[0][c], red[1][c]) : red[0][c] + red[1][c];
  }
  if constexpr (DIAG) {
    if (lane == 0) {
      DiagSlot *sl = &diag->s[(blockIdx.x * 2 + w) % DIAG_SLOTS];
      atomicMax(&sl->u2max, (unsigned long long)__double_as_longlong(du2));
      atomicMax(&sl->drmax, (unsigned long long)__double_as_longlong(ddr));
      if (drho_min > 0.0) atomicMin(&sl->rhomin, (unsigned long long)__double_as_longlong(drho_min));
      else atomicMax(&sl->nonfinite, 2ull);
      if (bad) atomicMax(&sl->nonfinite, 1ull);
    }
  }
}

// ------------------------------------------------------------------ accumulators
// Lane sums (lanes 0..63 of one wave) of this step's wall mass exchange (wall partials,
// component 4) and of the ring source partials: acc[0][lane] and acc[1 + c][lane].
__device__ __forceinline__ void ext_lane_sums(const double *__restrict__ wall_part, int wall_blocks, int nelem,
                                              const ExtDev &E, int ring_blocks, double (*acc)[64], int lane) {
  double wm = 0.0, s[5] = {0, 0, 0, 0, 0};
  for (int b = lane; b < wall_blocks; b += 64) {
    const double *p = wall_part + (long long)b * nelem * 5 + 4;
#pragma unroll
    for (int e = 0; e < MAX_ELEM; ++e)
      if (e < nelem) wm += p[e * 5];
  }
  for (int b = lane; b < ring_blocks; b += 64)
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      const double x = E.part[(long long)b * 5 + c];
      s[c] = c == 4 ? fmax(s[c], x) : s[c] + x;
    }
  acc[0][lane] = wm;
  for (int c = 0; c < 5; ++c) acc[1 + c][lane] = s[c];
}

// One lane: fixed-order totals of the lane sums into the accumulators and, with a row, the
// exterior columns of that history row.
__device__ __forceinline__ void ext_acc_row(const double (*acc)[64], const ExtDev &E, double *__restrict__ row) {
  double tot[6] = {0, 0, 0, 0, 0, 0};
  for (int k = 0; k < 64; ++k)
    for (int c = 0; c < 6; ++c) tot[c] = c == 5 ? fmax(tot[c], acc[c][k]) : tot[c] + acc[c][k];
  ExtAcc &a = *E.acc;
  a.wall_mass = tot[0];
  a.wall_total += tot[0];
  for (int c = 0; c < 4; ++c) {
    a.src[c] = tot[1 + c];
    a.src_total[c] += tot[1 + c];
  }
  a.moment_err = tot[5];
  a.steps += 1;
  if (row) {
    row[EXT_COL] = a.wall_total;
    row[EXT_COL + 1] = a.src[0];
    row[EXT_COL + 2] = a.src_total[0];
  }
}

// End of an exterior step, one block of three waves whose load chains overlap: wave 2 sums
// the wall mass exchange and the ring source into the accumulators; on sampled steps wave 0
// reduces the diagnostics slots and wave 1 the wall loads.  After one barrier lane 0 of each
// wave writes its columns of the current history row (sampled steps), wave 0 clears the
// slots for the next sampled sweep and thread 0 advances the cursor (read by every thread
// before the barrier).
__global__ __launch_bounds__(192) void k_ext_end(const double *__restrict__ wall_part, int wall_blocks, int nelem,
                                                 ExtDev E, int ring_blocks, StepCtl *ctl, double *__restrict__ hist,
                                                 int sample, double fscale, double mscale, double vscale,
                                                 DiagAcc *diag) {
  __shared__ double acc[LOADS_LDS];
  __shared__ double eacc[6][64];
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const long long step = ctl->step;
  const int rowi = ctl->row;
  double *row = hist + (long long)rowi * ctl->width;
  DiagTot t{};
  if (w == 0) { if (sample) t = diag_reduce_wave(diag, lane); }
  else if (w == 1) { if (sample) loads_lane_sums(wall_part, wall_blocks, nelem, acc, lane); }
  else ext_lane_sums(wall_part, wall_blocks, nelem, E, ring_blocks, eacc, lane);
  __syncthreads();
  if (lane == 0) {
    if (w == 2) ext_acc_row(eacc, E, sample ? row : nullptr);
    else if (sample && w == 0) diag_row(t, row);
    else if (sample) loads_row(acc, nelem, fscale, mscale, vscale, row, step);
  }
  if (sample && w == 0) diag_clear_lanes(diag, lane);
  if (threadIdx.x == 0) {
    ctl->step = step + 1;
    if (sample) ctl->row = rowi + 1;
  }
}

// ------------------------------------------------------------------ set-up
// Ring cells at the free-stream equilibrium (rho = 1, u = (U, 0, 0)), post-even form
// (the reference sets every cell outside the physical box to the free stream,
// solver.py:196-199).
template <int S>
__global__ void k_ext_fill_ring(typename St<S>::T *__restrict__ A, Desc L, ExtDev E, float scale) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)E.R2 * E.Z) return;
  const int r = (int)(t / E.Z), z = (int)(t % E.Z);
  const int xy = E.band_xy[r], x = xy & 0xffff, y = xy >> 16;
  const long long c = L.base + cell_of(L, x, y, z);
#pragma unroll
  for (int i = 0; i < NQ; ++i) {
    const double g = equilibrium<double>(i, 1.0, E.U, 0.0, 0.0) - wq(i);
    stv<S>(A + c + (long long)opp(i) * L.QS, (typename St<S>::C)g, scale);
  }
}

// Initial incoming invariant of the two y rows, rows[side][z][x] = u.n - cs ln rho with
// n = (0, -1, 0) / (0, 1, 0), from the band state of the first step (exterior.py:96-97).
__global__ void k_ext_rows(ExtDev E, double *__restrict__ rows) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2LL * E.Z * E.X) return;
  const int x = (int)(t % E.X), z = (int)((t / E.X) % E.Z), side = (int)(t / ((long long)E.X * E.Z));
  const int y = side == 0 ? 0 : E.Y - 1;
  double v[4];
  ext_at(E, E.M, x, y, z, v);
  rows[t] = (side == 0 ? -1.0 : 1.0) * v[2] - EXT_CS * v[0];
}

// One collision in place of every non-solid cell of a post-even state (the reference's
// collide_stream order collides the initial state before its first streaming).
template <int S>
__global__ __launch_bounds__(256) void k_collide_inplace(typename St<S>::T *__restrict__ A,
                                                         const uint8_t *__restrict__ solid, Desc L,
                                                         typename St<S>::C ws, typename St<S>::C wb, float scale,
                                                         float inv_scale) {
  using C = typename St<S>::C;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)L.nx * L.ny * L.nz) return;
  const int x = (int)(t % L.nx), y = (int)((t / L.nx) % L.ny), z = (int)(t / ((long long)L.nx * L.ny));
  if (solid[(long long)(y + 1) * (L.nx + 2) + (x + 1)] == 1) return;
  const long long c = L.base + cell_of(L, x, y, z);
  C g[NQ];
#pragma unroll
  for (int i = 0; i < NQ; ++i) g[i] = ld<S>(A + c + (long long)opp(i) * L.QS, inv_scale);
  C dr, u2;
  collide<C>(g, ws, wb, dr, u2);
  conserve_fp16<S>(g, scale, inv_scale);
#pragma unroll
  for (int i = 0; i < NQ; ++i) stv<S>(A + c + (long long)opp(i) * L.QS, g[i], scale);
}

}  // namespace ak
