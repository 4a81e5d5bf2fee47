// ak_refined_capi.hpp -- C ABI of the two-level 2:1 refined step (included by ak_capi.hip).
//
// One refined case = a coarse lattice (legacy farfield, z periodic, covered cells kept
// as donors) + a fine lattice over the box plus two donor layers in x and y (z
// periodic, Bouzidi walls).  One coarse step (10 launches):
//   coarse sweep -> [coarse face flux, coarse boundary, donor moments at t + dt]
//   -> prolongation x pass of both substeps -> 2 x [prolongation y/z + inject, fine sweep,
//   [fine face flux, fine walls + load partials]] -> [reflux, loads fold or tick,
//   restriction, donor moments at t of the next step]
// (brackets: one launch).  A run call starts with the donor moments of the current state.
#include "ak_refined.hpp"

namespace {

struct Refined {
  Lattice *c = nullptr, *f = nullptr;
  BoxGeom B{};
  Levels V{};
  FreeState F{};
  double dxc = 0, dtc = 0, dxf = 0, dtf = 0, span = 0, U = 0;
  // transfer buffers hold values of the transfer arithmetic TR<S> (FP64 for FP64 storage,
  // FP32 otherwise), sized for doubles and read through tr<S>()
  double *d_before = nullptr, *d_after = nullptr, *d_regc = nullptr, *d_regf = nullptr;
  double *d_px = nullptr;                   // prolongation x pass of both substeps
  double *d_part = nullptr;                 // wall load partials (FP64)
  int32_t *d_lpx = nullptr, *d_cols = nullptr;  // x-pass work items; fine interface-band columns
  int n_lpx = 0, n_cols = 0;
  int *d_err = nullptr;
  FluxTable T{};
  FluxTable *d_T = nullptr;
  int nt = 0, wall_blocks = 0;
  long long steps = 0;
  hipStream_t own_stream = nullptr;  // created when the caller passes the null stream
  bool use_graphs = true;
  GraphCache graphs;  // slot = coarse parity * 4 + fine parity * 2 + sampled
  // conservation budget: accumulators, outer-face flux partials of a step, integral partials
  BudgetAcc *d_acc = nullptr;
  double *d_opart = nullptr, *d_ipart = nullptr;
  int nb_outer = 0, nb_int_c = 0, nb_int_f = 0;
  bool budget_complete = true;  // false once a batch has stepped the case (no budget there)
};

Refined *R(void *h) { return (Refined *)h; }

// A transfer buffer as the kernels of storage S read it.
template <int S>
TR<S> *tr(double *p) { return (TR<S> *)p; }

// Device address of a double field of the case's budget accumulators.
double *acc_field(Refined *r, size_t offset) { return (double *)((char *)r->d_acc + offset); }

// host copies of the transfer algebra (same formulas as the device)
void host_populations(const double m[10], double g[27]) {
  const double rho = m[0];
  const double u[3] = {m[1] / rho, m[2] / rho, m[3] / rho};
  double ph[3][3];
  for (int a = 0; a < 3; ++a) {
    ph[a][1] = 2.0 / 3.0 - u[a] * u[a];
    ph[a][0] = (1.0 / 3.0 + u[a] * u[a] - u[a]) * 0.5;
    ph[a][2] = (1.0 / 3.0 + u[a] * u[a] + u[a]) * 0.5;
  }
  for (int i = 0; i < 27; ++i) {
    const int cx = cxq(i), cy = cyq(i), cz = czq(i);
    const double eq = rho * ph[0][cx + 1] * ph[1][cy + 1] * ph[2][cz + 1];
    const double h = (cx * cx - 1.0 / 3.0) * m[4] + (cy * cy - 1.0 / 3.0) * m[5] + (cz * cz - 1.0 / 3.0) * m[6]
                   + 2.0 * (cx * cy * m[7] + cx * cz * m[8] + cy * cz * m[9]);
    g[i] = eq + wq(i) * 4.5 * h - wq(i);
  }
}
void host_moments(const double g[27], double m[10]) {
  double d = 0, j[3] = {0, 0, 0}, s[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 27; ++i) {
    const int c[3] = {cxq(i), cyq(i), czq(i)};
    d += g[i];
    for (int a = 0; a < 3; ++a) j[a] += c[a] * g[i];
    s[0] += c[0] * c[0] * g[i]; s[1] += c[1] * c[1] * g[i]; s[2] += c[2] * c[2] * g[i];
    s[3] += c[0] * c[1] * g[i]; s[4] += c[0] * c[2] * g[i]; s[5] += c[1] * c[2] * g[i];
  }
  const double rho = d + 1.0;
  m[0] = rho; m[1] = j[0]; m[2] = j[1]; m[3] = j[2];
  m[4] = s[0] - d / 3.0 - j[0] * j[0] / rho;
  m[5] = s[1] - d / 3.0 - j[1] * j[1] / rho;
  m[6] = s[2] - d / 3.0 - j[2] * j[2] / rho;
  m[7] = s[3] - j[0] * j[1] / rho;
  m[8] = s[4] - j[0] * j[2] / rho;
  m[9] = s[5] - j[1] * j[2] / rho;
}

void build_flux_table(FluxTable &T) {
  for (int axis = 0; axis < 2; ++axis) {
    const int other = 1 - axis;
    int n = 0;
    for (int d = 0; d < 27; ++d) {
      const int c[3] = {cxq(d), cyq(d), czq(d)};
      const int nrm = c[axis];
      if (!nrm) continue;
      // tangential axes in ascending order: the in-plane one and z
      int moving[2], m = 0;
      if (c[other]) moving[m++] = other;
      if (c[2]) moving[m++] = 2;
      auto fact = [](int k) { int r = 1; for (int i = 2; i <= k; ++i) r *= i; return r; };
      for (int mask = 0; mask < (1 << m); ++mask) {
        const int cnt = __builtin_popcount(mask);
        const double w = (double)fact(cnt) * fact(m - cnt) / fact(m + 1);
        FluxTerm &tm = T.term[axis][n++];
        tm.d = (int8_t)d;
        tm.s0 = 0; tm.s1 = 0;
        for (int k = 0; k < m; ++k)
          if (mask & (1 << k)) {
            if (moving[k] == other) tm.s0 = (int8_t)c[other];
            else tm.s1 = (int8_t)c[2];
          }
        tm.coef[0] = w * nrm;
        for (int a = 0; a < 3; ++a) tm.coef[1 + a] = w * nrm * c[a];
      }
    }
    T.n[axis] = n;
  }
}

// Work items of the prolongation: the x-pass indices for which prolong_x_body does work
// (same condition), and the fine columns outside the owned box (y/z passes + injection).
int build_prolong_lists(Refined *r) {
  const BoxGeom &B = r->B;
  const int dy = B.nb[1] + 6, dz = B.nb[2] + 6, cxn = 2 * B.nb[0] + 4, p = B.pad;
  const Desc &L = r->f->L;
  if ((long long)cxn * dy * dz >= INT32_MAX || (long long)L.nx * L.ny >= INT32_MAX)
    return fail(nullptr, "refined box too large for 32-bit prolongation work items");
  if (prolong_yz_lds(B, r->c->storage == 0 ? sizeof(double) : sizeof(float)) > 65536)
    return fail(nullptr, "periodic span too long for the prolongation tiles");
  std::vector<int32_t> lx, cols;
  for (int qz = 0; qz < dz; ++qz)
    for (int qy = 0; qy < dy; ++qy)
      for (int cx = 0; cx < cxn; ++cx)
        if (!(cx >= p && cx < cxn - p && qy > 4 && qy < B.nb[1] + 1))
          lx.push_back((int32_t)(((long long)qz * dy + qy) * cxn + cx));
  for (int y = 0; y < L.ny; ++y)
    for (int x = 0; x < L.nx; ++x)
      if (!(x >= p && x < L.nx - p && y >= p && y < L.ny - p)) cols.push_back((int32_t)((long long)y * L.nx + x));
  r->n_lpx = (int)lx.size();
  r->n_cols = (int)cols.size();
  return (upload(r->c, &r->d_lpx, lx) || upload(r->c, &r->d_cols, cols)) ? -1 : 0;
}

// restriction + the next step's donor moments (at t) in one launch
template <int S>
void launch_restrict(Refined *r) {
  Lattice *c = r->c, *f = r->f;
  using T_ = typename St<S>::T;
  const int ntiles = restrict_tiles(r->B);
  const long long ndon = (long long)(r->B.nb[0] + 6) * (r->B.nb[1] + 6) * (r->B.nb[2] + 6);
  const dim3 grid(ntiles + blocks_for(ndon, 256)), block(256);
  const bool pf = f->sweeps % 2 == 0, pc = c->sweeps % 2 == 0;
#define AK_RT(PF, PC)                                                                                            \
  hipLaunchKernelGGL((k_restrict_tiles<S, PF, PC>), grid, block, 0, c->stream, (const T_ *)f->pop, f->L,         \
                     f->inv_scale, (T_ *)c->pop, c->L, r->B, r->V, r->F, c->scale, c->inv_scale,                \
                     tr<S>(r->d_before), r->d_err, ntiles)
  if (pf) { if (pc) AK_RT(true, true); else AK_RT(true, false); }
  else { if (pc) AK_RT(false, true); else AK_RT(false, false); }
#undef AK_RT
}

// The coarse step in phases.  A single case runs them in order with its own sweeps in
// between (refined_coarse_step); a batch runs each phase for every member and one
// batched sweep launch in place of the members' sweeps (ak_batch_run).  Every launch
// goes to the member's coarse-lattice stream, which both levels share.

// donor moments at time t (pre-collision coarse state, before the coarse sweep): at the
// start of a run; every step's restriction launch then leaves those of the next step
template <int S>
void phase_before_coarse(Refined *r) {
  Lattice *c = r->c;
  using T_ = typename St<S>::T;
  const long long ndon = (long long)(r->B.nb[0] + 6) * (r->B.nb[1] + 6) * (r->B.nb[2] + 6);
  if (c->sweeps % 2 == 0)
    hipLaunchKernelGGL((k_donor_moments<S, true>), dim3(blocks_for(ndon, 128)), dim3(128), 0, c->stream, (const T_ *)c->pop, c->L, r->B, r->F, c->inv_scale, tr<S>(r->d_before), r->d_err);
  else
    hipLaunchKernelGGL((k_donor_moments<S, false>), dim3(blocks_for(ndon, 128)), dim3(128), 0, c->stream, (const T_ *)c->pop, c->L, r->B, r->F, c->inv_scale, tr<S>(r->d_before), r->d_err);
}

// after the coarse sweep: coarse face flux, coarse farfield, donor moments at t + dt
template <int S>
void phase_after_coarse(Refined *r) {
  Lattice *c = r->c;
  using T_ = typename St<S>::T;
  const long long ndon = (long long)(r->B.nb[0] + 6) * (r->B.nb[1] + 6) * (r->B.nb[2] + 6);
  const unsigned cflux_blocks = blocks_for(4LL * r->nt * c->nz * FLUX_LANES, 256);
  const TR<S> volc = TR<S>(r->dxc * r->dxc * r->dxc);
  // one launch: face flux, legacy farfield and donor moments (disjoint slots)
  const LegacyLinks K{c->d_legx, c->d_legz, c->d_legdir, c->d_legkind, c->d_pdf, c->d_pdfnd, c->nleg};
  const int nb_leg = (int)blocks_for(c->nleg, 256);
  const dim3 grid(cflux_blocks + nb_leg + blocks_for(ndon, 256)), block(256);
  if (c->sweeps % 2 == 0)
    hipLaunchKernelGGL((k_after_coarse<S, true>), grid, block, 0, c->stream, (T_ *)c->pop, c->L, r->B, r->d_T, volc,
                       c->scale, c->inv_scale, tr<S>(r->d_regc), K, c->U, r->F, tr<S>(r->d_after), r->d_err,
                       (int)cflux_blocks, nb_leg);
  else
    hipLaunchKernelGGL((k_after_coarse<S, false>), grid, block, 0, c->stream, (T_ *)c->pop, c->L, r->B, r->d_T, volc,
                       c->scale, c->inv_scale, tr<S>(r->d_regc), K, c->U, r->F, tr<S>(r->d_after), r->d_err,
                       (int)cflux_blocks, nb_leg);
  // x pass of the prolongation for both fine substeps (time fractions 0 and 1/2); 64-thread
  // blocks spread the few thousand items over every CU
  hipLaunchKernelGGL(k_prolong_x<TR<S>>, dim3(blocks_for(r->n_lpx, 64)), dim3(64), 0, c->stream,
                     (const TR<S> *)tr<S>(r->d_before), (const TR<S> *)tr<S>(r->d_after), r->B, tr<S>(r->d_px),
                     r->d_lpx, r->n_lpx);
}

// fine substep `sub` before its sweep: y and z passes of the time-interpolated prolongation
// (x pass from phase_after_coarse) into the donor layers
template <int S>
void phase_before_fine(Refined *r, int sub) {
  Lattice *f = r->f;
  using T_ = typename St<S>::T;
  const TR<S> *X = tr<S>(r->d_px) + sub * prolong_x_half(r->B);
  const dim3 gyz(blocks_for(r->n_cols, PZ_COLS));
  const size_t lds = prolong_yz_lds(r->B, sizeof(TR<S>));
  if (f->sweeps % 2 == 0)
    hipLaunchKernelGGL((k_prolong_yz<S, true>), gyz, dim3(64), lds, f->stream, X, (T_ *)f->pop, f->L, r->B, r->V, f->scale, r->d_err, r->d_cols, r->n_cols);
  else
    hipLaunchKernelGGL((k_prolong_yz<S, false>), gyz, dim3(64), lds, f->stream, X, (T_ *)f->pop, f->L, r->B, r->V, f->scale, r->d_err, r->d_cols, r->n_cols);
}

// fine substep `sub` after its sweep: fine face flux, walls + load partials, and after the
// first substep the coarse domain's outer-face flux of this coarse step (budget)
template <int S>
void phase_after_fine(Refined *r, int sub) {
  Lattice *f = r->f, *c = r->c;
  using T_ = typename St<S>::T;
  const TR<S> volf = TR<S>(r->dxf * r->dxf * r->dxf);
  const unsigned fflux_blocks = blocks_for(4LL * (2 * r->nt) * f->nz * FLUX_LANES, 256);
  WallLinks W{f->d_wxy, f->d_wdir, f->d_wq, f->d_welem, f->nlinks, f->nelem, f->rx_off, f->ry_off, r->dxf};
  double *part = r->d_part + (size_t)sub * r->wall_blocks * f->nelem * 5;
  const int nb_walls = f->nlinks ? r->wall_blocks : 0;
  const int nb_outer = sub == 0 ? r->nb_outer : 0;
  // one launch: fine face flux, the walls with their loads and the outer flux (disjoint slots)
  const dim3 grid(fflux_blocks + nb_walls + nb_outer), block(256);
  const bool pf = f->sweeps % 2 == 0, pc = c->sweeps % 2 == 0;
#define AK_AF(PF, PC)                                                                                              \
  hipLaunchKernelGGL((k_after_fine<S, PF, PC>), grid, block, 0, f->stream, (T_ *)f->pop, f->L, r->B, r->d_T, volf, \
                     f->scale, f->inv_scale, sub, tr<S>(r->d_regf), W, part, (int)fflux_blocks, nb_walls,          \
                     (const T_ *)c->pop, c->L, c->inv_scale, r->F, r->d_opart)
  if (pf) { if (pc) AK_AF(true, true); else AK_AF(true, false); }
  else { if (pc) AK_AF(false, true); else AK_AF(false, false); }
#undef AK_AF
}

// end of the coarse step: reflux, restriction, loads (sampled) or the step tick
template <int S>
void phase_end(Refined *r, bool sample) {
  Lattice *c = r->c, *f = r->f;
  using T_ = typename St<S>::T;
  const unsigned cwave_blocks = blocks_for(4LL * r->nt * c->nz * 64, 256);
  const double inv_volc = 1.0 / (r->dxc * r->dxc * r->dxc), volf = r->dxf * r->dxf * r->dxf;
  // impulse over the two substeps: F dx_f^3 (dx_f/dt_f) / dt_c; q_area = span / 2.  The refined
  // history carries loads only (field diagnostics are read on demand).  The reflux launch's
  // last block folds the loads (sampled step) or ticks the step; the loads read only the
  // wall partials, the reflux only the coarse populations.
  const double fscale = volf * (r->dxf / r->dtf) / r->dtc / (0.5 * r->span);
  const int nparts = f->nlinks ? 2 * r->wall_blocks : 0;
  // reflux, loads / tick, restriction and the next step's donor moments: one launch
  const int ntiles = restrict_tiles(r->B);
  const long long ndon = (long long)(r->B.nb[0] + 6) * (r->B.nb[1] + 6) * (r->B.nb[2] + 6);
  const dim3 grid(cwave_blocks + 1 + ntiles + blocks_for(ndon, 256)), block(256);
  const bool pf = f->sweeps % 2 == 0, pc = c->sweeps % 2 == 0;
#define AK_END(PF, PC)                                                                                             \
  hipLaunchKernelGGL((k_step_end<S, PF, PC>), grid, block, 0, c->stream, (const T_ *)f->pop, f->L, f->inv_scale,  \
                     (T_ *)c->pop, c->L, r->B, r->V, r->F, c->scale, c->inv_scale, tr<S>(r->d_before),            \
                     tr<S>(r->d_regc), tr<S>(r->d_regf), TR<S>(inv_volc), r->d_err, (int)cwave_blocks, ntiles,    \
                     (const double *)r->d_part, nparts,                                                           \
                     f->nelem, fscale, volf, c->d_ctl, c->d_hist, sample ? 1 : 0, G)
  const BudgetArgs G{r->d_opart, r->nb_outer, r->dxc * r->dxc * r->dxc, r->dxc / r->dtc, r->d_acc};
  if (pf) { if (pc) AK_END(true, true); else AK_END(true, false); }
  else { if (pc) AK_END(false, true); else AK_END(false, false); }
#undef AK_END
}

// Donor moments of the current state, launched directly (not in a step graph) at the start
// of every run call: the state may have been initialised, loaded or set since the last
// step's restriction wrote them.
void run_start(Refined *r) {
  switch (r->c->storage) {
    case 0: phase_before_coarse<0>(r); break;
    case 1: phase_before_coarse<1>(r); break;
    default: phase_before_coarse<2>(r); break;
  }
}

// Conserved integrals of both levels' owned cells (pre-collision state) into dst[4] (a
// field of the device accumulators); with step > 0 the fold also records that step.
template <int S>
void launch_integrals(Refined *r, double *dst, long long step) {
  Lattice *c = r->c, *f = r->f;
  using T_ = typename St<S>::T;
  double *pc = r->d_ipart, *pf = r->d_ipart + (size_t)r->nb_int_c * 4;
  if (c->sweeps % 2 == 0)
    hipLaunchKernelGGL((k_integrals<S, true>), dim3(r->nb_int_c), dim3(256), 0, c->stream, (const T_ *)c->pop, c->d_solid, c->L, c->inv_scale, pc);
  else
    hipLaunchKernelGGL((k_integrals<S, false>), dim3(r->nb_int_c), dim3(256), 0, c->stream, (const T_ *)c->pop, c->d_solid, c->L, c->inv_scale, pc);
  if (f->sweeps % 2 == 0)
    hipLaunchKernelGGL((k_integrals<S, true>), dim3(r->nb_int_f), dim3(256), 0, c->stream, (const T_ *)f->pop, f->d_solid, f->L, f->inv_scale, pf);
  else
    hipLaunchKernelGGL((k_integrals<S, false>), dim3(r->nb_int_f), dim3(256), 0, c->stream, (const T_ *)f->pop, f->d_solid, f->L, f->inv_scale, pf);
  hipLaunchKernelGGL(k_integrals_final, dim3(1), dim3(64), 0, c->stream, (const double *)r->d_ipart, r->nb_int_c,
                     r->nb_int_f, r->dxc * r->dxc * r->dxc, r->dxf * r->dxf * r->dxf, r->dxc / r->dtc, dst,
                     step > 0 ? r->d_acc : nullptr, step);
}
void integrals_any(Refined *r, double *dst, long long step) {
  switch (r->c->storage) {
    case 0: launch_integrals<0>(r, dst, step); break;
    case 1: launch_integrals<1>(r, dst, step); break;
    default: launch_integrals<2>(r, dst, step); break;
  }
}

// One coarse step of one case; its number and history row come from the coarse lattice's
// d_ctl.  ev (optional, 8 events): step start/end, coarse sweep start/end, fine sweep 0 and
// 1 start/end (the sweep events hold the kernels' own start and end times).
template <int S>
void refined_coarse_step(Refined *r, bool sample, hipEvent_t *ev = nullptr) {
  Lattice *c = r->c, *f = r->f;
  if (ev) (void)hipEventRecord(ev[0], c->stream);
  launch_sweep<S>(c, c->sweeps % 2 == 0, false, ev ? ev[2] : nullptr, ev ? ev[3] : nullptr);
  c->sweeps++;
  phase_after_coarse<S>(r);
  for (int sub = 0; sub < 2; ++sub) {
    phase_before_fine<S>(r, sub);
    launch_sweep<S>(f, f->sweeps % 2 == 0, false, ev ? ev[4 + 2 * sub] : nullptr, ev ? ev[5 + 2 * sub] : nullptr);
    f->sweeps++;
    phase_after_fine<S>(r, sub);
  }
  phase_end<S>(r, sample);
  if (ev) (void)hipEventRecord(ev[1], c->stream);
}

}  // namespace

extern "C" {

int ak_refined_create(void **handle, int device, int nx, int ny, int nz, const int32_t *box_lo,
                      const int32_t *box_hi, int storage, const double *level, double lattice_speed,
                      double fp16_scale, int history_capacity, void *stream) {
  if (!handle) return fail(nullptr, "null handle pointer");
  *handle = nullptr;
  // level = {dx_c, dt_c, tau_shear_c, tau_bulk_c}
  const double dxc = level[0], dtc = level[1], tsc = level[2], tbc = level[3];
  if (!(dxc > 0 && dtc > 0 && tsc > 0.5 && tbc > 0.5)) return fail(nullptr, "invalid coarse level");
  for (int a = 0; a < 2; ++a)
    if (box_lo[a] < 3 || box_hi[a] > (a ? ny : nx) - 3 || box_hi[a] - box_lo[a] < 8)
      return fail(nullptr, "fine box must keep three donor cells inside the domain and span >= 8 cells");
  if (box_lo[2] != 0 || box_hi[2] != nz) return fail(nullptr, "fine box must span the periodic z axis");
  Refined *r = new Refined();
  if (!stream) {
    // both levels share one stream (the coarse and fine kernels of a step are ordered on it)
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&r->own_stream) != hipSuccess) {
      delete r;
      return fail(nullptr, "hipStreamCreate failed");
    }
    stream = r->own_stream;
  }
  const double tsf = 0.5 + 2 * (tsc - 0.5), tbf = 0.5 + 2 * (tbc - 0.5);
  r->dxc = dxc; r->dtc = dtc; r->dxf = dxc / 2; r->dtf = dtc / 2; r->U = lattice_speed;
  r->span = nz * dxc;
  r->V.shear_cf = tsf * r->dtf / (tsc * dtc);
  r->V.bulk_cf = tbf * r->dtf / (tbc * dtc);
  r->V.shear_fc = tsc * dtc / (tsf * r->dtf);
  r->V.bulk_fc = tbc * dtc / (tbf * r->dtf);
  for (int a = 0; a < 3; ++a) { r->B.lo[a] = box_lo[a]; r->B.hi[a] = box_hi[a]; r->B.nb[a] = box_hi[a] - box_lo[a]; }
  r->B.pad = 2;
  double ff[10] = {1.0, lattice_speed, 0, 0, 0, 0, 0, 0, 0, 0};
  host_populations(ff, r->F.g);
  host_moments(r->F.g, r->F.m);
  build_flux_table(r->T);
  r->nt = std::max(r->B.nb[0], r->B.nb[1]);
  void *hc = nullptr, *hf = nullptr;
  if (ak_lattice_create(&hc, device, nx, ny, nz, storage, 1.0 / tsc, 1.0 / tbc, fp16_scale, history_capacity, stream)) {
    if (r->own_stream) (void)hipStreamDestroy(r->own_stream);
    delete r;
    return -1;
  }
  if (ak_lattice_create(&hf, device, 2 * r->B.nb[0] + 4, 2 * r->B.nb[1] + 4, 2 * nz, storage, 1.0 / tsf, 1.0 / tbf,
                        fp16_scale, 16, stream)) {
    ak_lattice_destroy(hc);
    if (r->own_stream) (void)hipStreamDestroy(r->own_stream);
    delete r;
    return -1;
  }
  r->c = (Lattice *)hc;
  r->f = (Lattice *)hf;
  r->f->ff_kind = AK_FARFIELD_NONE;
  const size_t ndon = (size_t)(r->B.nb[0] + 6) * (r->B.nb[1] + 6) * (r->B.nb[2] + 6) * 10;
  const size_t nreg = (size_t)4 * r->nt * nz * 4;
  if (dalloc(r->c, &r->d_before, ndon) || dalloc(r->c, &r->d_after, ndon) || dalloc(r->c, &r->d_regc, nreg) ||
      dalloc(r->c, &r->d_regf, 4 * nreg) ||
      dalloc(r->c, &r->d_px, 2 * prolong_x_half(r->B)) ||
      dalloc(r->c, &r->d_err, 1) || dalloc(r->c, &r->d_T, 1)) {
    ak_lattice_destroy(hf); ak_lattice_destroy(hc);
    if (r->own_stream) (void)hipStreamDestroy(r->own_stream);
    delete r;
    return -1;
  }
  // budget: outer-face flux partials (one per block of the outer range), integral partials
  r->nb_outer = (int)blocks_for(2LL * ((long long)ny * nz + (long long)nx * nz) * FLUX_LANES, 256);
  r->nb_int_c = (int)blocks_for((long long)r->c->nx * r->c->ny * r->c->nz, 256);
  r->nb_int_f = (int)blocks_for((long long)r->f->nx * r->f->ny * r->f->nz, 256);
  if (dalloc(r->c, &r->d_acc, 1) || dalloc(r->c, &r->d_opart, (size_t)r->nb_outer * 4) ||
      dalloc(r->c, &r->d_ipart, (size_t)(r->nb_int_c + r->nb_int_f) * 4)) {
    ak_lattice_destroy(hf); ak_lattice_destroy(hc);
    if (r->own_stream) (void)hipStreamDestroy(r->own_stream);
    delete r;
    return -1;
  }
  r->c->pre_collision = r->f->pre_collision = true;
  hipLaunchKernelGGL(k_budget_reset, dim3(1), dim3(64), 0, r->c->stream, r->d_acc);
  (void)hipMemsetAsync(r->d_err, 0, sizeof(int), r->c->stream);
  if (build_prolong_lists(r)) {
    ak_lattice_destroy(hf); ak_lattice_destroy(hc);
    if (r->own_stream) (void)hipStreamDestroy(r->own_stream);
    delete r;
    return -1;
  }
  if (hipMemcpy(r->d_T, &r->T, sizeof(FluxTable), hipMemcpyHostToDevice) != hipSuccess) {
    ak_lattice_destroy(hf); ak_lattice_destroy(hc);
    if (r->own_stream) (void)hipStreamDestroy(r->own_stream);
    delete r;
    return fail(nullptr, "flux table upload failed");
  }
  *handle = r;
  return 0;
}

int ak_refined_destroy(void *handle) {
  Refined *r = R(handle);
  if (!r) return 0;
  (void)hipSetDevice(r->c->dev);
  (void)hipStreamSynchronize(r->c->stream);
  r->graphs.clear();
  ak_lattice_destroy(r->f);  // partials are owned by the fine lattice
  ak_lattice_destroy(r->c);  // transfer buffers are owned by the coarse lattice
  if (r->own_stream) (void)hipStreamDestroy(r->own_stream);
  delete r;
  return 0;
}

void *ak_refined_level(void *handle, int level) {
  Refined *r = R(handle);
  return r ? (level ? (void *)r->f : (void *)r->c) : nullptr;
}

// Fine-level walls (padded fine coordinates); lever arm in common units about
// the common origin: ((index + 0.5) dx_f + origin_f).
int ak_refined_set_walls(void *handle, int n, const int32_t *x, const int32_t *y, const int32_t *d, const double *q,
                         const int32_t *element, int nelem, double origin_fx, double origin_fy) {
  Refined *r = R(handle);
  if (!r) return fail(nullptr, "null refined handle");
  if (ak_set_walls(r->f, n, x, y, d, q, element, nelem, origin_fx / r->dxf, origin_fy / r->dxf, 1.0, r->U)) return -1;
  r->wall_blocks = r->f->wall_blocks;
  if (r->d_part) { dfree(r->f, r->d_part); r->d_part = nullptr; }
  if (dalloc(r->f, &r->d_part, (size_t)std::max(r->wall_blocks, 1) * nelem * 5 * 2)) return -1;
  return 0;
}

// Initial state: seeded field on both levels (p = {U, amp, phy, phz, ky, kz, lex, ley,
// plan_ox, plan_oy}), restriction of the fine state into the covered coarse cells,
// legacy farfield on the coarse level.
int ak_refined_init_seeded(void *handle, const double *p) {
  Refined *r = R(handle);
  if (!r) return fail(nullptr, "null refined handle");
  Lattice *c = r->c, *f = r->f;
  if (check_ready(c) || check_ready(f)) return -1;
  for (Lattice *s : {c, f}) {
    hipLaunchKernelGGL((k_fill<uint8_t>), dim3(1024), dim3(256), 0, s->stream, (uint8_t *)s->pop, (long long)s->pop_bytes, (uint8_t)0);
    s->sweeps = 1;  // pre-collision state stored in streamed form: the next sweep is even
    s->gen++;
    s->hist_rows = 0;
    s->host_hist.clear();
  }
  LevelSeed sc{p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], 0.0, 0.0, 0.0, r->dxc, p[8], p[9]};
  LevelSeed sf = sc;
  sf.dx = r->dxf;
  sf.ox = r->B.lo[0] * r->dxc - r->B.pad * r->dxf;
  sf.oy = r->B.lo[1] * r->dxc - r->B.pad * r->dxf;
  (void)hipMemsetAsync(r->d_err, 0, sizeof(int), r->c->stream);
  const long long ncv = (long long)c->nx * c->ny * c->nz, nfv = (long long)f->nx * f->ny * f->nz;
  switch (c->storage) {
    case 0:
      hipLaunchKernelGGL((k_init_level<0>), dim3(blocks_for(ncv, 256)), dim3(256), 0, c->stream, (double *)c->pop, c->d_solid, c->L, sc, c->scale, r->d_err);
      hipLaunchKernelGGL((k_init_level<0>), dim3(blocks_for(nfv, 256)), dim3(256), 0, f->stream, (double *)f->pop, f->d_solid, f->L, sf, f->scale, r->d_err);
      launch_restrict<0>(r);
      break;
    case 1:
      hipLaunchKernelGGL((k_init_level<1>), dim3(blocks_for(ncv, 256)), dim3(256), 0, c->stream, (float *)c->pop, c->d_solid, c->L, sc, c->scale, r->d_err);
      hipLaunchKernelGGL((k_init_level<1>), dim3(blocks_for(nfv, 256)), dim3(256), 0, f->stream, (float *)f->pop, f->d_solid, f->L, sf, f->scale, r->d_err);
      launch_restrict<1>(r);
      break;
    default:
      hipLaunchKernelGGL((k_init_level<2>), dim3(blocks_for(ncv, 256)), dim3(256), 0, c->stream, (_Float16 *)c->pop, c->d_solid, c->L, sc, c->scale, r->d_err);
      hipLaunchKernelGGL((k_init_level<2>), dim3(blocks_for(nfv, 256)), dim3(256), 0, f->stream, (_Float16 *)f->pop, f->d_solid, f->L, sf, f->scale, r->d_err);
      launch_restrict<2>(r);
      break;
  }
  HIPCHK(c, hipGetLastError());
  if (ak_set_farfield(c, AK_FARFIELD_LEGACY, r->U)) return -1;
  r->steps = 0;
  // budget from the initial state: accumulators cleared, integrals of the initial state
  hipLaunchKernelGGL(k_budget_reset, dim3(1), dim3(64), 0, c->stream, r->d_acc);
  integrals_any(r, acc_field(r, offsetof(BudgetAcc, initial)), 0);
  r->budget_complete = true;
  HIPCHK(c, hipGetLastError());
  return 0;
}

int ak_refined_run(void *handle, int nsteps, long long first_step, int sample_every) {
  Refined *r = R(handle);
  if (!r) return fail(nullptr, "null refined handle");
  Lattice *c = r->c, *f = r->f;
  if (check_ready(c) || check_ready(f)) return -1;
  if (nsteps > 0) {
    set_ctl(c, first_step + 1, c->hist_rows);
    run_start(r);
  }
  for (int k = 0; k < nsteps; ++k) {
    const long long number = first_step + k + 1;
    const bool sample = sample_every > 0 && number % sample_every == 0;
    if (sample && c->hist_rows >= c->hist_cap) {
      if (flush_history(c)) return -1;
      set_ctl(c, number, 0);
    }
    auto launch = [&] {
      switch (c->storage) {
        case 0: refined_coarse_step<0>(r, sample); break;
        case 1: refined_coarse_step<1>(r, sample); break;
        default: refined_coarse_step<2>(r, sample); break;
      }
    };
    const bool last = k == nsteps - 1;
    if (last) integrals_any(r, acc_field(r, offsetof(BudgetAcc, before)), 0);  // the budget brackets the call's last step
    if (r->use_graphs) {
      const int slot = (int)(c->sweeps % 2) * 4 + (int)(f->sweeps % 2) * 2 + (sample ? 1 : 0);
      const unsigned long long key = (c->gen << 32) ^ f->gen;
      if (graph_step(c, c->stream, r->graphs, slot, key, launch, [&] { c->sweeps += 1; f->sweeps += 2; })) return -1;
    } else {
      launch();
    }
    if (sample) c->hist_rows++;
    r->steps++;
    if (last) integrals_any(r, acc_field(r, offsetof(BudgetAcc, after)), number);
  }
  HIPCHK(c, hipGetLastError());
  return 0;
}

int ak_refined_set_graph_mode(void *handle, int enable) {
  Refined *r = R(handle);
  if (!r) return fail(nullptr, "null refined handle");
  r->use_graphs = enable != 0;
  return 0;
}

// ak_refined_run with HIP events (measurement): out = {total ms, coarse sweep ms, fine sweep ms,
// coarse steps}; synchronises at the end.
int ak_refined_run_profiled(void *handle, int nsteps, long long first_step, int sample_every, double *out) {
  Refined *r = R(handle);
  if (!r) return fail(nullptr, "null refined handle");
  Lattice *c = r->c;
  if (check_ready(c) || check_ready(r->f)) return -1;
  std::vector<hipEvent_t> ev(8 * (size_t)nsteps);
  for (auto &e : ev) HIPCHK(c, hipEventCreate(&e));
  if (nsteps > 0) {
    set_ctl(c, first_step + 1, c->hist_rows);
    run_start(r);
  }
  for (int k = 0; k < nsteps; ++k) {
    const long long number = first_step + k + 1;
    const bool sample = sample_every > 0 && number % sample_every == 0;
    if (sample && c->hist_rows >= c->hist_cap) {
      if (flush_history(c)) return -1;
      set_ctl(c, number, 0);
    }
    const bool last = k == nsteps - 1;
    if (last) integrals_any(r, acc_field(r, offsetof(BudgetAcc, before)), 0);
    switch (c->storage) {
      case 0: refined_coarse_step<0>(r, sample, &ev[8 * k]); break;
      case 1: refined_coarse_step<1>(r, sample, &ev[8 * k]); break;
      default: refined_coarse_step<2>(r, sample, &ev[8 * k]); break;
    }
    if (sample) c->hist_rows++;
    r->steps++;
    if (last) integrals_any(r, acc_field(r, offsetof(BudgetAcc, after)), number);
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double tot = 0, cs = 0, fs = 0;
  for (int k = 0; k < nsteps; ++k) {
    float a = 0, b = 0, d0 = 0, d1 = 0;
    hipEvent_t *e = &ev[8 * k];
    HIPCHK(c, hipEventElapsedTime(&a, e[0], e[1]));
    HIPCHK(c, hipEventElapsedTime(&b, e[2], e[3]));
    HIPCHK(c, hipEventElapsedTime(&d0, e[4], e[5]));
    HIPCHK(c, hipEventElapsedTime(&d1, e[6], e[7]));
    tot += a; cs += b; fs += d0 + d1;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  out[0] = tot; out[1] = cs; out[2] = fs; out[3] = nsteps;
  return 0;
}

// Restartable state of a two-level case: the coarse-step count, then the coarse and fine
// lattice states (ak_save_state layout).  The transfer buffers are rebuilt every step.
size_t ak_refined_state_bytes(void *handle) {
  Refined *r = R(handle);
  return r ? sizeof(long long) + ak_state_bytes(r->c) + ak_state_bytes(r->f) + sizeof(BudgetAcc) + sizeof(int) : 0;
}

int ak_refined_save_state(void *handle, void *host, size_t bytes) {
  Refined *r = R(handle);
  if (!r) return fail(nullptr, "null refined handle");
  if (!host || bytes < ak_refined_state_bytes(handle)) return fail(nullptr, "refined state buffer too short");
  char *o = (char *)host;
  std::memcpy(o, &r->steps, sizeof(long long));
  o += sizeof(long long);
  if (ak_save_state(r->c, o, ak_state_bytes(r->c))) return -1;
  o += ak_state_bytes(r->c);
  if (ak_save_state(r->f, o, ak_state_bytes(r->f))) return -1;
  o += ak_state_bytes(r->f);
  // the budget accumulators (the stream was synchronised by the state copies)
  HIPCHK(r->c, hipMemcpy(o, r->d_acc, sizeof(BudgetAcc), hipMemcpyDeviceToHost));
  const int complete = r->budget_complete ? 1 : 0;
  std::memcpy(o + sizeof(BudgetAcc), &complete, sizeof(int));
  return 0;
}

int ak_refined_load_state(void *handle, const void *host, size_t bytes) {
  Refined *r = R(handle);
  if (!r) return fail(nullptr, "null refined handle");
  if (!host || bytes < sizeof(long long)) return fail(nullptr, "refined state too short");
  const char *i = (const char *)host;
  long long steps = 0;
  std::memcpy(&steps, i, sizeof steps);
  i += sizeof(long long);
  const size_t nc = state_blob_size(r->c, i, bytes - sizeof(long long));
  const size_t nf = nc ? state_blob_size(r->f, i + nc, bytes - sizeof(long long) - std::min(nc, bytes - sizeof(long long))) : 0;
  const size_t tail = sizeof(BudgetAcc) + sizeof(int);
  if (!nc || !nf || bytes != sizeof(long long) + nc + nf + tail) return fail(nullptr, "refined state of another size or layout");
  if (ak_load_state(r->c, i, nc) || ak_load_state(r->f, i + nc, nf)) return -1;
  HIPCHK(r->c, hipMemcpy(r->d_acc, i + nc + nf, sizeof(BudgetAcc), hipMemcpyHostToDevice));
  int complete = 0;
  std::memcpy(&complete, i + nc + nf + sizeof(BudgetAcc), sizeof(int));
  r->budget_complete = complete != 0;
  r->steps = steps;
  return 0;
}

// Conservation budget of the case (refined_solver.py:291-361), FP64, common units:
// out[0..2] wall impulse of the last step, [3] its wall moment impulse, [4] its wall mass
// exchange, [5..8] mass and momentum in through the outer faces in the last step, [9]
// cumulative wall mass exchange, [10..12] cumulative wall impulse, [13..16] cumulative outer
// in-flux, [17..20] conserved integrals (mass, momentum) of the initial state, [21..24] and
// [25..28] before and after the last step of the last run call, [29] coarse steps accumulated,
// [30] the step [21..28] bracket (0: none), [31] 1 if every step since the initial state was
// accumulated (a batch does not keep the budget).  Synchronises.
int ak_refined_budget(void *handle, double *out) {
  Refined *r = R(handle);
  if (!r) return fail(nullptr, "null refined handle");
  if (!out) return fail(nullptr, "null output");
  BudgetAcc a{};
  HIPCHK(r->c, hipSetDevice(r->c->dev));
  HIPCHK(r->c, hipStreamSynchronize(r->c->stream));
  HIPCHK(r->c, hipMemcpy(&a, r->d_acc, sizeof a, hipMemcpyDeviceToHost));
  for (int k = 0; k < 5; ++k) out[k] = a.wall[k];
  for (int k = 0; k < 4; ++k) out[5 + k] = a.ext[k];
  out[9] = a.cum_wall_mass;
  for (int k = 0; k < 3; ++k) out[10 + k] = a.cum_wall_impulse[k];
  for (int k = 0; k < 4; ++k) out[13 + k] = a.cum_ext[k];
  for (int k = 0; k < 4; ++k) {
    out[17 + k] = a.initial[k];
    out[21 + k] = a.before[k];
    out[25 + k] = a.after[k];
  }
  out[29] = (double)a.steps;
  out[30] = (double)a.budget_step;
  out[31] = r->budget_complete ? 1.0 : 0.0;
  return 0;
}

// Transfer error flags raised so far (bit 1 density, 2 negative population, 4 Mach, 8 non-finite); clears them.
int ak_refined_transfer_errors(void *handle) {
  Refined *r = R(handle);
  if (!r) return fail(nullptr, "null refined handle");
  int e = 0;
  HIPCHK(r->c, hipStreamSynchronize(r->c->stream));
  HIPCHK(r->c, hipMemcpy(&e, r->d_err, sizeof(int), hipMemcpyDeviceToHost));
  (void)hipMemsetAsync(r->d_err, 0, sizeof(int), r->c->stream);
  return e;
}

// Pre-collision populations of one level (0 coarse, 1 fine padded), exchange layout.
int ak_refined_get_populations(void *handle, int level, double *host_out) {
  Refined *r = R(handle);
  if (!r) return fail(nullptr, "null refined handle");
  Lattice *s = level ? r->f : r->c;
  if (check_ready(s)) return -1;
  const size_t n = (size_t)s->nx * s->ny * s->nz;
  double *d = nullptr;
  HIPCHK(s, hipMalloc((void **)&d, n * 27 * sizeof(double)));
  const unsigned nb = blocks_for((long long)n, 256);
  const bool pe = s->sweeps % 2 == 0;
  switch (s->storage) {
    case 0: if (pe) hipLaunchKernelGGL((k_export_in<0, true>), dim3(nb), dim3(256), 0, s->stream, (const double *)s->pop, s->L, d, s->inv_scale);
            else hipLaunchKernelGGL((k_export_in<0, false>), dim3(nb), dim3(256), 0, s->stream, (const double *)s->pop, s->L, d, s->inv_scale); break;
    case 1: if (pe) hipLaunchKernelGGL((k_export_in<1, true>), dim3(nb), dim3(256), 0, s->stream, (const float *)s->pop, s->L, d, s->inv_scale);
            else hipLaunchKernelGGL((k_export_in<1, false>), dim3(nb), dim3(256), 0, s->stream, (const float *)s->pop, s->L, d, s->inv_scale); break;
    default: if (pe) hipLaunchKernelGGL((k_export_in<2, true>), dim3(nb), dim3(256), 0, s->stream, (const _Float16 *)s->pop, s->L, d, s->inv_scale);
             else hipLaunchKernelGGL((k_export_in<2, false>), dim3(nb), dim3(256), 0, s->stream, (const _Float16 *)s->pop, s->L, d, s->inv_scale); break;
  }
  hipError_t e = hipStreamSynchronize(s->stream);
  if (e == hipSuccess) e = hipMemcpy(host_out, d, n * 27 * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(s, "refined export: %s", hipGetErrorString(e));
  return 0;
}

}  // extern "C"
