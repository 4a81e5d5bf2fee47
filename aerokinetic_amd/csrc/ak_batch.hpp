// ak_batch.hpp -- device side of a batch of two-level cases.
//
// Every kernel of the coarse step runs for all members of a batch in ONE launch over the
// members' concatenated work lists: block b belongs to the member whose first block is
// the last one <= b (bisection over a per-phase table), and that member's arguments come
// from its CaseDev record.  The per-item work is the same device function the single-case
// kernels call (the *_body functions), so a member computes bit-identically alone, in a
// batch of one, or in any mix.
#pragma once

#include "ak_kernels.hpp"
#include "ak_refined.hpp"

namespace ak {

// Everything the step kernels of one two-level case read.
struct CaseDev {
  void *Ac, *Af;  // coarse / fine populations
  Desc Lc, Lf;
  float c_scale, c_inv, f_scale, f_inv;
  BoxGeom B;
  Levels V;
  FreeState F;
  const FluxTable *T;
  double *before, *after, *regc, *regf, *px;  // transfer buffers, read as TR<S> through tb<S>()
  double *part;                               // wall load partials (FP64)
  const int32_t *lpx, *cols;  // prolongation x-pass items; fine interface-band columns
  int npx, ncols;
  int *err;
  double volc, volf, inv_volc, loads_scale;
  WallLinks W;  // fine-level walls (lever arm dx_f)
  int wall_blocks;
  LegacyLinks K;  // coarse-level legacy farfield
  double U;
  StepCtl *ctl;
  double *hist;
  int cpar;  // coarse sweep-count parity at the start of a step when the lead member's is even
  int fpar;  // fine sweep-count parity at the start of a step
};

// The step's kernels, each with a table of first blocks per member.
enum BatchPhase {
  BP_DONOR, BP_FLUXC, BP_LEGACY, BP_PX, BP_PYZ, BP_FLUXF, BP_WALLS, BP_REFLUX, BP_RESTRICT,
  BP_COUNT
};

// A transfer buffer of a member as the kernels of storage S read it.
template <int S>
__device__ __forceinline__ TR<S> *tb(double *p) { return (TR<S> *)p; }

// Member owning block b: the last m with first[m] <= b (members with no blocks are skipped).
__device__ __forceinline__ int batch_owner(const int *__restrict__ first, int nm, int b) {
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

#define AK_BATCH_MEMBER                                                          \
  const int m_ = batch_owner(first, nm, (int)blockIdx.x);                        \
  const CaseDev &c = C[m_];                                                      \
  const int lb_ = (int)blockIdx.x - first[m_];                                   \
  const long long t = (long long)lb_ * blockDim.x + threadIdx.x;                 \
  (void)lb_;
// the same for a block `blk` of one range of a fused launch, with that range's table
#define AK_BATCH_MEMBER_IN(tbl, blk)                                             \
  const int m_ = batch_owner(tbl, nm, blk);                                      \
  const CaseDev &c = C[m_];                                                      \
  const int lb_ = (blk) - (tbl)[m_];                                             \
  const long long t = (long long)lb_ * blockDim.x + threadIdx.x;                 \
  (void)lb_;
// AA form of a coarse / fine kernel: the sweep count (at step start + delta) is even
#define AK_PE_COARSE(delta) (((((c.cpar ^ lead) + (delta)) & 1)) == 0)
#define AK_PE_FINE(delta) ((((c.fpar + (delta)) & 1)) == 0)

template <int S>
__global__ __launch_bounds__(256) void kb_donor(const CaseDev *__restrict__ C, const int *__restrict__ first, int nm,
                                                int lead, int delta, int after) {
  using T = typename St<S>::T;
  AK_BATCH_MEMBER
  TR<S> *out = tb<S>(after ? c.after : c.before);
  if (AK_PE_COARSE(delta)) donor_moments_body<S, true>((const T *)c.Ac, c.Lc, c.B, c.F, c.c_inv, out, c.err, t);
  else donor_moments_body<S, false>((const T *)c.Ac, c.Lc, c.B, c.F, c.c_inv, out, c.err, t);
}

// x pass of both substeps (after the coarse step's donor moments at t + dt)
template <int S>
__global__ __launch_bounds__(256) void kb_prolong_x(const CaseDev *__restrict__ C, const int *__restrict__ first,
                                                    int nm) {
  AK_BATCH_MEMBER
  if (t < c.npx) prolong_x_body<TR<S>>(tb<S>(c.before), tb<S>(c.after), c.B, tb<S>(c.px), c.lpx[t]);
}

// y and z passes + injection: one 64-thread block per group of PZ_COLS band columns
// (substep `sub`: fine AA form at step start + sub, x pass of time fraction sub / 2)
template <int S>
__global__ __launch_bounds__(64) void kb_prolong_yz(const CaseDev *__restrict__ C, const int *__restrict__ first,
                                                    int nm, int sub) {
  using T = typename St<S>::T;
  AK_BATCH_MEMBER
  (void)t;
  const TR<S> *X = tb<S>(c.px) + sub * prolong_x_half(c.B);
  if (AK_PE_FINE(sub)) prolong_yz_body<S, true>(X, (T *)c.Af, c.Lf, c.B, c.V, c.f_scale, c.err, c.cols, c.ncols, lb_);
  else prolong_yz_body<S, false>(X, (T *)c.Af, c.Lf, c.B, c.V, c.f_scale, c.err, c.cols, c.ncols, lb_);
}

template <int S>
__global__ __launch_bounds__(256) void kb_reflux(const CaseDev *__restrict__ C, const int *__restrict__ first, int nm,
                                                 int lead, int delta) {
  using T = typename St<S>::T;
  AK_BATCH_MEMBER
  if (AK_PE_COARSE(delta))
    reflux_body<S, true>((T *)c.Ac, c.Lc, c.B, tb<S>(c.regc), tb<S>(c.regf), TR<S>(c.inv_volc), c.c_scale, c.c_inv,
                         c.err, t);
  else
    reflux_body<S, false>((T *)c.Ac, c.Lc, c.B, tb<S>(c.regc), tb<S>(c.regf), TR<S>(c.inv_volc), c.c_scale, c.c_inv,
                          c.err, t);
}


// Fine moments + restriction, one block per restriction tile of a member (fine parity at
// delta 2, coarse at delta 1 from the step start), then (blocks from nb_rt on) the next
// step's donor moments of the donor cells the restriction does not write (as
// k_restrict_tiles of a single case).
template <int S>
__global__ __launch_bounds__(256) void kb_restrict(const CaseDev *__restrict__ C, const int *__restrict__ f_rt,
                                                   const int *__restrict__ f_don, int nm, int lead, int nb_rt) {
  using T = typename St<S>::T;
  __shared__ double sm_[RESTRICT_LDS];  // one array for the four parity instances
  TR<S> *sm = (TR<S> *)sm_;
  const int b = (int)blockIdx.x;
  if (b < nb_rt) {
    AK_BATCH_MEMBER_IN(f_rt, b)
    (void)t;
    const bool pf = AK_PE_FINE(2), pc = AK_PE_COARSE(1);
    if (pf && pc) restrict_tile_body<S, true, true>((const T *)c.Af, c.Lf, c.f_inv, (T *)c.Ac, c.Lc, c.B, c.V, c.F, c.c_scale, c.err, lb_, sm, tb<S>(c.before), c.c_inv);
    else if (pf) restrict_tile_body<S, true, false>((const T *)c.Af, c.Lf, c.f_inv, (T *)c.Ac, c.Lc, c.B, c.V, c.F, c.c_scale, c.err, lb_, sm, tb<S>(c.before), c.c_inv);
    else if (pc) restrict_tile_body<S, false, true>((const T *)c.Af, c.Lf, c.f_inv, (T *)c.Ac, c.Lc, c.B, c.V, c.F, c.c_scale, c.err, lb_, sm, tb<S>(c.before), c.c_inv);
    else restrict_tile_body<S, false, false>((const T *)c.Af, c.Lf, c.f_inv, (T *)c.Ac, c.Lc, c.B, c.V, c.F, c.c_scale, c.err, lb_, sm, tb<S>(c.before), c.c_inv);
  } else {
    AK_BATCH_MEMBER_IN(f_don, b - nb_rt)
    if (AK_PE_COARSE(1)) donor_moments_body<S, true, true>((const T *)c.Ac, c.Lc, c.B, c.F, c.c_inv, tb<S>(c.before), c.err, t);
    else donor_moments_body<S, false, true>((const T *)c.Ac, c.Lc, c.B, c.F, c.c_inv, tb<S>(c.before), c.err, t);
  }
}

// One 64-thread block per member: the loads of a sampled step, or the step tick.
__global__ __launch_bounds__(64) void kb_loads(const CaseDev *__restrict__ C, int sample) {
  __shared__ double acc[LOADS_LDS];
  const CaseDev &c = C[blockIdx.x];
  if (sample)
    loads_final_body(c.part, c.W.n ? 2 * c.wall_blocks : 0, c.W.nelem, c.loads_scale, c.loads_scale, c.volf, c.ctl,
                     c.hist, nullptr, acc);
  else if (threadIdx.x == 0)
    c.ctl->step += 1;
}

// Fused phases (as k_after_coarse / k_after_fine of a single case): the block ranges of
// several phases, each over all members, in one launch; the member of a block comes from
// that phase's table.

template <int S>
__global__ __launch_bounds__(256) void kb_after_coarse(const CaseDev *__restrict__ C, const int *__restrict__ f_flux,
                                                       const int *__restrict__ f_leg, const int *__restrict__ f_don,
                                                       int nm, int lead, int nb_flux, int nb_leg) {
  using T = typename St<S>::T;
  const int delta = 1;
  const int b = (int)blockIdx.x;
  if (b < nb_flux) {
    AK_BATCH_MEMBER_IN(f_flux, b)
    if (AK_PE_COARSE(delta)) flux_coarse_body<S, true>((const T *)c.Ac, c.Lc, c.B, c.T, TR<S>(c.volc), c.c_inv, tb<S>(c.regc), t);
    else flux_coarse_body<S, false>((const T *)c.Ac, c.Lc, c.B, c.T, TR<S>(c.volc), c.c_inv, tb<S>(c.regc), t);
  } else if (b < nb_flux + nb_leg) {
    AK_BATCH_MEMBER_IN(f_leg, b - nb_flux)
    if (AK_PE_COARSE(delta)) legacy_body<S, true>((T *)c.Ac, c.Lc, c.K, c.U, c.c_scale, c.c_inv, t);
    else legacy_body<S, false>((T *)c.Ac, c.Lc, c.K, c.U, c.c_scale, c.c_inv, t);
  } else {
    AK_BATCH_MEMBER_IN(f_don, b - nb_flux - nb_leg)
    if (AK_PE_COARSE(delta)) donor_moments_body<S, true>((const T *)c.Ac, c.Lc, c.B, c.F, c.c_inv, tb<S>(c.after), c.err, t);
    else donor_moments_body<S, false>((const T *)c.Ac, c.Lc, c.B, c.F, c.c_inv, tb<S>(c.after), c.err, t);
  }
}

template <int S>
__global__ __launch_bounds__(256) void kb_after_fine(const CaseDev *__restrict__ C, const int *__restrict__ f_flux,
                                                     const int *__restrict__ f_walls, int nm, int delta, int sub,
                                                     int nb_flux) {
  using T = typename St<S>::T;
  const int b = (int)blockIdx.x;
  if (b < nb_flux) {
    AK_BATCH_MEMBER_IN(f_flux, b)
    if (AK_PE_FINE(delta)) flux_fine_body<S, true>((const T *)c.Af, c.Lf, c.B, c.T, TR<S>(c.volf), c.f_inv, sub, tb<S>(c.regf), t);
    else flux_fine_body<S, false>((const T *)c.Af, c.Lf, c.B, c.T, TR<S>(c.volf), c.f_inv, sub, tb<S>(c.regf), t);
  } else {
    AK_BATCH_MEMBER_IN(f_walls, b - nb_flux)
    double *part = c.part + (size_t)sub * c.wall_blocks * c.W.nelem * 5;
    if (AK_PE_FINE(delta)) walls_body<S, true, true>((T *)c.Af, c.Lf, c.W, c.f_scale, c.f_inv, part, t, lb_);
    else walls_body<S, false, true>((T *)c.Af, c.Lf, c.W, c.f_scale, c.f_inv, part, t, lb_);
  }
}
#undef AK_BATCH_MEMBER_IN

#undef AK_BATCH_MEMBER
#undef AK_PE_COARSE
#undef AK_PE_FINE

}  // namespace ak
