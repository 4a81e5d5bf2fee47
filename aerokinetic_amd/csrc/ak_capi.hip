// ak_capi.hip -- C ABI of the gfx950 D3Q27 cumulant step (declared in include/aerokinetic.h).
//
// Host-side orchestration only: buffer sizing, boundary item tables, launch order.
// Per step: boundary kernels (farfield, walls; the walls share a launch with the
// edge phase that does not depend on them) write the incoming populations of
// boundary links where the next AA sweep reads them; on sampled steps the walls
// also emit load partials and the sweep accumulates the diagnostics (max |u|,
// max |rho-1|, min rho, non-finite flag) into slots, and a one-block kernel folds
// both into a history row.  No host synchronisation in the step path.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/aerokinetic.h"
#include "ak_kernels.hpp"
#include "ak_exterior.hpp"

using namespace ak;

namespace {

std::string g_error;

// Instantiated HIP graphs of one step, one slot per launch variant (AA parity,
// sampled or not, ...).  A slot is valid while `key` matches the generation of
// the state it captured (kernel arguments: buffers, tables, tile shape).
struct GraphCache {
  hipGraphExec_t exec[8] = {};
  unsigned long long key[8] = {};
  void clear() {
    for (auto &e : exec)
      if (e) { (void)hipGraphExecDestroy(e); e = nullptr; }
  }
};

struct Lattice {
  int dev = 0, nx = 0, ny = 0, nz = 0, storage = 0;
  double ws = 1.0, wb = 1.0;
  float scale = 1.0f, inv_scale = 1.0f;
  Desc L{};
  size_t pop_elems = 0, pop_bytes = 0, elem_size = 8;
  void *pop = nullptr;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  bool pre_collision = false;  // levels of a two-level case: the stored state is pre-collision
  uint8_t *d_solid = nullptr;
  // walls
  int nlinks = 0, nelem = 1;
  int32_t *d_wxy = nullptr;
  int8_t *d_wdir = nullptr, *d_welem = nullptr;
  double *d_wq = nullptr;
  double rx_off = 0, ry_off = 0, chord = 1, U = 0;
  double *d_partials = nullptr;
  int wall_blocks = 0;
  bool walls_clear = false;  // every wall link at least 3 cells from the x and y edges
  // farfield
  int ff_kind = AK_FARFIELD_PERIODIC;
  FFParams ffp{};
  int nff = 0;
  int32_t *d_ffg = nullptr, *d_ffz = nullptr, *d_ffmask = nullptr;
  int8_t *d_ffobj = nullptr;
  double *d_inc = nullptr, *d_der = nullptr, *d_rows = nullptr;
  // perimeter moments of the characteristic farfield (compute type): written by every sweep
  // and part of the restartable state; computed from the populations when the state was set
  // from outside (pm_valid false).  Used while no perimeter cell is solid.
  uint8_t *d_pm = nullptr;
  bool perim_fluid = true, pm_valid = false;
  // outgoing populations of the x-face cells (storage type), kept by the sweep beside the
  // perimeter moments; xo_valid false: recomputed from the lattice at the next run call
  uint8_t *d_xo = nullptr;
  bool xo_valid = false;
  // populations entering the x-face cells, written by the farfield for the next sweep only
  uint8_t *d_xi = nullptr;
  long long *d_clock = nullptr;
  int nleg = 0;
  int32_t *d_legx = nullptr, *d_legz = nullptr;
  int8_t *d_legdir = nullptr, *d_legkind = nullptr;
  double *d_pdf = nullptr, *d_pdfnd = nullptr;
  // exterior farfield: band tables, band / ring states, source partials and accumulators
  int ext_R2 = 0, ext_B2 = 0, ext_ring_blocks = 0;
  int32_t *d_ext_bxy = nullptr, *d_ext_map = nullptr;
  double *d_ext_M = nullptr, *d_ext_Y1 = nullptr, *d_ext_Y2 = nullptr, *d_ext_pred = nullptr;
  double *d_ext_rows = nullptr, *d_ext_part = nullptr;
  ExtAcc *d_ext_acc = nullptr;
  int *d_ext_err = nullptr;
  // diagnostics / history
  DiagAcc *d_diag = nullptr;
  double *d_hist = nullptr;
  int hist_cap = 0, hist_rows = 0, hist_width = 0;
  std::vector<double> host_hist;
  long long sweeps = 0;  // sweeps since the initial state: parity of the AA form
  StepCtl *d_ctl = nullptr;
  // stationarity residual: previous (rho, j / U) of every cell, block partials, result
  double *d_prev = nullptr, *d_rpart = nullptr, *d_rout = nullptr;
  bool prev_valid = false;
  // step graphs: replayed by ak_run while no setter has changed a kernel argument
  bool use_graphs = true;
  unsigned long long gen = 1;
  GraphCache graphs;
  // sweep tiles of the even (in-place) and odd (gathering) AA forms; geom_small: the periodic
  // form's bound of 256-thread blocks
  SweepGeom geom{}, geom_odd{}, geom_small{};
  // cap on the sweep's resident waves per SIMD of each AA form (0: none), enforced through the
  // LDS each sweep block reserves (the kernel uses none): fewer concurrent 27-row streams can
  // move more bytes per second than a full CU; lds_per_cu: the device's LDS per CU
  int waves_even = 0, waves_odd = 0;
  int lds_per_cu = 0;
  std::string err;
  std::vector<void *> owned;
};

int fail(Lattice *s, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (s) s->err = buf;
  g_error = buf;
  return -1;
}

#define HIPCHK(s, call)                                                                    \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) return fail((s), "%s: %s", #call, hipGetErrorString(e_));        \
  } while (0)

template <typename T>
int dalloc(Lattice *s, T **p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) return fail(s, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
  s->owned.push_back(*p);
  return 0;
}

void dfree(Lattice *s, void *p) {
  if (!p) return;
  auto it = std::find(s->owned.begin(), s->owned.end(), p);
  if (it != s->owned.end()) s->owned.erase(it);
  (void)hipFree(p);
}

template <typename T>
int upload(Lattice *s, T **dst, const std::vector<T> &v) {
  if (*dst) { dfree(s, *dst); *dst = nullptr; }
  if (dalloc(s, dst, v.size())) return -1;
  if (!v.empty()) HIPCHK(s, hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

inline unsigned blocks_for(long long n, int t) { return (unsigned)((n + t - 1) / t); }

// Sweep geometry: bx x by tiles, or with bx == 0 the row-linear mode with `by` threads.
SweepGeom make_geom(int nx, int ny, int nz, int bx, int by) {
  SweepGeom G{};
  if (bx == 0) {
    G.lin = 1;
    G.bx = by;
    G.by = 1;
    G.gx = (int)(((long long)nx * ny + by - 1) / by);
    G.gy = 1;
  } else {
    G.bx = bx;
    G.by = by;
    G.gx = (nx + bx - 1) / bx;
    G.gy = (ny + by - 1) / by;
  }
  G.gz = nz;
  G.nblocks = G.gx * G.gy * G.gz;
  magic_of((unsigned)G.gx, G.m_gx, G.s_gx);
  magic_of((unsigned)(G.gx * G.gy), G.m_gxy, G.s_gxy);
  magic_of((unsigned)G.bx, G.m_bx, G.s_bx);
  magic_of((unsigned)nx, G.m_nx, G.s_nx);
  return G;
}
// default: 64 x 4 tiles when rows are whole waves, row-linear blocks of 256 otherwise
SweepGeom default_geom(int nx, int ny, int nz) {
  return nx % 64 == 0 ? make_geom(nx, ny, nz, 64, 4) : make_geom(nx, ny, nz, 0, 256);
}
void set_geom(Lattice *s, int bx, int by) { s->geom = s->geom_odd = make_geom(s->nx, s->ny, s->nz, bx, by); }

// ---------------------------------------------------------------- launches

// Dynamic LDS per sweep block that caps the resident waves per SIMD at `waves` (0: no cap):
// a CU holds 4 SIMDs, so waves * 4 / (waves per block) blocks share its LDS.
size_t sweep_lds(const Lattice *s, int waves, int threads) {
  if (waves <= 0 || s->lds_per_cu <= 0) return 0;
  const int per_block = (threads + 63) / 64;
  const int blocks = waves * 4 / per_block;
  if (blocks < 1) return 0;
  return (size_t)(s->lds_per_cu / blocks) & ~(size_t)255;
}

// The perimeter moments the sweep keeps and the farfield reads, or null (no characteristic
// farfield, or a solid cell on the perimeter: the farfield then reads the lattice).
template <int S>
typename St<S>::C *perimeter_moments(Lattice *s) {
  const bool on = s->ff_kind == AK_FARFIELD_CHARACTERISTIC && s->nff && s->d_pm && s->perim_fluid;
  return on ? (typename St<S>::C *)s->d_pm : nullptr;
}
// The x-face copies of the outgoing and of the entering populations, used where the perimeter
// moments are.  Measured on the 1536x1024x64 lattice they cut the boundary phase from 0.100 /
// 0.115 / 0.118 to 0.082 / 0.087 / 0.085 ms (FP64 / FP32 / FP16) and the sweep by 2.1 / 2.8 / 2.4 %
// (the farfield no longer leaves partly written lines in the rows the sweep streams).
template <int S>
typename St<S>::T *xface_copy(Lattice *s) {
  return perimeter_moments<S>(s) && s->d_xo ? (typename St<S>::T *)s->d_xo : nullptr;
}
template <int S>
typename St<S>::T *xface_in(Lattice *s) {
  return perimeter_moments<S>(s) && s->d_xi ? (typename St<S>::T *)s->d_xi : nullptr;
}

// Perimeter moments of the current populations, when the state was set from outside since
// the last sweep (a sweep writes them for the state it leaves).
template <int S>
void fill_perimeter(Lattice *s) {
  using T = typename St<S>::T;
  typename St<S>::C *pm = perimeter_moments<S>(s);
  if (!pm) return;
  if (!s->pm_valid) {
    s->pm_valid = true;
    const dim3 g(blocks_for(2LL * (s->ny + s->nx) * s->nz, 256)), b(256);
    if (s->sweeps % 2 == 0)
      hipLaunchKernelGGL((k_perimeter_fill<S, true>), g, b, 0, s->stream, (const T *)s->pop, s->L, pm, s->inv_scale);
    else
      hipLaunchKernelGGL((k_perimeter_fill<S, false>), g, b, 0, s->stream, (const T *)s->pop, s->L, pm, s->inv_scale);
  }
  T *xo = xface_copy<S>(s);
  if (xo && !s->xo_valid) {
    s->xo_valid = true;
    const dim3 g(blocks_for(2LL * s->ny * s->nz, 256)), b(256);
    if (s->sweeps % 2 == 0)
      hipLaunchKernelGGL((k_xface_fill<S, true>), g, b, 0, s->stream, (const T *)s->pop, s->L, xo);
    else
      hipLaunchKernelGGL((k_xface_fill<S, false>), g, b, 0, s->stream, (const T *)s->pop, s->L, xo);
  }
}

// With e0/e1 the launch records the kernel's own start and end on those events
// (hipExtLaunchKernel), so the measured time is the sweep alone.
template <int S>
void launch_sweep(Lattice *s, bool odd, bool diag, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  using T = typename St<S>::T;
  using C = typename St<S>::C;
  T *A = (T *)s->pop;
  const bool pxy = s->ff_kind == AK_FARFIELD_PERIODIC;  // AK_FARFIELD_NONE keeps ghost indexing
  // each AA form has its own tile; the periodic form is bounded to 256-thread blocks, and
  // larger tiles fall back to 64x4 there
  const SweepGeom &GF = odd ? s->geom_odd : s->geom;
  const SweepGeom &G = pxy && GF.bx * GF.by > 256 ? s->geom_small : GF;
  const dim3 grid(G.nblocks), block(G.bx * G.by);
  const size_t lds = sweep_lds(s, odd ? s->waves_odd : s->waves_even, G.bx * G.by);
  const C ws = (C)s->ws, wb = (C)s->wb;
  C *pm = perimeter_moments<S>(s);
  T *xo = xface_copy<S>(s);
  const T *xi = xface_in<S>(s);
#define AK_SWEEP(O, P, D)                                                                                           \
  do {                                                                                                              \
    if (e0)                                                                                                         \
      hipExtLaunchKernelGGL((k_sweep<S, O, P, D>), grid, block, lds, s->stream, e0, e1, 0, A, s->d_solid, s->L, G, ws, \
                            wb, s->scale, s->inv_scale, s->d_diag, pm, xo, xi);                                     \
    else                                                                                                            \
      hipLaunchKernelGGL((k_sweep<S, O, P, D>), grid, block, lds, s->stream, A, s->d_solid, s->L, G, ws, wb, s->scale, \
                         s->inv_scale, s->d_diag, pm, xo, xi);                                                      \
  } while (0)
  if (odd) {
    if (pxy) { if (diag) AK_SWEEP(true, true, true); else AK_SWEEP(true, true, false); }
    else { if (diag) AK_SWEEP(true, false, true); else AK_SWEEP(true, false, false); }
  } else {
    if (pxy) { if (diag) AK_SWEEP(false, true, true); else AK_SWEEP(false, true, false); }
    else { if (diag) AK_SWEEP(false, false, true); else AK_SWEEP(false, false, false); }
  }
#undef AK_SWEEP
}

template <int S, bool PE>
void launch_boundary(Lattice *s, bool loads) {
  using T = typename St<S>::T;
  T *A = (T *)s->pop;
  const bool charff = s->ff_kind == AK_FARFIELD_CHARACTERISTIC && s->nff;
  const bool leg = (s->ff_kind == AK_FARFIELD_LEGACY || s->ff_kind == AK_FARFIELD_EXTERIOR) && s->nleg;
  const FFItems I{s->d_ffg, s->d_ffz, s->d_ffobj, s->d_ffmask, s->d_inc, s->d_der, s->d_clock, s->d_rows, s->nff,
                  perimeter_moments<S>(s), xface_copy<S>(s), xface_in<S>(s)};
  if (s->nlinks && s->walls_clear && (charff || leg)) {
    // walls clear of the edges: the walls and the edge phase (farfield items or legacy
    // links) in one launch
    WallLinks W{s->d_wxy, s->d_wdir, s->d_wq, s->d_welem, s->nlinks, s->nelem, s->rx_off, s->ry_off, 1.0};
    LegacyLinks K{s->d_legx, s->d_legz, s->d_legdir, s->d_legkind, s->d_pdf, s->d_pdfnd, s->nleg};
    const long long nb_edge = charff ? blocks_for(s->nff, 256) : blocks_for(s->nleg, 256);
    const dim3 g((unsigned)(s->wall_blocks + nb_edge)), b(256);
#define AK_WE(LD, E) hipLaunchKernelGGL((k_walls_edge<S, PE, LD, E>), g, b, 0, s->stream, A, s->L, W, s->scale, \
                                        s->inv_scale, s->d_partials, s->wall_blocks, K, s->U, I, s->ffp)
    if (charff) { if (loads) AK_WE(true, 1); else AK_WE(false, 1); }
    else { if (loads) AK_WE(true, 0); else AK_WE(false, 0); }
#undef AK_WE
    return;
  }
  if (charff) {
    hipLaunchKernelGGL((k_farfield<S, PE>), dim3(blocks_for(s->nff, 128)), dim3(128), 0, s->stream, A, s->L, I,
                       s->ffp, s->scale, s->inv_scale);
  } else if ((s->ff_kind == AK_FARFIELD_LEGACY || s->ff_kind == AK_FARFIELD_EXTERIOR) && s->nleg) {
    // the exterior reservoir is the same velocity bounce-back on every side
    LegacyLinks K{s->d_legx, s->d_legz, s->d_legdir, s->d_legkind, s->d_pdf, s->d_pdfnd, s->nleg};
    hipLaunchKernelGGL((k_legacy<S, PE>), dim3(blocks_for(s->nleg, 128)), dim3(128), 0, s->stream, A, s->L, K,
                       s->U, s->scale, s->inv_scale);
  }
  if (s->nlinks) {
    WallLinks W{s->d_wxy, s->d_wdir, s->d_wq, s->d_welem, s->nlinks, s->nelem, s->rx_off, s->ry_off, 1.0};
    if (loads)
      hipLaunchKernelGGL((k_walls<S, PE, true>), dim3(s->wall_blocks), dim3(256), 0, s->stream, A, s->L, W, s->scale,
                         s->inv_scale, s->d_partials);
    else
      hipLaunchKernelGGL((k_walls<S, PE, false>), dim3(s->wall_blocks), dim3(256), 0, s->stream, A, s->L, W, s->scale,
                         s->inv_scale, s->d_partials);
  }
}

// end of a sampled step: diagnostics and loads into the history row, slots cleared, cursor
// advanced (one block)
void launch_step_final(Lattice *s) {
  const double qa = 0.5 * s->U * s->U * s->chord * s->nz;
  hipLaunchKernelGGL(k_step_final, dim3(1), dim3(128), 0, s->stream, s->d_partials, s->nlinks ? s->wall_blocks : 0,
                     s->nelem, 1.0 / qa, 1.0 / (qa * s->chord), 1.0, s->d_ctl, s->d_hist, s->d_diag);
}

ExtDev ext_dev(Lattice *s) {
  ExtDev E{};
  E.band_xy = s->d_ext_bxy; E.map = s->d_ext_map;
  E.R2 = s->ext_R2; E.B2 = s->ext_B2; E.X = s->nx; E.Y = s->ny; E.Z = s->nz;
  E.U = s->U; E.omega = s->ws;
  E.M = s->d_ext_M; E.Y1 = s->d_ext_Y1; E.Y2 = s->d_ext_Y2; E.PRED = s->d_ext_pred;
  E.rows = s->d_ext_rows; E.part = s->d_ext_part; E.acc = s->d_ext_acc; E.err = s->d_ext_err;
  return E;
}

// The three predictor stages of the ring; t from the step cursor (ctl) or t0.
void launch_ext_predict(Lattice *s, const StepCtl *ctl, double t0) {
  const ExtDev E = ext_dev(s);
  const dim3 g(s->ext_ring_blocks), b(128);
  hipLaunchKernelGGL(k_ext_stage<1>, g, b, 0, s->stream, E, ctl, t0);
  hipLaunchKernelGGL(k_ext_stage<2>, g, b, 0, s->stream, E, ctl, t0);
  hipLaunchKernelGGL(k_ext_stage<3>, g, b, 0, s->stream, E, ctl, t0);
}

// Exterior phases after the sweep of form `odd` (the ring was skipped by it); they end the step.
template <int S>
void launch_exterior(Lattice *s, bool odd, bool sample) {
  using T = typename St<S>::T;
  using C = typename St<S>::C;
  T *A = (T *)s->pop;
  const ExtDev E = ext_dev(s);
  const long long nb = (long long)(s->ext_B2 - s->ext_R2) * s->nz;
  // after an odd sweep the state is in streamed form, after an even one in post-even form
  if (odd)
    hipLaunchKernelGGL((k_ext_band<S, false>), dim3(blocks_for(nb, 128)), dim3(128), 0, s->stream, (const T *)A, s->L, E, s->ext_R2, s->ext_B2, s->inv_scale);
  else
    hipLaunchKernelGGL((k_ext_band<S, true>), dim3(blocks_for(nb, 128)), dim3(128), 0, s->stream, (const T *)A, s->L, E, s->ext_R2, s->ext_B2, s->inv_scale);
  const dim3 g(s->ext_ring_blocks), b(128);
  const C ws = (C)s->ws, wb = (C)s->wb;
#define AK_RING(O, D) hipLaunchKernelGGL((k_ext_ring<S, O, D>), g, b, 0, s->stream, A, s->L, E, ws, wb, s->scale, s->inv_scale, s->d_diag)
  if (odd) { if (sample) AK_RING(true, true); else AK_RING(true, false); }
  else { if (sample) AK_RING(false, true); else AK_RING(false, false); }
#undef AK_RING
  launch_ext_predict(s, s->d_ctl, 0.0);
  // accumulators and the end of the step
  const double qa = 0.5 * s->U * s->U * s->chord * s->nz;
  hipLaunchKernelGGL(k_ext_end, dim3(1), dim3(192), 0, s->stream, (const double *)s->d_partials,
                     s->nlinks ? s->wall_blocks : 0, s->nelem, E, s->ext_ring_blocks, s->d_ctl, s->d_hist,
                     sample ? 1 : 0, 1.0 / qa, 1.0 / (qa * s->chord), 1.0, s->d_diag);
}

// One uniform step; the step number and history row come from s->d_ctl.  With `ev`, four
// events: around the boundary phase (ev[0], ev[1]) and the sweep kernel itself (ev[2], ev[3]).
template <int S>
void one_step(Lattice *s, bool sample, hipEvent_t *ev = nullptr) {
  const bool pe = (s->sweeps % 2) == 0;
  const bool ext = s->ff_kind == AK_FARFIELD_EXTERIOR;
  if (ev) (void)hipEventRecord(ev[0], s->stream);
  // the exterior farfield accumulates the wall mass exchange on every step
  if (pe) launch_boundary<S, true>(s, sample || ext);
  else launch_boundary<S, false>(s, sample || ext);
  if (ev) (void)hipEventRecord(ev[1], s->stream);
  launch_sweep<S>(s, pe, sample, ev ? ev[2] : nullptr, ev ? ev[3] : nullptr);
  if (ext) launch_exterior<S>(s, pe, sample);  // ends the step
  else if (sample) launch_step_final(s);
  else hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, s->stream, s->d_ctl);
  s->sweeps++;
}

void set_ctl_on(Lattice *s, hipStream_t st, long long step, int row) {
  hipLaunchKernelGGL(k_set_ctl, dim3(1), dim3(64), 0, st, s->d_ctl, step, row, s->hist_width);
}
void set_ctl(Lattice *s, long long step, int row) { set_ctl_on(s, s->stream, step, row); }

// Enqueue one step on `st` through the graph in slot `slot` of `gc`: captured from
// `launch` (which also advances the host-side counters) on first use or after a
// generation change, replayed otherwise (then `advance` updates the counters).
template <class Launch, class Advance>
int graph_step(Lattice *s, hipStream_t st, GraphCache &gc, int slot, unsigned long long key, Launch &&launch,
               Advance &&advance) {
  if (gc.exec[slot] && gc.key[slot] != key) {
    (void)hipGraphExecDestroy(gc.exec[slot]);
    gc.exec[slot] = nullptr;
  }
  if (!gc.exec[slot]) {
    hipGraph_t g = nullptr;
    HIPCHK(s, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    launch();
    hipError_t e = hipStreamEndCapture(st, &g);
    if (e != hipSuccess) return fail(s, "step graph capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiateWithFlags(&gc.exec[slot], g, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      gc.exec[slot] = nullptr;
      return fail(s, "step graph instantiation: %s", hipGetErrorString(e));
    }
    gc.key[slot] = key;
  } else {
    advance();
  }
  HIPCHK(s, hipGraphLaunch(gc.exec[slot], st));
  return 0;
}

int flush_history(Lattice *s) {
  if (!s->hist_rows) return 0;
  HIPCHK(s, hipStreamSynchronize(s->stream));
  const size_t n = (size_t)s->hist_rows * s->hist_width;
  const size_t old = s->host_hist.size();
  s->host_hist.resize(old + n);
  HIPCHK(s, hipMemcpy(s->host_hist.data() + old, s->d_hist, n * sizeof(double), hipMemcpyDeviceToHost));
  s->hist_rows = 0;
  return 0;
}

template <int S, bool PE>
void launch_export(Lattice *s, double *d_out) {
  using T = typename St<S>::T;
  const long long n = (long long)s->nx * s->ny * s->nz;
  if (s->ff_kind == AK_FARFIELD_PERIODIC)
    hipLaunchKernelGGL((k_export<S, PE, true>), dim3(blocks_for(n, 256)), dim3(256), 0, s->stream, (const T *)s->pop, s->L, d_out, s->inv_scale);
  else
    hipLaunchKernelGGL((k_export<S, PE, false>), dim3(blocks_for(n, 256)), dim3(256), 0, s->stream, (const T *)s->pop, s->L, d_out, s->inv_scale);
}

template <int S, bool PE>
void launch_import(Lattice *s, const double *d_in) {
  using T = typename St<S>::T;
  const long long n = (long long)s->nx * s->ny * s->nz;
  if (s->ff_kind == AK_FARFIELD_PERIODIC)
    hipLaunchKernelGGL((k_import<S, PE, true>), dim3(blocks_for(n, 256)), dim3(256), 0, s->stream, (T *)s->pop, s->L, d_in, s->scale);
  else
    hipLaunchKernelGGL((k_import<S, PE, false>), dim3(blocks_for(n, 256)), dim3(256), 0, s->stream, (T *)s->pop, s->L, d_in, s->scale);
}

#define AK_DISPATCH(s, FN, ...)            \
  do {                                     \
    switch ((s)->storage) {                \
      case 0: FN<0>(__VA_ARGS__); break;   \
      case 1: FN<1>(__VA_ARGS__); break;   \
      default: FN<2>(__VA_ARGS__); break;  \
    }                                      \
  } while (0)

template <int S>
void export_any(Lattice *s, double *d) {
  if (s->sweeps % 2 == 0) launch_export<S, true>(s, d);
  else launch_export<S, false>(s, d);
}
template <int S>
void import_any(Lattice *s, const double *d) {
  if (s->sweeps % 2 == 0) launch_import<S, true>(s, d);
  else launch_import<S, false>(s, d);
}

Lattice *H(void *h) { return (Lattice *)h; }

int check_ready(Lattice *s) {
  if (!s) return fail(nullptr, "null lattice handle");
  if (!s->pop) return fail(s, "population buffer not bound (ak_bind_populations)");
  HIPCHK(s, hipSetDevice(s->dev));
  return 0;
}

// Exterior farfield set-up on the initial (post-even) state: ring mask, reservoir links,
// band tables and buffers; the ring at the free stream; the first step's band state, trace
// rows and prediction; then one collision of every cell (collide_stream order).
template <int S>
void ext_initial_launches(Lattice *s) {
  using T = typename St<S>::T;
  using C = typename St<S>::C;
  T *A = (T *)s->pop;
  const ExtDev E = ext_dev(s);
  const long long nr = (long long)s->ext_R2 * s->nz, nb = (long long)s->ext_B2 * s->nz;
  hipLaunchKernelGGL((k_ext_fill_ring<S>), dim3(blocks_for(nr, 128)), dim3(128), 0, s->stream, A, s->L, E, s->scale);
  hipLaunchKernelGGL((k_ext_band<S, true>), dim3(blocks_for(nb, 128)), dim3(128), 0, s->stream, (const T *)A, s->L, E, 0, s->ext_B2, s->inv_scale);
  hipLaunchKernelGGL(k_ext_rows, dim3(blocks_for(2LL * s->nz * s->nx, 128)), dim3(128), 0, s->stream, E, s->d_ext_rows);
  launch_ext_predict(s, nullptr, 0.0);
  const long long n = (long long)s->nx * s->ny * s->nz;
  hipLaunchKernelGGL((k_collide_inplace<S>), dim3(blocks_for(n, 256)), dim3(256), 0, s->stream, A, s->d_solid, s->L,
                     (C)s->ws, (C)s->wb, s->scale, s->inv_scale);
}

// Every link of the legacy farfield is independent (a corner's outflow recomputes the
// bounce-back value it needs), so their order only serves the memory system: velocity
// bounce-back links first, then by z, row and x, with the links of one cell adjacent (they
// share its 27 loads), so that the lanes of a wave take a few neighbouring cells of one z
// plane rather than cells 64 planes apart.
void order_links(int nx, int ny, int nz, std::vector<int32_t> &xs, std::vector<int32_t> &zs,
                 std::vector<int8_t> &dir, std::vector<int8_t> &knd, std::vector<double> &pdf,
                 std::vector<double> &pdfnd) {
  std::vector<size_t> order(xs.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  auto key = [&](size_t i) {
    const long long y = xs[i] >> 16, x = xs[i] & 0xffff;
    return ((((long long)knd[i] * nz + zs[i]) * (ny + 2) + y) * (nx + 2) + x) * 27 + dir[i];
  };
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return key(a) < key(b); });
  auto perm = [&](auto &v) { auto c = v; for (size_t i = 0; i < order.size(); ++i) v[i] = c[order[i]]; };
  perm(xs); perm(zs); perm(dir); perm(knd); perm(pdf); perm(pdfnd);
}

int set_exterior(Lattice *s) {
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  std::vector<uint8_t> m((size_t)(nx + 2) * (ny + 2));
  HIPCHK(s, hipMemcpy(m.data(), s->d_solid, m.size(), hipMemcpyDeviceToHost));
  // band: ring cells first, then the cells one and two inside it
  std::vector<int32_t> bxy, map((size_t)nx * ny, -1);
  for (int pass = 0; pass < 2; ++pass)
    for (int x = 0; x < nx; ++x)
      for (int y = 0; y < ny; ++y) {
        const int depth = std::min(std::min(x, nx - 1 - x), std::min(y, ny - 1 - y));
        if (depth > 2 || (pass == 0) != (depth == 0)) continue;
        map[(size_t)x * ny + y] = (int32_t)bxy.size();
        bxy.push_back(x | (y << 16));
        if (depth == 0) {
          uint8_t &mk = m[(size_t)(y + 1) * (nx + 2) + (x + 1)];
          if (mk == 1) return fail(s, "solid cell (%d, %d) on the exterior ring", x, y);
          mk = 3;
        }
      }
  s->ext_R2 = 2 * nx + 2 * (ny - 2);
  s->ext_B2 = (int)bxy.size();
  s->ext_ring_blocks = (int)blocks_for((long long)s->ext_R2 * nz, 128);
  HIPCHK(s, hipMemcpy(s->d_solid, m.data(), m.size(), hipMemcpyHostToDevice));
  // reservoir: velocity bounce-back of every ring link leaving the storage in x or y
  std::vector<int32_t> xs, zs;
  std::vector<int8_t> dir, knd;
  for (int r = 0; r < s->ext_R2; ++r) {
    const int x = bxy[r] & 0xffff, y = bxy[r] >> 16;
    for (int z = 0; z < nz; ++z)
      for (int d = 0; d < 27; ++d) {
        const int gx = x + cxq(d), gy = y + cyq(d);
        if (gx >= 0 && gx < nx && gy >= 0 && gy < ny) continue;
        xs.push_back((x + 1) | ((y + 1) << 16));
        zs.push_back(z);
        dir.push_back((int8_t)d);
        knd.push_back(0);
      }
  }
  std::vector<double> none(xs.size(), 0.0), none2(xs.size(), 0.0);
  order_links(nx, ny, nz, xs, zs, dir, knd, none, none2);
  s->nleg = (int)xs.size();
  const size_t rz = (size_t)s->ext_R2 * nz * 4, bz = (size_t)s->ext_B2 * nz * 4;
  const std::vector<double> zb(bz, 0.0), zr(rz, 0.0), zrows((size_t)2 * nz * nx, 0.0),
      zpart((size_t)s->ext_ring_blocks * 5, 0.0);
  const std::vector<ExtAcc> acc(1, ExtAcc{});
  const std::vector<int> err(1, 0);
  if (upload(s, &s->d_legx, xs) || upload(s, &s->d_legz, zs) || upload(s, &s->d_legdir, dir) ||
      upload(s, &s->d_legkind, knd) || upload(s, &s->d_pdf, none) || upload(s, &s->d_pdfnd, none) ||
      upload(s, &s->d_ext_bxy, bxy) || upload(s, &s->d_ext_map, map) || upload(s, &s->d_ext_M, zb) ||
      upload(s, &s->d_ext_Y1, zr) || upload(s, &s->d_ext_Y2, zr) || upload(s, &s->d_ext_pred, zr) ||
      upload(s, &s->d_ext_rows, zrows) || upload(s, &s->d_ext_part, zpart) || upload(s, &s->d_ext_acc, acc) ||
      upload(s, &s->d_ext_err, err))
    return -1;
  AK_DISPATCH(s, ext_initial_launches, s);
  HIPCHK(s, hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

const char *ak_version(void) { return "aerokinetic-mi355x 0.4 (gfx950 HIP, AA-pattern D3Q27 cumulant)"; }

int ak_last_error(void *handle, char *buf, size_t len) {
  const std::string &e = handle ? H(handle)->err : g_error;
  if (buf && len) {
    std::snprintf(buf, len, "%s", e.c_str());
  }
  return (int)e.size();
}

int ak_lattice_create(void **handle, int device, int nx, int ny, int nz, int storage, double omega_shear,
                      double omega_bulk, double fp16_scale, int history_capacity, void *stream) {
  if (!handle) return fail(nullptr, "null handle pointer");
  *handle = nullptr;
  if (nx < 4 || ny < 4 || nz < 1 || nx > 65000 || ny > 65000)
    return fail(nullptr, "lattice %dx%dx%d outside 4..65000 x 4..65000 x >=1", nx, ny, nz);
  if (storage < 0 || storage > 2) return fail(nullptr, "unknown storage precision %d", storage);
  if (!(omega_shear > 0 && omega_shear < 2) || !(omega_bulk > 0 && omega_bulk < 2))
    return fail(nullptr, "relaxation rates must lie in (0, 2)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0)
    return fail(nullptr, "HIP device %d not available (%d devices)", device, ndev);
  if (hipSetDevice(device) != hipSuccess) return fail(nullptr, "hipSetDevice(%d) failed", device);
  Lattice *s = new Lattice();
  s->dev = device; s->nx = nx; s->ny = ny; s->nz = nz; s->storage = storage;
  s->ws = omega_shear; s->wb = omega_bulk;
  s->stream = (hipStream_t)stream;
  if (!s->stream) {
    // step graphs are captured on the lattice's stream, which cannot be the null stream;
    // a blocking stream keeps the ordering against work on the null stream
    if (hipStreamCreate(&s->stream) != hipSuccess) {
      delete s;
      return fail(nullptr, "hipStreamCreate failed");
    }
    s->own_stream = true;
  }
  s->elem_size = storage == 0 ? 8 : storage == 1 ? 4 : 2;
  if (storage == 2) {
    const double sc = fp16_scale > 0 ? fp16_scale : 65536.0;
    const double lg = std::log2(sc);
    if (lg != std::floor(lg) || sc < 1.0 || sc > 1048576.0) {
      if (s->own_stream) (void)hipStreamDestroy(s->stream);
      delete s;
      return fail(nullptr, "fp16 scale must be a power of two in [1, 2^20]");
    }
    s->scale = (float)sc;
    s->inv_scale = (float)(1.0 / sc);
  }
  // interior rows start on a 128-byte cache line (x = 0 first; the x = -1 ghost is the
  // last element of the previous row's pad), so a wave's 64 consecutive cells of one
  // direction plane cover whole lines; pitch >= nx + 2 holds both x ghosts
  const long long align = 128 / (long long)s->elem_size;
  long long P = ((long long)nx + 2 + align - 1) / align * align;
  s->L.nx = nx; s->L.ny = ny; s->L.nz = nz;
  s->L.base = align;
  // The 27 direction rows of one y row are adjacent ([z][y][q][x]) whenever a z plane of
  // that layout fits the sweep's 32-bit in-plane byte offsets: a wave's 27 loads and 27
  // stores then fall in one 27-row group instead of 27 planes far apart.  Measured on the
  // sweep's exact in-place access pattern (even + odd, 1536x1024x64, 64x4 tiles), TB/s
  // with one plane per direction -> interleaved rows: FP64 5.02 -> 5.31, FP32 4.87 ->
  // 5.20, FP16 4.57 -> 5.01.  Larger planes keep one plane per direction ([q][z][y][x]).
  const bool row_groups = (unsigned long long)27 * P * (ny + 2) * s->elem_size < (1ull << 31);
  if (row_groups) {
    s->L.QS = P;
    s->L.P = 27 * P;
    s->L.PS = s->L.P * (ny + 2);
    s->pop_elems = (size_t)(s->L.base + (long long)nz * s->L.PS + align);
  } else {
    s->L.P = P;
    s->L.PS = P * (ny + 2);
    s->L.QS = s->L.PS * nz;
    s->pop_elems = (size_t)(s->L.base + 27 * s->L.QS + align);
  }
  // the sweep addresses a cell inside a z plane with a 32-bit byte offset
  if ((unsigned long long)s->L.PS * s->elem_size >= (1ull << 31)) {
    if (s->own_stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return fail(nullptr, "lattice plane of %lld elements too large for 32-bit in-plane offsets", (long long)(P * (ny + 2)));
  }
  s->pop_bytes = s->pop_elems * s->elem_size;
  s->geom = s->geom_odd = default_geom(nx, ny, nz);
  s->geom_small = s->geom;
  if (hipDeviceGetAttribute(&s->lds_per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) != hipSuccess)
    s->lds_per_cu = 0;
  s->hist_cap = history_capacity > 0 ? history_capacity : 4096;
  s->hist_width = AK_HIST_WIDTH;
  if (dalloc(s, &s->d_solid, (size_t)(nx + 2) * (ny + 2)) || dalloc(s, &s->d_diag, 1) ||
      dalloc(s, &s->d_hist, (size_t)s->hist_cap * s->hist_width) || dalloc(s, &s->d_ctl, 1)) {
    for (void *p : s->owned) (void)hipFree(p);
    if (s->own_stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return -1;
  }
  (void)hipMemset(s->d_solid, 0, (size_t)(nx + 2) * (ny + 2));
  (void)hipMemset(s->d_hist, 0, (size_t)s->hist_cap * s->hist_width * sizeof(double));
  // the first sampled sweep accumulates into clean slots; every sampled step's end clears them
  hipLaunchKernelGGL(k_diag_clear, dim3(1), dim3(256), 0, s->stream, s->d_diag);
  *handle = s;
  return 0;
}

int ak_lattice_destroy(void *handle) {
  Lattice *s = H(handle);
  if (!s) return 0;
  (void)hipSetDevice(s->dev);
  (void)hipStreamSynchronize(s->stream);
  s->graphs.clear();
  for (void *p : s->owned) (void)hipFree(p);
  if (s->own_stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return 0;
}

size_t ak_population_bytes(void *handle) { return handle ? H(handle)->pop_bytes : 0; }

int ak_bind_populations(void *handle, void *device_ptr, size_t bytes) {
  Lattice *s = H(handle);
  if (!s) return fail(nullptr, "null lattice handle");
  s->gen++;  // invalidates captured step graphs
  if (!device_ptr) return fail(s, "null population buffer");
  if (bytes < s->pop_bytes) return fail(s, "population buffer of %zu bytes < required %zu", bytes, s->pop_bytes);
  if (((uintptr_t)device_ptr) % 128) return fail(s, "population buffer must be 128-byte aligned");
  s->pop = device_ptr;
  s->pm_valid = s->xo_valid = false;
  HIPCHK(s, hipSetDevice(s->dev));
  // ghosts and padding start at zero
  hipLaunchKernelGGL((k_fill<uint8_t>), dim3(1024), dim3(256), 0, s->stream, (uint8_t *)s->pop, (long long)s->pop_bytes, (uint8_t)0);
  HIPCHK(s, hipGetLastError());
  return 0;
}

int ak_populations_written(void *handle) {
  Lattice *s = H(handle);
  if (!s) return fail(nullptr, "null lattice handle");
  s->pm_valid = s->xo_valid = false;  // the next run call recomputes the perimeter moments from the populations
  return 0;
}

int ak_set_solid(void *handle, const uint8_t *mask_xy) {
  Lattice *s = H(handle);
  if (!s) return fail(nullptr, "null lattice handle");
  s->gen++;  // invalidates captured step graphs
  std::vector<uint8_t> m((size_t)(s->nx + 2) * (s->ny + 2), 0);
  for (int x = 0; x < s->nx; ++x)
    for (int y = 0; y < s->ny; ++y) m[(size_t)(y + 1) * (s->nx + 2) + (x + 1)] = mask_xy[(size_t)x * s->ny + y] > 2 ? 1 : mask_xy[(size_t)x * s->ny + y];
  s->perim_fluid = true;
  s->pm_valid = s->xo_valid = false;
  for (int x = 0; x < s->nx; ++x)
    for (int y = 0; y < s->ny; ++y)
      if ((x == 0 || y == 0 || x == s->nx - 1 || y == s->ny - 1) && m[(size_t)(y + 1) * (s->nx + 2) + (x + 1)] == 1)
        s->perim_fluid = false;
  HIPCHK(s, hipSetDevice(s->dev));
  HIPCHK(s, hipMemcpy(s->d_solid, m.data(), m.size(), hipMemcpyHostToDevice));
  return 0;
}

int ak_set_walls(void *handle, int n, const int32_t *x, const int32_t *y, const int32_t *d, const double *q,
                 const int32_t *element, int nelem, double rx_off, double ry_off, double chord_cells,
                 double lattice_speed) {
  Lattice *s = H(handle);
  if (!s) return fail(nullptr, "null lattice handle");
  s->gen++;  // invalidates captured step graphs
  if (nelem < 1 || nelem > MAX_ELEM) return fail(s, "element count %d outside 1..%d", nelem, MAX_ELEM);
  std::vector<int32_t> xy(n);
  std::vector<int8_t> dir(n), el(n);
  std::vector<double> qq(n);
  for (int k = 0; k < n; ++k) {
    if (x[k] < 0 || x[k] >= s->nx || y[k] < 0 || y[k] >= s->ny) return fail(s, "wall link %d outside the lattice", k);
    if (d[k] < 0 || d[k] > 26 || d[k] == 13) return fail(s, "wall link %d has direction %d", k, d[k]);
    if (element[k] < 0 || element[k] >= nelem) return fail(s, "wall link %d has element %d", k, element[k]);
    const int gx = x[k] + cxq(d[k]), gy = y[k] + cyq(d[k]);
    if (gx < -1 || gx > s->nx || gy < -1 || gy > s->ny) return fail(s, "wall link %d leaves the ghost layer", k);
    if (!(q[k] == -1.0 || (q[k] > 0.0 && q[k] <= 1.0))) return fail(s, "wall link %d has q=%g", k, q[k]);
    xy[k] = x[k] | (y[k] << 16);
    dir[k] = (int8_t)d[k];
    el[k] = (int8_t)element[k];
    qq[k] = q[k];
  }
  HIPCHK(s, hipSetDevice(s->dev));
  if (upload(s, &s->d_wxy, xy) || upload(s, &s->d_wdir, dir) || upload(s, &s->d_wq, qq) || upload(s, &s->d_welem, el))
    return -1;
  s->nlinks = n;
  s->nelem = nelem;
  s->rx_off = rx_off; s->ry_off = ry_off; s->chord = chord_cells; s->U = lattice_speed;
  s->wall_blocks = (int)blocks_for((long long)n * s->nz, 256);
  s->walls_clear = true;
  for (int k = 0; k < n; ++k)
    if (x[k] < 3 || x[k] > s->nx - 4 || y[k] < 3 || y[k] > s->ny - 4) s->walls_clear = false;
  if (s->d_partials) { dfree(s, s->d_partials); s->d_partials = nullptr; }
  if (dalloc(s, &s->d_partials, (size_t)std::max(s->wall_blocks, 1) * nelem * 5)) return -1;
  return 0;
}

int ak_init_seeded(void *handle, const double *p) {
  Lattice *s = H(handle);
  if (check_ready(s)) return -1;
  s->gen++;  // invalidates captured step graphs
  SeedParams sp{p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], p[10]};
  s->pm_valid = s->xo_valid = false;
  const long long n = (long long)s->nx * s->ny * s->nz;
  hipLaunchKernelGGL((k_fill<uint8_t>), dim3(1024), dim3(256), 0, s->stream, (uint8_t *)s->pop, (long long)s->pop_bytes, (uint8_t)0);
  switch (s->storage) {
    case 0: hipLaunchKernelGGL((k_init_seeded<0>), dim3(blocks_for(n, 256)), dim3(256), 0, s->stream, (double *)s->pop, s->d_solid, s->L, sp, s->scale); break;
    case 1: hipLaunchKernelGGL((k_init_seeded<1>), dim3(blocks_for(n, 256)), dim3(256), 0, s->stream, (float *)s->pop, s->d_solid, s->L, sp, s->scale); break;
    default: hipLaunchKernelGGL((k_init_seeded<2>), dim3(blocks_for(n, 256)), dim3(256), 0, s->stream, (_Float16 *)s->pop, s->d_solid, s->L, sp, s->scale); break;
  }
  HIPCHK(s, hipGetLastError());
  s->sweeps = 0;
  s->hist_rows = 0;
  s->host_hist.clear();
  s->U = p[0];
  return 0;
}

int ak_set_farfield(void *handle, int kind, double U) {
  Lattice *s = H(handle);
  if (check_ready(s)) return -1;
  s->gen++;  // invalidates captured step graphs
  if (kind < 0 || kind > 4) return fail(s, "unknown farfield kind %d", kind);
  if ((kind == AK_FARFIELD_CHARACTERISTIC || kind == AK_FARFIELD_EXTERIOR) && s->sweeps % 2 != 0)
    return fail(s, "set the characteristic or exterior farfield on the initial state");
  if (kind == AK_FARFIELD_EXTERIOR && (s->nx < 6 || s->ny < 6 || s->nz < 4))
    return fail(s, "the exterior farfield needs a padded lattice of at least 6 x 6 x 4");
  s->ff_kind = kind;
  s->U = U;
  dfree(s, s->d_pm);
  s->d_pm = nullptr;
  dfree(s, s->d_xo);
  s->d_xo = nullptr;
  dfree(s, s->d_xi);
  s->d_xi = nullptr;
  s->pm_valid = s->xo_valid = false;
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  if (kind == AK_FARFIELD_CHARACTERISTIC) {
    // objects in the order normals (-1,-1),(-1,0),(-1,1),(0,-1),(0,1),(1,-1),(1,0),(1,1)
    int k = 0;
    for (int sx = -1; sx <= 1; ++sx)
      for (int sy = -1; sy <= 1; ++sy) {
        if (!sx && !sy) continue;
        FFObj &o = s->ffp.o[k++];
        o.n[0] = sx; o.n[1] = sy; o.n[2] = 0;
        const double len = std::sqrt((double)(sx * sx + sy * sy));
        o.nh[0] = sx / len; o.nh[1] = sy / len; o.nh[2] = 0.0;
        const double vel[3] = {U, 0.0, 0.0};
        const double ndv = sx * U;
        o.evolve = ndv >= 0 && U != 0.0;
        o.glancing = o.evolve && ndv == 0.0;
        o.un_inf = o.nh[0] * vel[0] + o.nh[1] * vel[1] + o.nh[2] * vel[2];
        for (int a = 0; a < 3; ++a) {
          o.ut_inf[a] = vel[a] - o.un_inf * o.nh[a];
          bool any = false;
          for (int b = 0; b < 3; ++b) {
            o.P[a][b] = (a == b ? 1.0 : 0.0) - o.nh[a] * o.nh[b];
            if (o.P[a][b] != 0.0) any = true;
          }
          o.use_axis[a] = !(o.ut_inf[a] == 0.0 && !any);
        }
        o.side = sy < 0 ? 0 : 1;
      }
    s->ffp.U = U;
    s->ffp.rho0 = 1.0;
    // items: every exterior ghost cell of each object, with the links entering the fluid.
    // Within an object the ghost cells of one z follow each other along the face, so the
    // lanes of a wave read consecutive x of a y face (coalesced) or consecutive rows of an
    // x face, never 64 z planes far apart.
    std::vector<int32_t> g, gz, mask;
    std::vector<int8_t> obj;
    for (int ob = 0; ob < 8; ++ob) {
      const FFObj &o = s->ffp.o[ob];
      std::vector<int32_t> cg, cm;
      for (int gx = -1; gx <= nx; ++gx)
        for (int gy = -1; gy <= ny; ++gy) {
          const int tx = gx < 0 ? -1 : gx >= nx ? 1 : 0, ty = gy < 0 ? -1 : gy >= ny ? 1 : 0;
          if ((!tx && !ty) || tx != o.n[0] || ty != o.n[1]) continue;
          int m = 0;
          for (int d = 0; d < 27; ++d) {
            const int fx = gx - cxq(d), fy = gy - cyq(d);
            if (fx >= 0 && fx < nx && fy >= 0 && fy < ny) m |= 1 << d;
          }
          if (!m) continue;
          cg.push_back((gx + 1) | ((gy + 1) << 16));
          cm.push_back(m);
        }
      for (int z = 0; z < nz; ++z)
        for (size_t c = 0; c < cg.size(); ++c) {
          g.push_back(cg[c]);
          gz.push_back(z);
          mask.push_back(cm[c]);
          obj.push_back((int8_t)ob);
        }
    }
    s->nff = (int)g.size();
    if (upload(s, &s->d_ffg, g) || upload(s, &s->d_ffz, gz) || upload(s, &s->d_ffmask, mask) || upload(s, &s->d_ffobj, obj))
      return -1;
    std::vector<double> zeros(s->nff, 0.0);
    std::vector<long long> zc(s->nff, 0);
    if (upload(s, &s->d_inc, zeros) || upload(s, &s->d_der, zeros) || upload(s, &s->d_clock, zc)) return -1;
    std::vector<double> rows((size_t)2 * nz * nx, 0.0);
    if (upload(s, &s->d_rows, rows)) return -1;
    if (dalloc(s, &s->d_pm, pm_count(nx, ny, nz) * (s->storage == 0 ? sizeof(double) : sizeof(float))) ||
        dalloc(s, &s->d_xo, xo_count(ny, nz) * s->elem_size) || dalloc(s, &s->d_xi, xo_count(ny, nz) * s->elem_size))
      return -1;
    FFItems I{s->d_ffg, s->d_ffz, s->d_ffobj, s->d_ffmask, s->d_inc, s->d_der, s->d_clock, s->d_rows, s->nff, nullptr,
              nullptr, nullptr};
    const long long nr = 2LL * nz * nx;
    switch (s->storage) {
      case 0:
        hipLaunchKernelGGL((k_glancing_rows<0>), dim3(blocks_for(nr, 256)), dim3(256), 0, s->stream, (const double *)s->pop, s->L, s->d_rows, 1.0, s->inv_scale);
        hipLaunchKernelGGL((k_farfield_init<0>), dim3(blocks_for(s->nff, 128)), dim3(128), 0, s->stream, (const double *)s->pop, s->L, I, s->ffp, s->inv_scale);
        break;
      case 1:
        hipLaunchKernelGGL((k_glancing_rows<1>), dim3(blocks_for(nr, 256)), dim3(256), 0, s->stream, (const float *)s->pop, s->L, s->d_rows, 1.0, s->inv_scale);
        hipLaunchKernelGGL((k_farfield_init<1>), dim3(blocks_for(s->nff, 128)), dim3(128), 0, s->stream, (const float *)s->pop, s->L, I, s->ffp, s->inv_scale);
        break;
      default:
        hipLaunchKernelGGL((k_glancing_rows<2>), dim3(blocks_for(nr, 256)), dim3(256), 0, s->stream, (const _Float16 *)s->pop, s->L, s->d_rows, 1.0, s->inv_scale);
        hipLaunchKernelGGL((k_farfield_init<2>), dim3(blocks_for(s->nff, 128)), dim3(128), 0, s->stream, (const _Float16 *)s->pop, s->L, I, s->ffp, s->inv_scale);
        break;
    }
    HIPCHK(s, hipGetLastError());
  } else if (kind == AK_FARFIELD_LEGACY) {
    // velocity bounce-back on x-, y-, y+ ghosts; outflow on x+ ghosts (corners included)
    std::vector<int32_t> xs, zs;
    std::vector<int8_t> dir, knd;
    std::vector<double> pdf, pdfnd;
    for (int x = 0; x < nx; ++x)
      for (int y = 0; y < ny; ++y) {
        if (x > 0 && x < nx - 1 && y > 0 && y < ny - 1) continue;
        for (int z = 0; z < nz; ++z)
          for (int d = 0; d < 27; ++d) {
            const int gx = x + cxq(d), gy = y + cyq(d);
            int kind_ = -1;
            if (gx >= nx) kind_ = 1;
            else if (gx < 0 || gy < 0 || gy >= ny) kind_ = 0;
            if (kind_ < 0) continue;
            xs.push_back((x + 1) | ((y + 1) << 16));
            zs.push_back(z);
            dir.push_back((int8_t)d);
            knd.push_back((int8_t)kind_);
            // outflow history starts at the shifted free-stream equilibrium of the incoming direction
            const int id = 26 - d;
            auto ph = [&](int c, double u) { return c == 0 ? 2.0 / 3.0 - u * u : (1.0 / 3.0 + u * u + c * u) * 0.5; };
            const double e = ph(cxq(id), U) * ph(cyq(id), 0.0) * ph(czq(id), 0.0) - wq(id);
            pdf.push_back(kind_ == 1 ? e : 0.0);
            pdfnd.push_back(kind_ == 1 ? e : 0.0);
          }
      }
    order_links(nx, ny, nz, xs, zs, dir, knd, pdf, pdfnd);
    s->nleg = (int)xs.size();
    if (upload(s, &s->d_legx, xs) || upload(s, &s->d_legz, zs) || upload(s, &s->d_legdir, dir) ||
        upload(s, &s->d_legkind, knd) || upload(s, &s->d_pdf, pdf) || upload(s, &s->d_pdfnd, pdfnd))
      return -1;
  } else if (kind == AK_FARFIELD_EXTERIOR) {
    if (set_exterior(s)) return -1;
  }
  return 0;
}

int ak_set_sweep_tile(void *handle, int bx, int by) {
  Lattice *s = H(handle);
  if (!s) return fail(nullptr, "null lattice handle");
  s->gen++;  // invalidates captured step graphs
  if (bx == 0) {
    if (by < 64 || by % 64 || by > 1024) return fail(s, "row-linear sweep blocks of %d threads: a multiple of 64, <= 1024", by);
  } else if (bx < 64 || bx % 64 || by < 1 || bx * by > 1024) {
    return fail(s, "tile %dx%d: bx multiple of 64, bx*by <= 1024", bx, by);
  }
  set_geom(s, bx, by);
  s->waves_even = s->waves_odd = 0;
  return 0;
}

int ak_set_sweep_waves(void *handle, int odd, int waves) {
  Lattice *s = H(handle);
  if (!s) return fail(nullptr, "null lattice handle");
  if (odd != 0 && odd != 1) return fail(s, "AA form %d: 0 (even) or 1 (odd)", odd);
  if (waves < 0 || waves > 8) return fail(s, "waves per SIMD %d: 0 (no cap) to 8", waves);
  const SweepGeom &G = odd ? s->geom_odd : s->geom;
  const int threads = G.bx * G.by;
  if (waves > 0 && (s->lds_per_cu <= 0 || waves * 4 < (threads + 63) / 64))
    return fail(s, "a cap of %d waves per SIMD does not fit %d-thread sweep blocks", waves, threads);
  if (sweep_lds(s, waves, threads) > 65536)
    return fail(s, "a cap of %d waves per SIMD needs more than 64 KiB of LDS per %d-thread block", waves, threads);
  s->gen++;  // invalidates captured step graphs
  (odd ? s->waves_odd : s->waves_even) = waves;
  return 0;
}

int ak_next_sweep_odd(void *handle) {
  Lattice *s = H(handle);
  if (!s) return fail(nullptr, "null lattice handle");
  return s->sweeps % 2 == 0 ? 1 : 0;
}

int ak_set_sweep_tile_form(void *handle, int odd, int bx, int by) {
  Lattice *s = H(handle);
  if (!s) return fail(nullptr, "null lattice handle");
  if (odd != 0 && odd != 1) return fail(s, "AA form %d: 0 (even) or 1 (odd)", odd);
  const SweepGeom keep = odd ? s->geom : s->geom_odd;
  const int keep_waves = odd ? s->waves_even : s->waves_odd;
  if (ak_set_sweep_tile(handle, bx, by)) return -1;
  (odd ? s->geom : s->geom_odd) = keep;
  (odd ? s->waves_even : s->waves_odd) = keep_waves;
  return 0;
}

int ak_run(void *handle, int nsteps, long long first_step, int sample_every) {
  Lattice *s = H(handle);
  if (check_ready(s)) return -1;
  if (nsteps > 0) {
    set_ctl(s, first_step + 1, s->hist_rows);
    AK_DISPATCH(s, fill_perimeter, s);
  }
  for (int k = 0; k < nsteps; ++k) {
    const long long number = first_step + k + 1;
    const bool sample = sample_every > 0 && number % sample_every == 0;
    if (sample && s->hist_rows >= s->hist_cap) {
      if (flush_history(s)) return -1;
      set_ctl(s, number, 0);
    }
    auto launch = [&] {
      switch (s->storage) {
        case 0: one_step<0>(s, sample); break;
        case 1: one_step<1>(s, sample); break;
        default: one_step<2>(s, sample); break;
      }
    };
    if (s->use_graphs) {
      const int slot = (int)(s->sweeps % 2) * 2 + (sample ? 1 : 0);
      if (graph_step(s, s->stream, s->graphs, slot, s->gen, launch, [&] { s->sweeps++; })) return -1;
    } else {
      launch();
    }
    if (sample) s->hist_rows++;
  }
  HIPCHK(s, hipGetLastError());
  return 0;
}

int ak_set_graph_mode(void *handle, int enable) {
  Lattice *s = H(handle);
  if (!s) return fail(nullptr, "null lattice handle");
  s->use_graphs = enable != 0;
  return 0;
}

// Same launches as ak_run, launched directly, with HIP events around each boundary phase
// on the lattice's stream and on each sweep launch itself (the kernel's own start and
// end).  out = {sweep ms total, boundary ms total, sweeps}.
int ak_run_profiled(void *handle, int nsteps, long long first_step, int sample_every, double *out) {
  Lattice *s = H(handle);
  if (check_ready(s)) return -1;
  std::vector<hipEvent_t> ev(4 * (size_t)nsteps);
  for (auto &e : ev) HIPCHK(s, hipEventCreate(&e));
  if (nsteps > 0) {
    set_ctl(s, first_step + 1, s->hist_rows);
    AK_DISPATCH(s, fill_perimeter, s);
  }
  for (int k = 0; k < nsteps; ++k) {
    const long long number = first_step + k + 1;
    const bool sample = sample_every > 0 && number % sample_every == 0;
    if (sample && s->hist_rows >= s->hist_cap) {
      if (flush_history(s)) return -1;
      set_ctl(s, number, 0);
    }
    switch (s->storage) {
      case 0: one_step<0>(s, sample, &ev[4 * k]); break;
      case 1: one_step<1>(s, sample, &ev[4 * k]); break;
      default: one_step<2>(s, sample, &ev[4 * k]); break;
    }
    if (sample) s->hist_rows++;
  }
  hipEvent_t last = nullptr;
  HIPCHK(s, hipEventCreate(&last));
  HIPCHK(s, hipEventRecord(last, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  float span = 0.f;
  if (nsteps > 0) HIPCHK(s, hipEventElapsedTime(&span, ev[0], last));
  (void)hipEventDestroy(last);
  double sweep = 0.0, bnd = 0.0;
  for (int k = 0; k < nsteps; ++k) {
    float a = 0.f, b = 0.f;
    HIPCHK(s, hipEventElapsedTime(&a, ev[4 * k], ev[4 * k + 1]));
    HIPCHK(s, hipEventElapsedTime(&b, ev[4 * k + 2], ev[4 * k + 3]));
    bnd += a;
    sweep += b;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  out[0] = sweep;
  out[1] = bnd;
  out[2] = nsteps;
  out[3] = span;
  return 0;
}

int ak_history_width(void *handle) { return handle ? H(handle)->hist_width : 0; }

int ak_read_history(void *handle, double *out, int max_rows) {
  Lattice *s = H(handle);
  if (check_ready(s)) return -1;
  if (flush_history(s)) return -1;
  const int rows = (int)(s->host_hist.size() / s->hist_width);
  const int n = std::min(rows, max_rows);
  if (out && n) std::memcpy(out, s->host_hist.data(), (size_t)n * s->hist_width * sizeof(double));
  s->host_hist.erase(s->host_hist.begin(), s->host_hist.begin() + (size_t)n * s->hist_width);
  return n;
}

int ak_get_populations(void *handle, double *host_out) {
  Lattice *s = H(handle);
  if (check_ready(s)) return -1;
  const size_t n = (size_t)s->nx * s->ny * s->nz * 27;
  double *d = nullptr;
  HIPCHK(s, hipMalloc((void **)&d, n * sizeof(double)));
  AK_DISPATCH(s, export_any, s, d);
  hipError_t e = hipStreamSynchronize(s->stream);
  if (e == hipSuccess) e = hipMemcpy(host_out, d, n * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(s, "population export: %s", hipGetErrorString(e));
  return 0;
}

int ak_set_populations(void *handle, const double *host_in) {
  Lattice *s = H(handle);
  if (check_ready(s)) return -1;
  s->gen++;  // invalidates captured step graphs
  s->pm_valid = s->xo_valid = false;
  const size_t n = (size_t)s->nx * s->ny * s->nz * 27;
  double *d = nullptr;
  HIPCHK(s, hipMalloc((void **)&d, n * sizeof(double)));
  hipError_t e = hipMemcpy(d, host_in, n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    AK_DISPATCH(s, import_any, s, d);
    e = hipStreamSynchronize(s->stream);
  }
  (void)hipFree(d);
  if (e != hipSuccess) return fail(s, "population import: %s", hipGetErrorString(e));
  return 0;
}

}  // extern "C"

namespace {
template <int S, bool PE>
void launch_diag_state(Lattice *s, double *part, unsigned nb) {
  using T = typename St<S>::T;
  if (s->pre_collision)
    hipLaunchKernelGGL((k_diag_state<S, PE, false, true>), dim3(nb), dim3(256), 0, s->stream, (const T *)s->pop, s->d_solid, s->L, s->inv_scale, part);
  else if (s->ff_kind == AK_FARFIELD_PERIODIC)
    hipLaunchKernelGGL((k_diag_state<S, PE, true, false>), dim3(nb), dim3(256), 0, s->stream, (const T *)s->pop, s->d_solid, s->L, s->inv_scale, part);
  else
    hipLaunchKernelGGL((k_diag_state<S, PE, false, false>), dim3(nb), dim3(256), 0, s->stream, (const T *)s->pop, s->d_solid, s->L, s->inv_scale, part);
}
// blocks of the diagnostics reduction: one thread per cell up to DIAG_MAX_BLOCKS blocks, which
// then stride over the lattice
unsigned diag_blocks(Lattice *s) {
  return (unsigned)std::min<long long>(blocks_for((long long)s->nx * s->ny * s->nz, 256), DIAG_MAX_BLOCKS);
}
template <int S>
void diag_any(Lattice *s, double *part, unsigned nb) {
  if (s->sweeps % 2 == 0) launch_diag_state<S, true>(s, part, nb);
  else launch_diag_state<S, false>(s, part, nb);
}
template <int S, bool PE>
void launch_plane(Lattice *s, int z, double *r, double *u) {
  using T = typename St<S>::T;
  const long long n = (long long)s->nx * s->ny;
  if (s->pre_collision)
    hipLaunchKernelGGL((k_plane<S, PE, false, true>), dim3(blocks_for(n, 256)), dim3(256), 0, s->stream, (const T *)s->pop, s->L, z, s->inv_scale, r, u);
  else if (s->ff_kind == AK_FARFIELD_PERIODIC)
    hipLaunchKernelGGL((k_plane<S, PE, true, false>), dim3(blocks_for(n, 256)), dim3(256), 0, s->stream, (const T *)s->pop, s->L, z, s->inv_scale, r, u);
  else
    hipLaunchKernelGGL((k_plane<S, PE, false, false>), dim3(blocks_for(n, 256)), dim3(256), 0, s->stream, (const T *)s->pop, s->L, z, s->inv_scale, r, u);
}
template <int S>
void plane_any(Lattice *s, int z, double *r, double *u) {
  if (s->sweeps % 2 == 0) launch_plane<S, true>(s, z, r, u);
  else launch_plane<S, false>(s, z, r, u);
}
}  // namespace

extern "C" {

int ak_diagnostics(void *handle, double *out) {
  Lattice *s = H(handle);
  if (check_ready(s)) return -1;
  const unsigned nb = diag_blocks(s);
  double *d = nullptr;
  HIPCHK(s, hipMalloc((void **)&d, ((size_t)nb * 5 + 8) * sizeof(double)));
  AK_DISPATCH(s, diag_any, s, d, nb);
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, s->stream, (const double *)d, (int)nb, d + (size_t)nb * 5);
  hipError_t e = hipMemcpyAsync(out, d + (size_t)nb * 5, 5 * sizeof(double), hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(s, "diagnostics: %s", hipGetErrorString(e));
  return 0;
}

int ak_get_macroscopic(void *handle, double *rho, double *u) {
  Lattice *s = H(handle);
  if (check_ready(s)) return -1;
  // one z plane at a time on the device (the plane kernel), scattered to [x][y][z] on the host
  const size_t n = (size_t)s->nx * s->ny;
  std::vector<double> h(4 * n);
  double *d = nullptr;
  HIPCHK(s, hipMalloc((void **)&d, n * 4 * sizeof(double)));
  hipError_t e = hipSuccess;
  for (int z = 0; z < s->nz && e == hipSuccess; ++z) {
    AK_DISPATCH(s, plane_any, s, z, d, d + n);
    e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) e = hipMemcpy(h.data(), d, 4 * n * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) break;
    for (size_t c = 0; c < n; ++c) {
      const size_t o = c * s->nz + z;
      rho[o] = h[c];
      u[3 * o] = h[n + 3 * c]; u[3 * o + 1] = h[n + 3 * c + 1]; u[3 * o + 2] = h[n + 3 * c + 2];
    }
  }
  (void)hipFree(d);
  if (e != hipSuccess) return fail(s, "macroscopic export: %s", hipGetErrorString(e));
  return 0;
}

int ak_get_plane(void *handle, int z, double *rho, double *u) {
  Lattice *s = H(handle);
  if (check_ready(s)) return -1;
  if (z < 0 || z >= s->nz) return fail(s, "plane z=%d outside 0..%d", z, s->nz - 1);
  const size_t n = (size_t)s->nx * s->ny;
  double *d = nullptr;
  HIPCHK(s, hipMalloc((void **)&d, n * 4 * sizeof(double)));
  AK_DISPATCH(s, plane_any, s, z, d, d + n);
  hipError_t e = hipStreamSynchronize(s->stream);
  if (e == hipSuccess) e = hipMemcpy(rho, d, n * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(u, d + n, 3 * n * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(s, "plane export: %s", hipGetErrorString(e));
  return 0;
}

int ak_synchronize(void *handle) {
  Lattice *s = H(handle);
  if (!s) return fail(nullptr, "null lattice handle");
  HIPCHK(s, hipSetDevice(s->dev));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return 0;
}

}  // extern "C"

namespace {

// Restartable state of a lattice: a fixed header, the raw population buffer (every storage
// precision bit-exact), then the boundary histories: characteristic farfield incoming and
// derivative values, clocks and glancing trace rows; legacy outflow pdf and pdf_nd; the
// exterior ring's prediction, band state, trace rows and accumulators.
struct StateHeader {
  char magic[8];
  int version, storage, nx, ny, nz, ff_kind, nff, nleg, has_prev;
  long long sweeps;
  unsigned long long pop_bytes;
};
constexpr int STATE_VERSION = 4;

struct StatePart {
  void *dev;
  size_t bytes;
};

std::vector<StatePart> state_parts(Lattice *s, bool with_prev) {
  std::vector<StatePart> p{{s->pop, s->pop_bytes}};
  if (s->ff_kind == AK_FARFIELD_CHARACTERISTIC && s->nff) {
    p.push_back({s->d_inc, (size_t)s->nff * sizeof(double)});
    p.push_back({s->d_der, (size_t)s->nff * sizeof(double)});
    p.push_back({s->d_clock, (size_t)s->nff * sizeof(long long)});
    p.push_back({s->d_rows, (size_t)2 * s->nz * s->nx * sizeof(double)});
    if (s->d_pm)
      p.push_back({s->d_pm, pm_count(s->nx, s->ny, s->nz) * (s->storage == 0 ? sizeof(double) : sizeof(float))});
  }
  if (s->ff_kind == AK_FARFIELD_LEGACY && s->nleg) {
    p.push_back({s->d_pdf, (size_t)s->nleg * sizeof(double)});
    p.push_back({s->d_pdfnd, (size_t)s->nleg * sizeof(double)});
  }
  if (s->ff_kind == AK_FARFIELD_EXTERIOR && s->ext_R2) {
    // predicted ring state, band state, initial trace rows, accumulators, error flag
    p.push_back({s->d_ext_pred, (size_t)s->ext_R2 * s->nz * 4 * sizeof(double)});
    p.push_back({s->d_ext_M, (size_t)s->ext_B2 * s->nz * 4 * sizeof(double)});
    p.push_back({s->d_ext_rows, (size_t)2 * s->nz * s->nx * sizeof(double)});
    p.push_back({s->d_ext_acc, sizeof(ExtAcc)});
    p.push_back({s->d_ext_err, sizeof(int)});
  }
  if (with_prev) p.push_back({s->d_prev, (size_t)4 * s->nx * s->ny * s->nz * sizeof(double)});
  return p;
}

// Bytes of a state of this lattice with or without the residual snapshot.
size_t state_size(Lattice *s, bool with_prev) {
  size_t n = sizeof(StateHeader);
  for (const StatePart &p : state_parts(s, with_prev)) n += p.bytes;
  return n;
}

StateHeader state_header(const Lattice *s) {
  StateHeader h{};
  std::memcpy(h.magic, "AKSTATE", 8);
  h.version = STATE_VERSION;
  h.storage = s->storage; h.nx = s->nx; h.ny = s->ny; h.nz = s->nz;
  h.ff_kind = s->ff_kind; h.nff = s->nff; h.nleg = s->nleg;
  h.has_prev = s->prev_valid ? 1 : 0;
  h.sweeps = s->sweeps;
  h.pop_bytes = s->pop_bytes;
  return h;
}

// Bytes of the lattice state blob at `host` (read from its header), 0 if it is not one.
size_t state_blob_size(Lattice *s, const void *host, size_t avail) {
  StateHeader h{};
  if (avail < sizeof h) return 0;
  std::memcpy(&h, host, sizeof h);
  if (std::memcmp(h.magic, "AKSTATE", 8)) return 0;
  return state_size(s, h.has_prev != 0);
}

}  // namespace

extern "C" {

size_t ak_state_bytes(void *handle) {
  Lattice *s = H(handle);
  return s ? state_size(s, s->prev_valid) : 0;
}

int ak_save_state(void *handle, void *host, size_t bytes) {
  Lattice *s = H(handle);
  if (check_ready(s)) return -1;
  if (!host || bytes < ak_state_bytes(handle)) return fail(s, "state buffer of %zu bytes < %zu", bytes, ak_state_bytes(handle));
  AK_DISPATCH(s, fill_perimeter, s);  // the saved perimeter moments are those of the saved populations
  HIPCHK(s, hipStreamSynchronize(s->stream));
  const StateHeader h = state_header(s);
  char *o = (char *)host;
  std::memcpy(o, &h, sizeof h);
  o += sizeof h;
  for (const StatePart &p : state_parts(s, s->prev_valid)) {
    HIPCHK(s, hipMemcpy(o, p.dev, p.bytes, hipMemcpyDeviceToHost));
    o += p.bytes;
  }
  return 0;
}

int ak_load_state(void *handle, const void *host, size_t bytes) {
  Lattice *s = H(handle);
  if (check_ready(s)) return -1;
  s->gen++;  // invalidates captured step graphs
  StateHeader h{};
  if (!host || bytes < sizeof h) return fail(s, "state buffer too short");
  std::memcpy(&h, host, sizeof h);
  const StateHeader c = state_header(s);
  if (std::memcmp(h.magic, c.magic, 8) || h.version != STATE_VERSION)
    return fail(s, "not a lattice state (or another state version)");
  if (h.storage != c.storage || h.nx != c.nx || h.ny != c.ny || h.nz != c.nz || h.ff_kind != c.ff_kind ||
      h.nff != c.nff || h.nleg != c.nleg || h.pop_bytes != c.pop_bytes)
    return fail(s, "state of another lattice (storage, shape or boundary set differ)");
  // the residual snapshot travels with the state when it was taken
  const bool with_prev = h.has_prev != 0;
  if (bytes != state_size(s, with_prev)) return fail(s, "state size %zu, expected %zu", bytes, state_size(s, with_prev));
  if (with_prev && !s->d_prev && dalloc(s, &s->d_prev, (size_t)4 * s->nx * s->ny * s->nz)) return -1;
  s->prev_valid = with_prev;
  HIPCHK(s, hipStreamSynchronize(s->stream));
  const char *i = (const char *)host + sizeof h;
  for (const StatePart &p : state_parts(s, with_prev)) {
    HIPCHK(s, hipMemcpy(p.dev, i, p.bytes, hipMemcpyHostToDevice));
    i += p.bytes;
  }
  s->sweeps = h.sweeps;
  s->pm_valid = s->d_pm != nullptr;
  s->xo_valid = false;  // not in the blob: copied again from the loaded populations (bit for bit)
  return 0;
}

}  // extern "C"

namespace {
template <int S, bool PE>
void launch_residual(Lattice *s, bool incoming, double inv_u, double inv_dt, int reset, unsigned nb) {
  using T = typename St<S>::T;
  const T *A = (const T *)s->pop;
  if (incoming)
    hipLaunchKernelGGL((k_residual<S, PE, false, true>), dim3(nb), dim3(256), 0, s->stream, A, s->d_solid, s->L, s->inv_scale, inv_u, inv_dt, reset, s->d_prev, s->d_rpart);
  else if (s->ff_kind == AK_FARFIELD_PERIODIC)
    hipLaunchKernelGGL((k_residual<S, PE, true, false>), dim3(nb), dim3(256), 0, s->stream, A, s->d_solid, s->L, s->inv_scale, inv_u, inv_dt, reset, s->d_prev, s->d_rpart);
  else
    hipLaunchKernelGGL((k_residual<S, PE, false, false>), dim3(nb), dim3(256), 0, s->stream, A, s->d_solid, s->L, s->inv_scale, inv_u, inv_dt, reset, s->d_prev, s->d_rpart);
}
template <int S>
void residual_any(Lattice *s, bool incoming, double inv_u, double inv_dt, int reset, unsigned nb) {
  if (s->sweeps % 2 == 0) launch_residual<S, true>(s, incoming, inv_u, inv_dt, reset, nb);
  else launch_residual<S, false>(s, incoming, inv_u, inv_dt, reset, nb);
}
}  // namespace

extern "C" {

int ak_residual(void *handle, int incoming, double lattice_speed, double dt, int reset, double *out) {
  Lattice *s = H(handle);
  if (check_ready(s)) return -1;
  if (!(lattice_speed > 0) || !(dt > 0)) return fail(s, "residual needs a positive lattice speed and interval");
  const long long n = (long long)s->nx * s->ny * s->nz;
  const unsigned nb = blocks_for(n, 256);
  if (!s->d_prev && dalloc(s, &s->d_prev, (size_t)4 * n)) return -1;
  if (!s->d_rpart && (dalloc(s, &s->d_rpart, (size_t)9 * nb) || dalloc(s, &s->d_rout, 9))) return -1;
  if (!s->prev_valid) reset = 1;
  AK_DISPATCH(s, residual_any, s, incoming != 0, 1.0 / lattice_speed, 1.0 / dt, reset, nb);
  hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(64), 0, s->stream, (const double *)s->d_rpart, (int)nb, s->d_rout);
  HIPCHK(s, hipStreamSynchronize(s->stream));
  HIPCHK(s, hipMemcpy(out, s->d_rout, 9 * sizeof(double), hipMemcpyDeviceToHost));
  s->prev_valid = true;
  return 0;
}

int ak_exterior_state(void *handle, double *out) {
  Lattice *s = H(handle);
  if (check_ready(s)) return -1;
  if (s->ff_kind != AK_FARFIELD_EXTERIOR) return fail(s, "the lattice has no exterior farfield");
  ExtAcc a{};
  int err = 0;
  HIPCHK(s, hipStreamSynchronize(s->stream));
  HIPCHK(s, hipMemcpy(&a, s->d_ext_acc, sizeof a, hipMemcpyDeviceToHost));
  HIPCHK(s, hipMemcpy(&err, s->d_ext_err, sizeof err, hipMemcpyDeviceToHost));
  out[0] = a.wall_mass; out[1] = a.wall_total;
  for (int c = 0; c < 4; ++c) { out[2 + c] = a.src[c]; out[6 + c] = a.src_total[c]; }
  out[10] = a.moment_err;
  out[11] = (double)a.steps;
  out[12] = err;
  return 0;
}

int ak_device_copy(void *dst, const void *src, size_t bytes, void *stream) {
  if (!dst || !src) return fail(nullptr, "null copy buffer");
  if (((uintptr_t)dst | (uintptr_t)src | bytes) % 8) return fail(nullptr, "copy needs 8-byte aligned buffers and size");
  // chunks of 2^33 bytes keep the grid below 2^32 work-items
  const size_t chunk = (size_t)1 << 33;
  for (size_t off = 0; off < bytes; off += chunk) {
    const long long n = (long long)(std::min(chunk, bytes - off) / 8);
    hipLaunchKernelGGL(k_stream_copy, dim3(blocks_for(n, 512)), dim3(256), 0, (hipStream_t)stream,
                       (const uint64_t *)((const char *)src + off), (uint64_t *)((char *)dst + off), n);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(nullptr, "ak_device_copy: %s", hipGetErrorString(e));
  return 0;
}

}  // extern "C"

#include "ak_refined_capi.hpp"
#include "ak_batch.hpp"
#include "ak_batch_capi.hpp"

#include "ak_noalloc.hpp"
