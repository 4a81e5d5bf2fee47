/*
 * aerokinetic.h -- C ABI of the MI355X D3Q27 cumulant lattice-Boltzmann step.
 *
 * One opaque lattice = one uniform D3Q27 level of one airfoil case on one GPU.
 * All entry points return 0 on success and a negative code on failure; the
 * message is available from ak_last_error().  No entry point throws, and none
 * allocates device memory in the step path (ak_run / ak_step_*): buffers are
 * created by ak_lattice_create / ak_set_* or bound by the caller.
 *
 * Each entry point replaces a piece of the reference's lbmpy/pystencils kernel
 * plug-in boundary (paths inside the reference repository):
 *
 *   ak_lattice_create        LatticeBoltzmannStep(...) kernel build
 *                            (src/aerokinetic/solver.py:58-91, interface.py:116-130)
 *   ak_set_solid/ak_set_walls  solid masks + NoSlipLinearBouzidi index lists
 *                            (solver.py:93-122,139-143, walls.py:21-49)
 *   ak_set_farfield          install_characteristic_farfield / UBB +
 *                            ExtrapolationOutflow (farfield.py:209-224, solver.py:130-138)
 *   ak_init_seeded           _initialize + set_pdf_fields_from_macroscopic_values
 *                            (solver.py:157-202)
 *   ak_run                   the per-step body of Solver.step: _sync_src,
 *                            boundary_handling, _loads, run_kernel(stream_collide),
 *                            swap (solver.py:252-279)
 *   ak_read_history          loads + diagnostics of sampled steps
 *                            (solver.py:204-250, 281-297)
 *   ak_diagnostics           post_run + the low-Mach guards of Solver.diagnostics
 *                            (solver.py:281-297)
 *   ak_get_plane             the mid-span gather of save_frame (solver.py:312-334)
 *   ak_get/set_populations   gather_array / cpu_arrays of the pdf field (checkpoint.py:32-68)
 *   ak_refined_*             RefinedSolver + D3Q27Transfer + StreamingFaceFlux: the two-level
 *                            2:1 sub-cycled step (refined_solver.py:116-345,
 *                            interface.py:116-352, refinement.py:117-254)
 *
 * Population exchange layout (host, FP64): g[x][y][z][27], zero-centred
 * g = f - w, direction i = 9 (cx+1) + 3 (cy+1) + (cz+1).
 */
#ifndef AEROKINETIC_H
#define AEROKINETIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* storage precision of the populations */
enum { AK_STORE_F64 = 0, AK_STORE_F32 = 1, AK_STORE_F16 = 2 };
/* outer boundary */
enum { AK_FARFIELD_PERIODIC = 0, AK_FARFIELD_LEGACY = 1, AK_FARFIELD_CHARACTERISTIC = 2,
       AK_FARFIELD_NONE = 3 /* x/y ghosts left to the caller (refined fine level) */,
       AK_FARFIELD_EXTERIOR = 4 /* exterior_regularized_fd_v1: the lattice is the physical box
                                   padded by one cell in x and y (src/aerokinetic/exterior.py:18-200) */ };
/* one history row per sampled step (doubles) */
enum {
  AK_HIST_STEP = 0, AK_HIST_CL = 1, AK_HIST_CD = 2, AK_HIST_CM = 3,
  AK_HIST_MACH = 4, AK_HIST_DENSITY_ERR = 5, AK_HIST_RHO_MIN = 6, AK_HIST_NONFINITE = 7,
  AK_HIST_MASS_EXCHANGE = 8, AK_HIST_FIXED = 9 /* then nelem*4 per-element loads Fx,Fy,Fz,Mz */,
  /* exterior farfield only, after room for 8 elements' loads (solver.py:443-445, 265-271) */
  AK_HIST_EXT_WALL_MASS_TOTAL = 41, AK_HIST_EXT_SOURCE = 42, AK_HIST_EXT_SOURCE_TOTAL = 43,
  AK_HIST_WIDTH = 44
};

/* Create a lattice of nx*ny*nz fluid cells (z periodic) on `device`.
 * omega_shear / omega_bulk: relaxation rates; fp16_scale: power-of-two scale of
 * FP16 shifted storage (ignored otherwise; 0 selects the default);
 * history_capacity: sampled rows kept on the device between reads;
 * stream: hipStream_t for every launch (NULL = default stream). */
int ak_lattice_create(void **handle, int device, int nx, int ny, int nz, int storage,
                      double omega_shear, double omega_bulk, double fp16_scale,
                      int history_capacity, void *stream);
int ak_lattice_destroy(void *handle);

/* Bytes the caller must provide for the population buffer, and binding of a
 * caller-owned, 128-byte aligned device buffer (e.g. a torch tensor's data_ptr()). */
size_t ak_population_bytes(void *handle);
int ak_bind_populations(void *handle, void *device_ptr, size_t bytes);

/* The caller wrote the bound population buffer directly (outside the library): state the
 * library derives from the populations and keeps between run calls (the characteristic
 * farfield's perimeter moments) is recomputed at the next run call.  The reference writes
 * its pdf arrays only through the lbmpy macroscopic setter (solver.py:157-202). */
int ak_populations_written(void *handle);

/* Cell mask of the interior xy cells (host, nx*ny bytes, x-major: m[x*ny+y]):
 * 0 fluid, 1 solid (not updated), 2 fluid excluded from the diagnostics. */
int ak_set_solid(void *handle, const uint8_t *mask_xy);

/* Fluid-to-solid links of the xy plane, extruded over z (host arrays of n):
 * fluid cell (x, y), direction d (solid at x + c_d), Bouzidi q (-1 = halfway
 * fallback), element index.  Lever arm = (index + 0.5 + r_off) in lattice units;
 * chord_cells, lattice_speed normalise the coefficients. */
int ak_set_walls(void *handle, int n, const int32_t *x, const int32_t *y, const int32_t *d,
                 const double *q, const int32_t *element, int nelem, double rx_off, double ry_off,
                 double chord_cells, double lattice_speed);

/* Outer boundary; reads the current populations, so call after the initial state.
 * AK_FARFIELD_EXTERIOR (ExteriorBoundary, exterior.py:18-50; solver.py:81-90): the lattice is
 * the physical box padded by one cell in x and y (at least 6 x 6 x 4).  The outermost x/y cells
 * are the exterior ring (mask value 3, no solid allowed there): they are set to the free
 * stream, the first step's ring prediction is taken from the initial state, and every cell is
 * collided once (the reference's collide -> boundary -> stream order); the storage edge's
 * ghosts are a velocity bounce-back reservoir.  History rows then also carry
 * AK_HIST_EXT_* columns and ak_exterior_state reports the accumulators. */
int ak_set_farfield(void *handle, int kind, double lattice_speed);

/* Seeded initial field (u_x = U plus the spanwise-periodic Gaussian-envelope
 * perturbation; u = 0 in solid cells; rho = 1) as product-form equilibrium.
 * p = {U, amp, phase_y, phase_z, ky, kz, lex, ley, ox, oy, dx}. */
int ak_init_seeded(void *handle, const double *p);

/* Thread tile of the sweep (bx a multiple of 64, bx*by <= 1024), or with bx = 0 row-linear
 * blocks of `by` threads (a multiple of 64) over consecutive cells of a plane; set by the tile
 * autotuner.  Default: 64 x 4 when nx is a multiple of 64, row-linear 256 otherwise. */
int ak_set_sweep_tile(void *handle, int bx, int by);

/* The same for one AA form only (odd = 1: the gathering sweep, 0: the in-place sweep), so the
 * two forms can run different tiles (their access patterns differ: the odd form's x-shifted
 * rows straddle cache lines that neighbouring waves share).  Replaces the sweep launch
 * configuration that lbmpy's generated kernel fixes at code generation (solver.py:58-91). */
int ak_set_sweep_tile_form(void *handle, int odd, int bx, int by);

/* Cap on the resident waves per SIMD of one AA form's sweep (0: no cap, else 1..8), enforced
 * by the LDS each sweep block reserves (the sweep itself uses none).  Fewer concurrent waves,
 * each streaming 27 population rows in and out, can reach a higher memory rate than a full
 * CU.  Setting a tile clears the caps.  Like the tile, a launch configuration of the sweep
 * that replaces the one lbmpy fixes at code generation (solver.py:58-91). */
int ak_set_sweep_waves(void *handle, int odd, int waves);

/* AA form of the lattice's next sweep: 1 odd (gathering), 0 even (in place), -1 on error.
 * The reference's two-array step has no such state (solver.py:252-279 swaps src and dst); the
 * tile autotuner uses it to time each form on its own. */
int ak_next_sweep_odd(void *handle);

/* Advance `nsteps` reference steps.  Steps with (first_step + k + 1) % sample_every == 0
 * (k = 0..nsteps-1; sample_every <= 0 disables) record a history row. */
int ak_run(void *handle, int nsteps, long long first_step, int sample_every);

/* Step graphs (default on): ak_run captures each launch variant of a step (AA parity,
 * sampled or not) as a HIP graph on first use and replays it; any setter call
 * (ak_set_*, ak_bind_populations, ak_init_seeded) invalidates the captured graphs.
 * enable = 0 launches every kernel directly.  Replaces the per-step Python kernel
 * dispatch of Solver.step (solver.py:252-279). */
int ak_set_graph_mode(void *handle, int enable);

/* ak_run with HIP events (measurement only; direct launches; synchronises at the end):
 * each sweep launch records its own kernel start and end, each boundary phase is
 * bracketed on the lattice's stream.
 * out = {total sweep milliseconds, total boundary milliseconds, steps, milliseconds of all the
 * steps from the first boundary phase to the end of the last step}. */
int ak_run_profiled(void *handle, int nsteps, long long first_step, int sample_every, double *out);

/* Copy up to max_rows recorded rows to host (row length ak_history_width) and
 * clear the device history; returns the row count (or < 0 on error). */
int ak_history_width(void *handle);
int ak_read_history(void *handle, double *out, int max_rows);

/* Populations of the interior, FP64, in the exchange layout (synchronises). */
int ak_get_populations(void *handle, double *host_out);
int ak_set_populations(void *handle, const double *host_in);

/* Macroscopic fields of the interior: rho[x][y][z], u[x][y][z][3] (host, FP64). */
int ak_get_macroscopic(void *handle, double *rho, double *u);

/* Diagnostics of the current state over fluid cells (synchronises):
 * out = {local Mach max, max |rho-1|, min rho, non-finite flag, sum of rho}. */
int ak_diagnostics(void *handle, double *out);

/* ak_diagnostics without any allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid it).  Same kernels, same reduction order and the same result as
 * ak_diagnostics; the block partials live in a file-scope device array of 160 KiB per GPU
 * (4096 blocks x 5 doubles + the result), held until the process ends.  The reduction strides
 * over the lattice with at most 4096 blocks, so every lattice fits.  Calls that overlap on one
 * GPU share the array and race. */
int ak_diagnostics_noalloc(void *handle, double *out);

/* rho[x][y] and u[x][y][3] of one z plane (host, FP64; synchronises). */
int ak_get_plane(void *handle, int z, double *rho, double *u);

/* Restartable device state (checkpoint.py:32-135, refined_checkpoint.py:54-142): the raw
 * population buffer (bit-exact in every storage precision), the AA parity and the boundary
 * histories (characteristic incoming/derivative/clock/glancing trace; legacy outflow
 * pdf/pdf_nd; exterior ring prediction, band state, trace rows and accumulators).
 * ak_load_state accepts only a state of a lattice with the same storage, shape and
 * boundary set. */
size_t ak_state_bytes(void *handle);
int ak_save_state(void *handle, void *host, size_t bytes);
int ak_load_state(void *handle, const void *host, size_t bytes);

/* Stationarity residual (convergence.py:55-96, 138-158): d = (q - q_prev) / dt with
 * q = (rho, jx / U, jy / U, jz / U) on owned fluid cells (mask 0), from the stored
 * post-collision state (incoming = 0, uniform lattices) or the incoming pre-collision state
 * (incoming = 1, refined levels); then q_prev = q.  out = {sum d^2 [4], max |d| [4], cells}.
 * reset = 1 (or no snapshot yet) only records q_prev (out = zeros).  The snapshot is part of
 * the restartable state once taken. */
int ak_residual(void *handle, int incoming, double lattice_speed, double dt, int reset, double *out);

/* Exterior farfield bookkeeping (ExteriorBoundary.step report and source_total,
 * exterior.py:187-200; Solver._wall_mass_exchange/_wall_mass_total, solver.py:262-271):
 * out[13] = {wall mass exchange of the last step, accumulated; exterior source of the last
 * step (mass, momentum x/y/z), accumulated [4]; largest rebuild mass/momentum mismatch of the
 * last step; exterior steps taken; 1 if a ring prediction was non-finite}. */
int ak_exterior_state(void *handle, double *out);

/* Wait for all work of this lattice. */
int ak_synchronize(void *handle);

/* ---- Two-level 2:1 refined step -------------------------------------------------------------
 * One refined handle owns a coarse lattice (nx*ny*nz, legacy farfield, z periodic) and a fine
 * lattice over the half-open coarse box [box_lo, box_hi) in x/y (full span in z), padded by two
 * fine cells on every x/y side.  Stored state on both levels is pre-collision
 * (collide_stream order).  One coarse step = donor moments, coarse sweep, coarse face flux and
 * farfield, two fine substeps (time-interpolated prolongation, fine sweep, fine face flux,
 * walls + loads), reflux of the face-flux mismatch, restriction into the covered coarse cells. */

/* RefinedSolver.__init__ / _build (refined_solver.py:116-215, interface.py:116-130,230-283).
 * level = {dx_c, dt_c, tau_shear_c, tau_bulk_c} of the coarse level (the fine level is its
 * acoustic refinement, refinement.py:58-65); lattice_speed = U (coarse lattice units). */
int ak_refined_create(void **handle, int device, int nx, int ny, int nz, const int32_t *box_lo,
                      const int32_t *box_hi, int storage, const double *level, double lattice_speed,
                      double fp16_scale, int history_capacity, void *stream);
int ak_refined_destroy(void *handle);

/* Lattice handle of one level (0 coarse, 1 fine) for the uniform entry points
 * (populations binding, masks, diagnostics, history reads). */
void *ak_refined_level(void *handle, int level);

/* Fine-level Bouzidi links in padded fine coordinates (refined_solver.py:186-203,
 * walls.py:21-49); lever arms about (origin_fx, origin_fy), common units. */
int ak_refined_set_walls(void *handle, int n, const int32_t *x, const int32_t *y, const int32_t *d,
                         const double *q, const int32_t *element, int nelem, double origin_fx,
                         double origin_fy);

/* Seeded initial field on both levels, restriction into the covered coarse cells and the
 * legacy farfield state (refined_solver.py:216-247, solver.py:157-202).
 * p = {U, amp, phase_y, phase_z, ky, kz, le_x, le_y, origin_x, origin_y}. */
int ak_refined_init_seeded(void *handle, const double *p);

/* nsteps coarse steps (RefinedSolver.step, refined_solver.py:311-345, interface.py:311-352);
 * loads and diagnostics of the sampled steps go to the coarse level's history rows. */
int ak_refined_run(void *handle, int nsteps, long long first_step, int sample_every);

/* Graph replay of the whole coarse step (both substeps and all transfers), as
 * ak_set_graph_mode; replaces the Python orchestration of _step_once
 * (interface.py:311-352) and RefinedSolver.step (refined_solver.py:311-345). */
int ak_refined_set_graph_mode(void *handle, int enable);

/* ak_refined_run with HIP events (measurement; synchronises):
 * out = {total ms, coarse sweep ms, fine sweep ms, steps}. */
int ak_refined_run_profiled(void *handle, int nsteps, long long first_step, int sample_every, double *out);

/* Transfer guard flags raised since the last call, then cleared (refinement.py:159-163,187-202):
 * 1 density <= 0, 2 negative population, 4 local Mach > 0.3, 8 non-finite. */
int ak_refined_transfer_errors(void *handle);

/* Restartable state of a two-level case: coarse-step count plus both lattice states. */
size_t ak_refined_state_bytes(void *handle);
int ak_refined_save_state(void *handle, void *host, size_t bytes);
int ak_refined_load_state(void *handle, const void *host, size_t bytes);

/* Pre-collision populations of one level (0 coarse, 1 fine padded), exchange layout. */
int ak_refined_get_populations(void *handle, int level, double *host_out);

/* FP64 conservation budget of the case, common units (RefinedSolver.conserved_integrals,
 * step and diagnostics bookkeeping: refined_solver.py:291-361; the outer face-flux samplers,
 * refined_solver.py:215-222, 271-280).  Kept on the device: every coarse step adds its wall
 * mass exchange, wall impulse and outer-face in-flux to the accumulators; the conserved
 * integrals (mass, momentum) of both levels' owned cells are taken at the initial state and
 * around the last step of every ak_refined_run call.  out[32] = {wall impulse x/y/z of the
 * last step, its wall moment impulse, its wall mass exchange, mass and momentum x/y/z in
 * through the outer faces in the last step, cumulative wall mass exchange, cumulative wall
 * impulse [3], cumulative outer in-flux [4], integrals of the initial state [4], before [4]
 * and after [4] the last step of the last run call, coarse steps accumulated, the step that
 * before/after bracket (0: none), 1 if every step since the initial state was accumulated
 * (0 once a batch has stepped the case: batches keep no budget)}.  Synchronises. */
int ak_refined_budget(void *handle, double *out);

/* ---- Batches of two-level cases (serving) --------------------------------------------------
 * No counterpart in the reference, which runs one case per process (Solver.run /
 * refined_run.run, solver.py:353-537, refined_run.py:19-157; polar sweeps loop over such
 * runs).  A batch steps several refined cases of one storage type together: the members'
 * coarse sweeps run as ONE launch over their concatenated tile lists, likewise each fine
 * substep's sweeps and each transfer, farfield, wall and load kernel (over the members'
 * concatenated work lists).  Each batch step is a captured HIP graph (per launch variant).  A member keeps
 * its own step count, history rows and buffers, and gives the same result alone or in any
 * mix (same kernels, same arithmetic, fixed-order reductions).  Errors: ak_last_error(NULL). */
int ak_batch_create(void **handle, int device, int storage, void *stream);
int ak_batch_destroy(void *handle);
/* Add / remove a refined case (ak_refined_create handle, at a coarse-step boundary; work
 * queued on its stream is ordered before the batch's next step).  Membership changes
 * re-capture the batch graph on the next run. */
int ak_batch_add(void *handle, void *refined);
int ak_batch_remove(void *handle, void *refined);
int ak_batch_size(void *handle);
int ak_batch_set_graph_mode(void *handle, int enable);
/* Launch shape of the batched sweep (no counterpart in the reference, which runs one case per
 * process): 0 (default) one block per 256-cell tile of the concatenated tile list; blocks > 0
 * a persistent grid of that many blocks (rounded down to a multiple of 8, at least 8) that
 * loops over the tiles.  Results are the same either way. */
int ak_batch_set_sweep_grid(void *handle, int blocks);
/* Enqueue nsteps coarse steps of every member on the batch stream (returns before they run);
 * sample_every 1 (loads of every step into each member's history) or 0.  Call
 * ak_batch_synchronize before reading a member's history (ak_read_history on
 * ak_refined_level(member, 0)) or state; cases may be created meanwhile. */
int ak_batch_run(void *handle, int nsteps, int sample_every);
int ak_batch_synchronize(void *handle);

/* Device-to-device streaming copy of `bytes` (8-byte aligned buffers and size) on
 * `stream`.  No counterpart in the reference: it measures the attainable HBM
 * bandwidth beside the stencil kernels and is the known-size kernel that calibrates
 * the memory traffic counters. */
int ak_device_copy(void *dst, const void *src, size_t bytes, void *stream);

/* Last error message of this handle (or of the library when handle is NULL). */
int ak_last_error(void *handle, char *buf, size_t len);

/* Library build identifier. */
const char *ak_version(void);

#ifdef __cplusplus
}
#endif
#endif
