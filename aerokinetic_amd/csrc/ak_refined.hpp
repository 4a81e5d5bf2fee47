// ak_refined.hpp -- kernels of the two-level 2:1 coupling (coarse lattice + padded fine box).
//
// Both levels are AA lattices whose state between sweeps is post-collision; the
// pre-collision value of a cell is what the next sweep gathers (loc_in).  Every
// interface operation therefore reads or writes loc_in slots between sweeps:
//   donor moments   -- gather pre-collision coarse values of the donor block (box +- 3)
//   prolong/inject  -- time-interpolated, 5-point prolonged, stress-rescaled and
//                      regularised populations written into the fine donor cells' slots
//   face flux       -- exact moment flux of pull streaming through the box faces,
//                      from post-collision values (coarse once, fine per substep)
//   reflux          -- equilibrium correction of the coarse cells outside the faces
//   restrict        -- 2x2x2 mean of fine moments, rescaled and regularised, written
//                      into the covered coarse cells' slots
// Transfer arithmetic runs in the storage's compute type (FP64 for FP64 storage, FP32 for
// FP32 and FP16 storage: TR<S>); the load sums and the conservation budget stay FP64.
#pragma once
#include "ak_kernels.hpp"

namespace ak {

struct Levels {
  // stress scales coarse->fine and fine->coarse: tau*dt ratios of shear and bulk
  double shear_cf, bulk_cf, shear_fc, bulk_fc;
};

struct BoxGeom {
  int lo[3], hi[3];  // coarse cell bounds (z spans the whole periodic span)
  int nb[3];         // box shape in coarse cells
  int pad;           // fine donor layers (2)
};

enum { TE_DENSITY = 1, TE_NEGATIVE = 2, TE_MACH = 4, TE_NONFINITE = 8 };

// compute type of the transfer arithmetic of storage S
template <int S>
using TR = typename St<S>::C;

// Moments of the shifted populations g = f - w as running sums in index order 0..26.  Every
// moments computation goes through add(), so a form that streams the populations (no
// 27-value array live in registers) is bit-identical to one that holds them.
template <typename R>
struct MomentSum {
  R d = 0, jx = 0, jy = 0, jz = 0, sxx = 0, syy = 0, szz = 0, sxy = 0, sxz = 0, syz = 0, mn = R(1e30);
  __device__ __forceinline__ void add(int i, R v) {
    d += v;
    jx += cxq(i) * v; jy += cyq(i) * v; jz += czq(i) * v;
    sxx += cxq(i) * cxq(i) * v; syy += cyq(i) * cyq(i) * v; szz += czq(i) * czq(i) * v;
    sxy += cxq(i) * cyq(i) * v; sxz += cxq(i) * czq(i) * v; syz += cyq(i) * czq(i) * v;
    const R f = v + R(wq(i));
    mn = f < mn ? f : mn;
  }
  __device__ __forceinline__ void finish(R (&m)[10], int &err) const {
    const R rho = d + R(1);
    if (!(rho > 0)) err |= TE_DENSITY;
    if (mn < 0) err |= TE_NEGATIVE;
    if (!isfinite(rho) || !isfinite(jx + jy + jz)) err |= TE_NONFINITE;
    if (sqrt(jx * jx + jy * jy + jz * jz) / rho / R(0.57735026918962576451) > R(0.3)) err |= TE_MACH;
    m[0] = rho; m[1] = jx; m[2] = jy; m[3] = jz;
    m[4] = sxx - d / R(3) - jx * jx / rho;
    m[5] = syy - d / R(3) - jy * jy / rho;
    m[6] = szz - d / R(3) - jz * jz / rho;
    m[7] = sxy - jx * jy / rho;
    m[8] = sxz - jx * jz / rho;
    m[9] = syz - jy * jz / rho;
  }
};

template <typename R>
__device__ __forceinline__ void moments_of(const R (&g)[NQ], R (&m)[10], int &err) {
  MomentSum<R> s;
#pragma unroll
  for (int i = 0; i < NQ; ++i) s.add(i, g[i]);
  s.finish(m, err);
}

// tensor-product equilibrium + second-order Hermite stress, shifted by the weights; pop(i)
// gives one population, so callers may store each as it is made
template <typename R>
struct PopBasis {
  R rho, ph[3][3];
  __device__ __forceinline__ PopBasis(const R (&m)[10], int &err) {
    rho = m[0];
    if (!(rho > 0)) err |= TE_DENSITY;
    const R u[3] = {m[1] / rho, m[2] / rho, m[3] / rho};
    if (sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) / R(0.57735026918962576451) > R(0.3)) err |= TE_MACH;
    for (int a = 0; a < 3; ++a) {
      ph[a][1] = R(2.0 / 3.0) - u[a] * u[a];
      ph[a][0] = (R(1.0 / 3.0) + u[a] * u[a] - u[a]) * R(0.5);
      ph[a][2] = (R(1.0 / 3.0) + u[a] * u[a] + u[a]) * R(0.5);
    }
  }
  __device__ __forceinline__ R pop(int i, const R (&m)[10], int &err) const {
    const int cx = cxq(i), cy = cyq(i), cz = czq(i);
    const R eq = rho * ph[0][cx + 1] * ph[1][cy + 1] * ph[2][cz + 1];
    const R t3 = R(1.0 / 3.0);
    const R h = (R(cx * cx) - t3) * m[4] + (R(cy * cy) - t3) * m[5] + (R(cz * cz) - t3) * m[6]
              + R(2) * (R(cx * cy) * m[7] + R(cx * cz) * m[8] + R(cy * cz) * m[9]);
    const R p = eq + R(wq(i) * 4.5) * h;
    if (!(p >= 0)) err |= (isfinite(p) ? TE_NEGATIVE : TE_NONFINITE);
    return p - R(wq(i));
  }
};

template <typename R>
__device__ __forceinline__ void populations_of(const R (&m)[10], R (&g)[NQ], int &err) {
  const PopBasis<R> b(m, err);
#pragma unroll
  for (int i = 0; i < NQ; ++i) g[i] = b.pop(i, m, err);
}

template <typename R>
__device__ __forceinline__ void rescale(R (&m)[10], R shear, R bulk) {
  const R iso = (m[4] + m[5] + m[6]) / R(3);
  for (int k = 4; k < 10; ++k) m[k] *= shear;
  for (int k = 4; k < 7; ++k) m[k] += (bulk - shear) * iso;
}

template <int S, bool PE, typename R>
__device__ __forceinline__ void scatter_in(typename St<S>::T *A, const Desc &L, int x, int y, int z, float scale,
                                           const R (&g)[NQ]) {
#pragma unroll
  for (int i = 0; i < NQ; ++i) stv<S>(A + loc_in<PE>(L, x, y, z, i), (typename St<S>::C)g[i], scale);
}

// A population as the next coarse sweep (and the donor moments) read it back after
// scatter_in stored it: the storage rounding, applied in registers.
template <int S>
__device__ __forceinline__ TR<S> stored_value(TR<S> v, float scale, float inv) {
  typename St<S>::T t;
  stv<S>(&t, v, scale);
  return ld<S>(&t, inv);
}

__device__ __forceinline__ bool in_core(const BoxGeom &B, int bx, int by) {
  // covered cell (box coordinates) more than three rows inside the box in x and y
  return bx >= 3 && bx < B.nb[0] - 3 && by >= 3 && by < B.nb[1] - 3;
}

struct FreeState {
  double g[NQ];  // free-stream shifted populations
  double m[10];  // their moments
};

// Pre-collision moments of the coarse donor block (box +- 3 cells; z wraps).
// SKIP 1: leave out the covered cells outside the core (the restriction that writes their
// populations writes their moments too); SKIP 2: also the reflux cells (the coarse cells
// next to the box faces, whose reflux writes theirs).
template <int S, bool PE, int SKIP = 0>
__device__ __forceinline__ void donor_moments_body(const typename St<S>::T *__restrict__ A, Desc L, BoxGeom B, FreeState F, float inv_scale, TR<S> *__restrict__ M, int *__restrict__ err, long long t) {
  using R = TR<S>;
  const int dx = B.nb[0] + 6, dy = B.nb[1] + 6, dz = B.nb[2] + 6;
  if (t >= (long long)dx * dy * dz) return;
  const int i = (int)(t % dx), j = (int)((t / dx) % dy), k = (int)(t / ((long long)dx * dy));
  const int x = B.lo[0] - 3 + i, y = B.lo[1] - 3 + j;
  const int z = ((B.lo[2] - 3 + k) % L.nz + L.nz) % L.nz;
  R m[10];
  int e = 0;
  const int bx = x - B.lo[0], by = y - B.lo[1];
  const bool inx = bx >= 0 && bx < B.nb[0], iny = by >= 0 && by < B.nb[1];
  const bool covered = inx && iny;
  const bool refluxed = (inx && (by == -1 || by == B.nb[1])) || (iny && (bx == -1 || bx == B.nb[0]));
  if (covered && in_core(B, bx, by)) {
    for (int c = 0; c < 10; ++c) m[c] = R(F.m[c]);
  } else if ((SKIP >= 1 && covered) || (SKIP >= 2 && refluxed)) {
    return;
  } else {
    MomentSum<R> ms;
#pragma unroll
    for (int i = 0; i < NQ; ++i) ms.add(i, ld<S>(A + loc_in<PE>(L, x, y, z, i), inv_scale));
    ms.finish(m, e);
  }
  for (int c = 0; c < 10; ++c) M[t * 10 + c] = m[c];
  if (e) atomicOr(err, e);
}

template <int S, bool PE>
__global__ void k_donor_moments(const typename St<S>::T *__restrict__ A, Desc L, BoxGeom B, FreeState F,
                                float inv_scale, TR<S> *__restrict__ M, int *__restrict__ err) {
  donor_moments_body<S, PE>(A, L, B, F, inv_scale, M, err, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Fine donor cells (the two pad layers in x and y) from the time-interpolated,
// prolonged, rescaled coarse moments.  The 5-point conservative prolongation is
// separable, so it runs as three one-axis passes over the donor block:
//   x pass (with the time interpolation): (nbx+6, nby+6, nz+6) parents -> (2nbx+4, nby+6, nz+6)
//   y pass: -> (2nbx+4, 2nby+4, nz+6), only columns that contain donor cells
//   z pass + stress rescale + regularised rebuild, only donor cells, written into
//   the gather slots of the fine lattice.
// Child 2P of parent P takes a0 - off, child 2P+1 takes a0 + off, with
// off = 11/64 (a1 - a_-1) - 3/128 (a2 - a_-2).
template <typename R>
__device__ __forceinline__ R prolong_weight(int side, int k) {
  // k = -2..2
  const R lo[5] = {R(-3.0 / 128.0), R(11.0 / 64.0), R(1), R(-11.0 / 64.0), R(3.0 / 128.0)};
  return side ? -lo[k + 2] + (k == 0 ? R(2) : R(0)) : lo[k + 2];
}

// Values of one substep's x-pass output (the x pass of both substeps is one launch: it
// needs only the coarse donor moments at t and t + dt, which the fine substeps leave alone).
__host__ __device__ inline size_t prolong_x_half(const BoxGeom &B) {
  return (size_t)(2 * B.nb[0] + 4) * (B.nb[1] + 6) * (B.nb[2] + 6) * 10;
}

// x pass of both substeps: substep s (time fraction s / 2) goes to X + s * prolong_x_half.
template <typename R>
__device__ __forceinline__ void prolong_x_body(const R *__restrict__ M0, const R *__restrict__ M1, BoxGeom B,
                                               R *__restrict__ X, long long t) {
  const int dx = B.nb[0] + 6, dy = B.nb[1] + 6, dz = B.nb[2] + 6, cxn = 2 * B.nb[0] + 4;
  if (t >= (long long)cxn * dy * dz) return;
  const int cx = (int)(t % cxn), qy = (int)((t / cxn) % dy), qz = (int)(t / ((long long)cxn * dy));
  // the y pass only reads columns of donor cells: every row next to an x pad, and only
  // the parent rows of the y pads (qy <= 4 or qy >= nby+1) elsewhere
  if (cx >= B.pad && cx < cxn - B.pad && qy > 4 && qy < B.nb[1] + 1) return;
  const int P = cx / 2 + 2, sd = cx & 1;
  // each donor value is loaded once for both time fractions (0 and 1/2); every term keeps
  // the form w * ((1 - frac) * M0 + frac * M1)
  R m0[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, m1[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int a = -2; a <= 2; ++a) {
    const R w = prolong_weight<R>(sd, a);
    const long long idx = (((long long)qz * dy + qy) * dx + (P + a)) * 10;
#pragma unroll
    for (int k = 0; k < 10; ++k) {
      const R v0 = M0[idx + k], v1 = M1[idx + k];
      m0[k] += w * ((R(1) - R(0)) * v0 + R(0) * v1);
      m1[k] += w * ((R(1) - R(0.5)) * v0 + R(0.5) * v1);
    }
  }
  R *o0 = X + t * 10, *o1 = X + prolong_x_half(B) + t * 10;
#pragma unroll
  for (int k = 0; k < 10; ++k) { o0[k] = m0[k]; o1[k] = m1[k]; }
}

// The prolongation kernels run over compact lists of the items that do work (the host
// builds them with the bodies' own conditions): the fine interface band is a few percent of
// the box volume.
template <typename R>
__global__ void k_prolong_x(const R *__restrict__ M0, const R *__restrict__ M1, BoxGeom B,
                            R *__restrict__ X, const int32_t *__restrict__ list, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) prolong_x_body(M0, M1, B, X, list[i]);
}

// The y pass and the z pass + injection of the prolongation for a group of PZ_COLS fine
// columns of the interface band (fine columns outside the owned box): the y pass
// of a column at every donor z goes to LDS (dz * 10 values per column, dynamic shared
// memory), then its fine cells (the fine cells outside the owned box, every z) interpolate
// in z from there, are rescaled fine <- coarse, regularised and written into the slots the
// next fine sweep reads.
constexpr int PZ_COLS = 4;
__host__ __device__ inline size_t prolong_yz_lds(const BoxGeom &B, size_t value_bytes) {
  return (size_t)PZ_COLS * (B.nb[2] + 6) * 10 * value_bytes;
}

template <int S, bool PE>
__device__ __forceinline__ void prolong_yz_body(const TR<S> *__restrict__ X, typename St<S>::T *__restrict__ A, Desc L,
                                                BoxGeom B, Levels V, float scale, int *__restrict__ err,
                                                const int32_t *__restrict__ cols, int ncol, int group) {
  using R = TR<S>;
  extern __shared__ __align__(16) unsigned char yz_bytes[];
  R *yz = (R *)yz_bytes;  // [PZ_COLS][dz][10]
  const int dy = B.nb[1] + 6, dz = B.nb[2] + 6, cxn = 2 * B.nb[0] + 4;
  const int c0 = group * PZ_COLS;
  const int nc = ncol - c0 < PZ_COLS ? ncol - c0 : PZ_COLS;
  for (int i = threadIdx.x; i < nc * dz; i += blockDim.x) {
    const int k = i / dz, qz = i % dz;
    const int col = cols[c0 + k], cx = col % cxn, cy = col / cxn;
    const int P = cy / 2 + 2, sd = cy & 1;
    R m[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = -2; b <= 2; ++b) {
      const R w = prolong_weight<R>(sd, b);
      const R *q = X + (((long long)qz * dy + (P + b)) * cxn + cx) * 10;
      for (int kk = 0; kk < 10; ++kk) m[kk] += w * q[kk];
    }
    R *o = yz + ((size_t)k * dz + qz) * 10;
    for (int kk = 0; kk < 10; ++kk) o[kk] = m[kk];
  }
  __syncthreads();
  const int p = B.pad;
  for (int i = threadIdx.x; i < nc * L.nz; i += blockDim.x) {
    const int k = i / L.nz, z = i % L.nz;
    const int col = cols[c0 + k], x = col % cxn, y = col / cxn;
    const int pz = z + p;
    const int P = pz / 2 + 2, sd = pz & 1;
    R m[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = -2; c <= 2; ++c) {
      const R w = prolong_weight<R>(sd, c);
      const R *q = yz + ((size_t)k * dz + (P + c)) * 10;
      for (int kk = 0; kk < 10; ++kk) m[kk] += w * q[kk];
    }
    rescale(m, R(V.shear_cf), R(V.bulk_cf));
    int e = 0;
    const PopBasis<R> pb(m, e);
#pragma unroll
    for (int i = 0; i < NQ; ++i) stv<S>(A + loc_in<PE>(L, x, y, z, i), (typename St<S>::C)pb.pop(i, m, e), scale);
    if (e) atomicOr(err, e);
  }
}

template <int S, bool PE>
__global__ __launch_bounds__(64) void k_prolong_yz(const TR<S> *__restrict__ X, typename St<S>::T *__restrict__ A,
                                                   Desc L, BoxGeom B, Levels V, float scale, int *__restrict__ err,
                                                   const int32_t *__restrict__ cols, int ncol) {
  prolong_yz_body<S, PE>(X, A, L, B, V, scale, err, cols, ncol, (int)blockIdx.x);
}

// Face flux registers.  Face f = 2*axis + (side > 0), axis in {0 (x faces), 1 (y faces)}.
struct FluxTerm {
  int8_t d, s0, s1;  // direction, tangential shifts (in-plane axis, z)
  double coef[4];    // weight * normal * [1, c]
};
struct FluxTable {  // host-built, copied once into a device buffer
  FluxTerm term[2][52];
  int n[2];
};

// FLUX_LANES lanes per face element: lane j sums terms j, j + FLUX_LANES, ... in that
// order, then a fixed-order butterfly over the element's lanes adds the partial sums, so
// the register value does not depend on scheduling.
constexpr int FLUX_LANES = 8;

template <int S, bool PE>
__device__ __forceinline__ void flux_terms(const typename St<S>::T *A, const Desc &L, const FluxTable *__restrict__ T,
                                           int axis, int face_cell, int tpos, int zpos, float inv, int sub_lane,
                                           bool active, TR<S> (&out)[4]) {
  using R = TR<S>;
  out[0] = out[1] = out[2] = out[3] = R(0);
  if (active) {
    for (int k = sub_lane; k < T->n[axis]; k += FLUX_LANES) {
      const FluxTerm &tm = T->term[axis][k];
      const int nrm = axis == 0 ? cxq(tm.d) : cyq(tm.d);
      const int a = face_cell + (nrm < 0 ? 1 : 0);
      const int tt = tpos - tm.s0;
      const int zz = ((zpos - tm.s1) % L.nz + L.nz) % L.nz;
      const int x = axis == 0 ? a : tt, y = axis == 0 ? tt : a;
      const R v = ld<S>(A + loc_out<PE>(L, x, y, zz, tm.d), inv);
      for (int c = 0; c < 4; ++c) out[c] += v * R(tm.coef[c]);
    }
  }
#pragma unroll
  for (int o = FLUX_LANES / 2; o >= 1; o >>= 1)
    for (int c = 0; c < 4; ++c) out[c] += __shfl_xor(out[c], o, 64);
}

// Coarse face registers reg[f][t][z][4] (integrated over one coarse step: x dx_c^3).
template <int S, bool PE>
__device__ __forceinline__ void flux_coarse_body(const typename St<S>::T *__restrict__ A, Desc L, BoxGeom B, const FluxTable *__restrict__ T, TR<S> vol, float inv, TR<S> *__restrict__ reg, long long t) {
  const int nt = B.nb[0] > B.nb[1] ? B.nb[0] : B.nb[1];
  const long long w = t / FLUX_LANES;  // face element; its lanes are consecutive within a wave
  const int sub_lane = (int)(t % FLUX_LANES);
  const long long ne = 4LL * nt * L.nz;
  const int zpos = (int)(w % L.nz), ti = (int)((w / L.nz) % nt), f = (int)(w / ((long long)L.nz * nt));
  const int axis = (f / 2) & 1, side = f % 2;
  const int other = 1 - axis;
  const bool active = w < ne && ti < B.nb[other];
  const int face_cell = side ? B.hi[axis] - 1 : B.lo[axis] - 1;
  TR<S> o[4];
  flux_terms<S, PE>(A, L, T, axis, face_cell, B.lo[other] + ti, zpos, inv, sub_lane, active, o);
  if (active && sub_lane == 0) {
    TR<S> *r = reg + (((long long)f * nt + ti) * L.nz + zpos) * 4;
    for (int c = 0; c < 4; ++c) r[c] = o[c] * vol;
  }
}

template <int S, bool PE>
__global__ __launch_bounds__(256) void k_flux_coarse(const typename St<S>::T *__restrict__ A, Desc L, BoxGeom B,
                                                     const FluxTable *__restrict__ T, TR<S> vol, float inv,
                                                     TR<S> *__restrict__ reg) {
  flux_coarse_body<S, PE>(A, L, B, T, vol, inv, reg, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Fine face registers per fine face element regf[f][t_f][z_f][4] (x dx_f^3), summed over
// the two substeps (accumulate = 1 on the second); the reflux folds the 2x2 children.
template <int S, bool PE>
__device__ __forceinline__ void flux_fine_body(const typename St<S>::T *__restrict__ A, Desc L, BoxGeom B, const FluxTable *__restrict__ T, TR<S> vol, float inv, int accumulate, TR<S> *__restrict__ regf, long long t) {
  const int nt = 2 * (B.nb[0] > B.nb[1] ? B.nb[0] : B.nb[1]);
  const long long w = t / FLUX_LANES;  // face element; its lanes are consecutive within a wave
  const int sub_lane = (int)(t % FLUX_LANES);
  const long long ne = 4LL * nt * L.nz;
  const int zf = (int)(w % L.nz), tf = (int)((w / L.nz) % nt), f = (int)(w / ((long long)L.nz * nt));
  const int axis = (f / 2) & 1, side = f % 2;
  const int other = 1 - axis;
  const bool active = w < ne && tf < 2 * B.nb[other];
  const int n_axis = 2 * B.nb[axis];
  const int face_cell = side ? B.pad + n_axis - 1 : B.pad - 1;
  TR<S> o[4];
  flux_terms<S, PE>(A, L, T, axis, face_cell, B.pad + tf, zf, inv, sub_lane, active, o);
  if (active && sub_lane == 0) {
    TR<S> *r = regf + (((long long)f * nt + tf) * L.nz + zf) * 4;
    for (int c = 0; c < 4; ++c) r[c] = (accumulate ? r[c] : TR<S>(0)) + o[c] * vol;
  }
}

template <int S, bool PE>
__global__ __launch_bounds__(256) void k_flux_fine(const typename St<S>::T *__restrict__ A, Desc L, BoxGeom B,
                                                   const FluxTable *__restrict__ T, TR<S> vol, float inv,
                                                   int accumulate, TR<S> *__restrict__ regf) {
  flux_fine_body<S, PE>(A, L, B, T, vol, inv, accumulate, regf, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// ------------------------------------------------------------------ conservation budget
// FP64 bookkeeping of one two-level case (refined_solver.py:291-345): per coarse step the
// wall mass exchange and wall impulse of both fine substeps and the mass and momentum that
// enter through the coarse domain's outer x/y faces, accumulated over the steps since the
// initial state; the conserved integrals (mass, momentum) of both levels' owned cells are
// taken at the initial state and around the last step of a run call.  Common units:
// volumes dx^3 of each level, momenta times dx/dt (the same on both levels).
struct BudgetAcc {
  double wall[5];   // last step: wall impulse x/y/z, wall moment impulse z, wall mass exchange
  double ext[4];    // last step: mass and momentum x/y/z entering through the outer faces
  double cum_wall_mass, cum_wall_impulse[3], cum_ext[4];
  double initial[4], before[4], after[4];  // conserved integrals: mass, momentum x/y/z
  long long steps;        // coarse steps folded since the initial state
  long long budget_step;  // the step that `before` and `after` bracket (0: none yet)
};

// Block sum of four values per thread (wave butterflies, then the waves in order by thread
// 0); the whole block must call it.  Returns the totals in thread 0.
__device__ __forceinline__ void block_sum4(double (&v)[4], double (&red)[16][4]) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1)
    for (int c = 0; c < 4; ++c) v[c] += __shfl_xor(v[c], o, 64);
  const int w = threadIdx.x / 64;
  if (threadIdx.x % 64 == 0)
    for (int c = 0; c < 4; ++c) red[w][c] = v[c];
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) / 64;
    for (int c = 0; c < 4; ++c) {
      double s = red[0][c];
      for (int k = 1; k < nw; ++k) s += red[k][c];
      v[c] = s;
    }
  }
}

// Mass and momentum entering the coarse domain through its four outer x/y faces during one
// coarse step: the pull-streaming face flux (the reference's outer StreamingFaceFlux
// samplers over the face between the ghost layer and the first interior row,
// refined_solver.py:215-222, 276-279) of the post-collision state after the farfield has
// written the incoming values.  A ghost cell's population pointing into the domain is the
// incoming value of its link, stored where the next sweep gathers it (loc_in of the
// interior cell it feeds); a ghost population whose target is another ghost cell is never
// written by a boundary and holds the free-stream value of the initial state.  The z faces
// are left out: on the periodic span they carry one and the same flux in and out.
// Face elements (face, tangential cell, z) take FLUX_LANES lanes each; each block writes one
// partial [4] (lattice units, sign -side applied per element).
template <int S, bool PE>
__device__ __forceinline__ void outer_flux_body(const typename St<S>::T *__restrict__ A, Desc L,
                                                const FluxTable *__restrict__ T, const FreeState &F, float inv,
                                                double *__restrict__ part, long long t, int blk) {
  __shared__ double red[16][4];
  const long long w = t / FLUX_LANES;
  const int sub_lane = (int)(t % FLUX_LANES);
  const long long nfx = (long long)L.ny * L.nz, nfy = (long long)L.nx * L.nz;
  const bool active = w < 2 * (nfx + nfy);
  int f = 0, ti = 0, zpos = 0;
  if (active) {
    long long r = w;
    if (r < 2 * nfx) {
      f = (int)(r / nfx);
      r %= nfx;
    } else {
      r -= 2 * nfx;
      f = 2 + (int)(r / nfy);
      r %= nfy;
    }
    zpos = (int)(r % L.nz);
    ti = (int)(r / L.nz);
  }
  const int axis = f / 2, side = f % 2;
  const int face_cell = side ? (axis == 0 ? L.nx : L.ny) - 1 : -1;
  double o[4] = {0.0, 0.0, 0.0, 0.0};
  if (active) {
    for (int k = sub_lane; k < T->n[axis]; k += FLUX_LANES) {
      const FluxTerm &tm = T->term[axis][k];
      const int d = tm.d;
      const int nrm = axis == 0 ? cxq(d) : cyq(d);
      const int a = face_cell + (nrm < 0 ? 1 : 0);
      const int tt = ti - tm.s0;
      const int zz = ((zpos - tm.s1) % L.nz + L.nz) % L.nz;
      const int x = axis == 0 ? a : tt, y = axis == 0 ? tt : a;
      double v;
      if (x >= 0 && x < L.nx && y >= 0 && y < L.ny) {
        v = (double)ld<S>(A + loc_out<PE>(L, x, y, zz, d), inv);
      } else {
        const int tx = x + cxq(d), ty = y + cyq(d);
        if (tx >= 0 && tx < L.nx && ty >= 0 && ty < L.ny)
          v = (double)ld<S>(A + loc_in<PE>(L, tx, ty, ((zz + czq(d)) % L.nz + L.nz) % L.nz, d), inv);
        else
          v = F.g[d];
      }
      for (int c = 0; c < 4; ++c) o[c] += v * tm.coef[c];
    }
  }
#pragma unroll
  for (int s = FLUX_LANES / 2; s >= 1; s >>= 1)
    for (int c = 0; c < 4; ++c) o[c] += __shfl_xor(o[c], s, 64);
  const double sgn = side ? -1.0 : 1.0;
  double v[4];
  for (int c = 0; c < 4; ++c) v[c] = (active && sub_lane == 0) ? sgn * o[c] : 0.0;
  block_sum4(v, red);
  if (threadIdx.x == 0)
    for (int c = 0; c < 4; ++c) part[(long long)blk * 4 + c] = v[c];
}

// Conserved integrals of one level (refined_solver.py:291-299): sums of rho = 1 + sum g and of
// j = sum c g over the owned cells (mask 0: coarse cells outside the box, fine fluid cells of
// the box) from the pre-collision state, FP64; one partial [4] per block.
template <int S, bool PE>
__global__ __launch_bounds__(256) void k_integrals(const typename St<S>::T *__restrict__ A,
                                                   const uint8_t *__restrict__ mask, Desc L, float inv,
                                                   double *__restrict__ part) {
  __shared__ double red[16][4];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (t < (long long)L.nx * L.ny * L.nz) {
    const int x = (int)(t % L.nx), y = (int)((t / L.nx) % L.ny), z = (int)(t / ((long long)L.nx * L.ny));
    if (mask[(long long)(y + 1) * (L.nx + 2) + (x + 1)] == 0) {
      double r = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
#pragma unroll
      for (int i = 0; i < NQ; ++i) {
        const double g = (double)ld<S>(A + loc_in<PE>(L, x, y, z, i), inv);
        r += g;
        if (cxq(i)) jx += cxq(i) * g;
        if (cyq(i)) jy += cyq(i) * g;
        if (czq(i)) jz += czq(i) * g;
      }
      v[0] = 1.0 + r; v[1] = jx; v[2] = jy; v[3] = jz;
    }
  }
  block_sum4(v, red);
  if (threadIdx.x == 0)
    for (int c = 0; c < 4; ++c) part[(long long)blockIdx.x * 4 + c] = v[c];
}

// Both levels' integral partials folded in a fixed order into dst[4] (common units: mass
// sum rho dx^3, momentum sum j dx^3 dx/dt; coarse level first).  One wave.  With acc, the
// fold also records the step it brackets.
__global__ void k_integrals_final(const double *__restrict__ part, int nb_c, int nb_f, double vol_c, double vol_f,
                                  double uv, double *__restrict__ dst, BudgetAcc *acc, long long step) {
  __shared__ double lane_sum[2][4][64];
  const int lane = threadIdx.x;
  for (int lev = 0; lev < 2; ++lev) {
    const int first = lev ? nb_c : 0, n = lev ? nb_f : nb_c;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = lane; b < n; b += 64)
      for (int c = 0; c < 4; ++c) s[c] += part[(long long)(first + b) * 4 + c];
    for (int c = 0; c < 4; ++c) lane_sum[lev][c][lane] = s[c];
  }
  __syncthreads();
  if (lane == 0) {
    double tot[2][4];
    for (int lev = 0; lev < 2; ++lev)
      for (int c = 0; c < 4; ++c) {
        double s = 0.0;
        for (int k = 0; k < 64; ++k) s += lane_sum[lev][c][k];
        tot[lev][c] = s;
      }
    dst[0] = tot[0][0] * vol_c + tot[1][0] * vol_f;
    for (int c = 1; c < 4; ++c) dst[c] = tot[0][c] * vol_c * uv + tot[1][c] * vol_f * uv;
    if (acc) acc->budget_step = step;
  }
}

// Budget accumulators back to the initial state (the integrals of the initial state are
// written by k_integrals_final into acc->initial afterwards).
__global__ void k_budget_reset(BudgetAcc *acc) {
  if (threadIdx.x == 0) {
    double *p = (double *)acc;
    for (int k = 0; k < (int)(offsetof(BudgetAcc, steps) / sizeof(double)); ++k) p[k] = 0.0;
    acc->steps = 0;
    acc->budget_step = 0;
  }
}

// Equilibrium reflux on the coarse cells just outside each face: one wavefront per
// cell, lane i (< 27) owns population i; density and momentum by fixed-order butterflies.
template <typename R>
__device__ __forceinline__ R eq_dir(int i, R rho, R ux, R uy, R uz) {
  const R u[3] = {ux, uy, uz};
  const int c[3] = {cxq(i), cyq(i), czq(i)};
  R e = rho;
  for (int a = 0; a < 3; ++a)
    e *= c[a] == 0 ? R(2.0 / 3.0) - u[a] * u[a] : (R(1.0 / 3.0) + u[a] * u[a] + R(c[a]) * u[a]) * R(0.5);
  return e;
}
template <typename R>
__device__ __forceinline__ R wave_sum(R v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int S, bool PE>
// With M, the wave also writes the donor moments of its cell's new state for the next coarse
// step (its populations as stored, gathered to lane 0; every donor-block z index that maps
// to its z).
__device__ __forceinline__ void reflux_body(typename St<S>::T *__restrict__ A, Desc L, BoxGeom B, const TR<S> *__restrict__ rc, const TR<S> *__restrict__ rf, TR<S> inv_vol, float scale, float inv, int *__restrict__ err, long long t, TR<S> *__restrict__ M = nullptr) {
  using R = TR<S>;
  // rf holds fine face elements [f][2 nt][2 nz][4]
  const int nt = B.nb[0] > B.nb[1] ? B.nb[0] : B.nb[1];
  const long long w = t / 64;
  const int lane = threadIdx.x % 64;
  if (w >= 4LL * nt * L.nz) return;
  const int zpos = (int)(w % L.nz), ti = (int)((w / L.nz) % nt), f = (int)(w / ((long long)L.nz * nt));
  const int axis = f / 2, side = f % 2;
  const int other = 1 - axis;
  if (ti >= B.nb[other]) return;  // uniform across the wave
  const int a = side ? B.hi[axis] : B.lo[axis] - 1;
  const int x = axis == 0 ? a : B.lo[other] + ti, y = axis == 0 ? B.lo[other] + ti : a;
  const long long ri = (((long long)f * nt + ti) * L.nz + zpos) * 4;
  const R sgn = side ? R(1) : R(-1);
  R corr[4];
  for (int c = 0; c < 4; ++c) {
    R ff = R(0);
    for (int p = 0; p < 2; ++p)
      for (int q = 0; q < 2; ++q) ff += rf[(((long long)f * 2 * nt + 2 * ti + p) * 2 * L.nz + 2 * zpos + q) * 4 + c];
    corr[c] = sgn * (ff - rc[ri + c]) * inv_vol;
  }
  const int i = lane < NQ ? lane : 0;
  const long long slot = loc_in<PE>(L, x, y, zpos, i);
  R g = lane < NQ ? R(ld<S>(A + slot, inv)) : R(0);
  const R on = lane < NQ ? R(1) : R(0);
  const R d = wave_sum(g);
  const R jx = wave_sum(R(cxq(i)) * g * on);
  const R jy = wave_sum(R(cyq(i)) * g * on);
  const R jz = wave_sum(R(czq(i)) * g * on);
  int e = 0;
  const R rho = d + R(1), rho2 = rho + corr[0];
  if (!(rho > 0) || !(rho2 > 0)) e |= TE_DENSITY;
  const R ux = jx / rho, uy = jy / rho, uz = jz / rho;
  const R vx = (jx + corr[1]) / rho2, vy = (jy + corr[2]) / rho2, vz = (jz + corr[3]) / rho2;
  const R mach_cap = R(0.3 * 0.57735026918962576451);
  if (sqrt(ux * ux + uy * uy + uz * uz) > mach_cap || sqrt(vx * vx + vy * vy + vz * vz) > mach_cap) e |= TE_MACH;
  const R eo = eq_dir(i, rho, ux, uy, uz), en = eq_dir(i, rho2, vx, vy, vz);
  if (lane < NQ) {
    if (!(eo >= 0) || !(en >= 0)) e |= TE_NEGATIVE;
    g += en - eo;
    if (!(g + wq(i) >= 0)) e |= isfinite(g) ? TE_NEGATIVE : TE_NONFINITE;
    stv<S>(A + slot, (typename St<S>::C)g, scale);
  }
  if (M != nullptr) {
    const R h = lane < NQ ? stored_value<S>(g, scale, inv) : R(0);
    MomentSum<R> ms;
#pragma unroll
    for (int q = 0; q < NQ; ++q) ms.add(q, __shfl(h, q, 64));
    if (lane == 0) {
      R md[10];
      ms.finish(md, e);
      const int dx = B.nb[0] + 6, dy = B.nb[1] + 6, dz = B.nb[2] + 6;
      const int bx = x - B.lo[0], by = y - B.lo[1];
      for (int k = ((zpos - B.lo[2] + 3) % L.nz + L.nz) % L.nz; k < dz; k += L.nz) {
        R *o = M + ((((long long)k * dy) + (by + 3)) * dx + (bx + 3)) * 10;
        for (int c = 0; c < 10; ++c) o[c] = md[c];
      }
    }
  }
  if (e) atomicOr(err, e);
}

// Fine moments and restriction in one pass: a block takes a tile of RT_X x RT_Y covered
// coarse cells of one coarse z layer; its 256 threads are their children, whose moments go
// to LDS; then one thread per coarse cell averages its eight children (fixed order) as in
// the reference's restriction (covered cells: free stream in the virtual core, otherwise the
// 2x2x2 mean of the children's moments, rescaled fine->coarse and regularised); the moments
// never leave the chip.
constexpr int RT_X = 16, RT_Y = 2;
constexpr int RESTRICT_LDS = 10 * 256;  // doubles: room for the children's moments of one tile
// shared memory of a kernel that fuses the restriction and the loads + budget fold (the
// wall partials' lane sums, then four lane-sum columns of the outer-face flux partials)
constexpr int FOLD_LDS = LOADS_LDS + 4 * 64;
constexpr int END_LDS = RESTRICT_LDS > FOLD_LDS ? RESTRICT_LDS : FOLD_LDS;
__host__ __device__ inline int restrict_tiles(const BoxGeom &B) {
  return ((B.nb[0] + RT_X - 1) / RT_X) * ((B.nb[1] + RT_Y - 1) / RT_Y) * B.nb[2];
}

// With M, each restricted cell outside the core also writes its donor moments for the next
// coarse step (the moments of its populations as stored, at every donor-block z index
// that maps to its z), so that no separate donor pass reads them back.
template <int S, bool PE, bool PEC>
__device__ __forceinline__ void restrict_tile_body(const typename St<S>::T *__restrict__ Af, Desc Lf, float finv,
                                                   typename St<S>::T *__restrict__ Ac, Desc Lc, BoxGeom B, Levels V,
                                                   FreeState F, float cscale, int *__restrict__ err, int tile,
                                                   TR<S> *__restrict__ sm, TR<S> *__restrict__ M = nullptr,
                                                   float cinv = 1.0f) {
  using R = TR<S>;
  const int ntx = (B.nb[0] + RT_X - 1) / RT_X, nty = (B.nb[1] + RT_Y - 1) / RT_Y;
  const int tbx = tile % ntx, tby = (tile / ntx) % nty, bz = tile / (ntx * nty);
  const int tid = threadIdx.x;
  const int cx = tid % (2 * RT_X), cy = (tid / (2 * RT_X)) % (2 * RT_Y), cz = tid / (4 * RT_X * RT_Y);
  const int fx = 2 * RT_X * tbx + cx, fy = 2 * RT_Y * tby + cy, fz = 2 * bz + cz;
  int e = 0;
  R m[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (fx < 2 * B.nb[0] && fy < 2 * B.nb[1] && !in_core(B, fx / 2, fy / 2)) {
    MomentSum<R> ms;
#pragma unroll
    for (int i = 0; i < NQ; ++i) ms.add(i, R(ld<S>(Af + loc_in<PE>(Lf, B.pad + fx, B.pad + fy, fz, i), finv)));
    ms.finish(m, e);
  }
#pragma unroll
  for (int k = 0; k < 10; ++k) sm[k * 256 + tid] = m[k];
  __syncthreads();
  if (tid < RT_X * RT_Y) {
    const int ox = tid % RT_X, oy = tid / RT_X;
    const int bx = RT_X * tbx + ox, by = RT_Y * tby + oy;
    if (bx < B.nb[0] && by < B.nb[1]) {
      const int X = B.lo[0] + bx, Y = B.lo[1] + by, Z = B.lo[2] + bz;
      if (in_core(B, bx, by)) {
#pragma unroll
        for (int i = 0; i < NQ; ++i) stv<S>(Ac + loc_in<PEC>(Lc, X, Y, Z, i), (typename St<S>::C)F.g[i], cscale);
      } else {
        R acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int c = 0; c < 2; ++c)
          for (int b = 0; b < 2; ++b)
            for (int a = 0; a < 2; ++a) {
              const int ch = (2 * ox + a) + 2 * RT_X * ((2 * oy + b) + 2 * RT_Y * c);
              for (int k = 0; k < 10; ++k) acc[k] += sm[k * 256 + ch];
            }
        for (int k = 0; k < 10; ++k) acc[k] *= R(0.125);
        rescale(acc, R(V.shear_fc), R(V.bulk_fc));
        // with M the stored values also feed the donor moments of the next step.  FP32/FP16
        // store each population as it is made (holding all 27 beside their rounded values
        // costs ~45 registers there); FP64 stores exactly what it computes and measured
        // faster with the 27 held
        const PopBasis<R> pb(acc, e);
        MomentSum<R> ms;
        if constexpr (S == 0) {
          R g[NQ];
#pragma unroll
          for (int i = 0; i < NQ; ++i) g[i] = pb.pop(i, acc, e);
          scatter_in<S, PEC>(Ac, Lc, X, Y, Z, cscale, g);
          if (M != nullptr) {
#pragma unroll
            for (int i = 0; i < NQ; ++i) ms.add(i, g[i]);
          }
        } else {
#pragma unroll
          for (int i = 0; i < NQ; ++i) {
            const R gi = pb.pop(i, acc, e);
            stv<S>(Ac + loc_in<PEC>(Lc, X, Y, Z, i), (typename St<S>::C)gi, cscale);
            if (M != nullptr) ms.add(i, stored_value<S>(gi, cscale, cinv));
          }
        }
        if (M != nullptr) {
          R md[10];
          ms.finish(md, e);
          const int dx = B.nb[0] + 6, dy = B.nb[1] + 6, dz = B.nb[2] + 6;
          for (int k = ((bz + 3) % Lc.nz + Lc.nz) % Lc.nz; k < dz; k += Lc.nz) {
            R *o = M + ((((long long)k * dy) + (by + 3)) * dx + (bx + 3)) * 10;
            for (int c = 0; c < 10; ++c) o[c] = md[c];
          }
        }
      }
    }
  }
  if (e) atomicOr(err, e);
}

// End of a coarse step: the restriction tiles, then (block range after `ntiles`) the donor
// moments of the next step's pre-collision state for the donor cells it does not write
// (outside the box, and the core's free-stream constants).  The two ranges touch disjoint
// slots: every slot is the incoming population of exactly one cell.
template <int S, bool PE, bool PEC>
__global__ __launch_bounds__(256) void k_restrict_tiles(const typename St<S>::T *__restrict__ Af, Desc Lf, float finv,
                                                        typename St<S>::T *__restrict__ Ac, Desc Lc, BoxGeom B,
                                                        Levels V, FreeState F, float cscale, float cinv,
                                                        TR<S> *__restrict__ M, int *__restrict__ err, int ntiles) {
  __shared__ double sm[RESTRICT_LDS];
  if ((int)blockIdx.x < ntiles)
    restrict_tile_body<S, PE, PEC>(Af, Lf, finv, Ac, Lc, B, V, F, cscale, err, (int)blockIdx.x, (TR<S> *)sm, M, cinv);
  else
    donor_moments_body<S, PEC, 1>(Ac, Lc, B, F, cinv, M, err, (long long)((int)blockIdx.x - ntiles) * 256 + threadIdx.x);
}

// End of a coarse step in one launch over consecutive block ranges: the reflux waves
// (coarse cells next to the box faces), one block for the loads (sampled step) or the step
// tick, the restriction tiles (covered cells), and the next step's donor moments of the
// remaining donor cells.  Reflux and restriction write their own cells' donor moments, so
// every range reads and writes slots no other range touches.
// Where the end of a step folds the budget: the outer-face flux partials of this step
// (written during the first fine substep) and the accumulators; acc == nullptr skips it.
struct BudgetArgs {
  const double *opart;
  int nb_outer;
  double volc, uv;  // coarse cell volume, common velocity unit dx/dt
  BudgetAcc *acc;
};

// The loads block of a step's end with the budget (one block, 256 threads): lanes 0..63 sum
// the wall partials, lanes 64..127 the outer-face flux partials (lane-strided, fixed order);
// thread 0 then writes the history row of a sampled step and adds this step's wall mass
// exchange, wall impulse and outer in-flux to the accumulators (refined_solver.py:318-343),
// and advances the step cursor.
__device__ __forceinline__ void refined_fold_body(const double *__restrict__ part, int nparts, int nelem,
                                                  double fscale, double volf, const BudgetArgs &G, StepCtl *ctl,
                                                  double *__restrict__ hist, int sample, double *__restrict__ lds) {
  const int lane = threadIdx.x;
  double *oacc = lds + LOADS_LDS;
  if (lane < 64) loads_lane_sums(part, nparts, nelem, lds, lane);
  if (lane >= 64 && lane < 128) {
    const int l = lane - 64;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = l; b < G.nb_outer; b += 64)
      for (int c = 0; c < 4; ++c) s[c] += G.opart[(long long)b * 4 + c];
    for (int c = 0; c < 4; ++c) oacc[c * 64 + l] = s[c];
  }
  __syncthreads();
  if (lane != 0) return;
  if (sample) loads_row(lds, nelem, fscale, fscale, volf, hist + (long long)ctl->row * ctl->width, ctl->step);
  // the same lane sums in the same order as loads_row
  double tot[5] = {0, 0, 0, 0, 0};
  for (int e = 0; e < nelem; ++e)
    for (int c = 0; c < 5; ++c) {
      double s = 0.0;
      for (int k = 0; k < 64; ++k) s += lds[(e * 5 + c) * 64 + k];
      tot[c] += s;
    }
  double ext[4];
  for (int c = 0; c < 4; ++c) {
    double s = 0.0;
    for (int k = 0; k < 64; ++k) s += oacc[c * 64 + k];
    ext[c] = s;
  }
  BudgetAcc *a = G.acc;
  for (int c = 0; c < 4; ++c) a->wall[c] = tot[c] * volf * G.uv;
  a->wall[4] = tot[4] * volf;
  a->ext[0] = ext[0] * G.volc;
  for (int c = 1; c < 4; ++c) a->ext[c] = ext[c] * G.volc * G.uv;
  a->cum_wall_mass += a->wall[4];
  for (int c = 0; c < 3; ++c) a->cum_wall_impulse[c] += a->wall[c];
  for (int c = 0; c < 4; ++c) a->cum_ext[c] += a->ext[c];
  a->steps += 1;
  ctl->step += 1;
  if (sample) ctl->row += 1;
}

template <int S, bool PF, bool PC>
__global__ __launch_bounds__(256) void k_step_end(const typename St<S>::T *__restrict__ Af, Desc Lf, float finv,
                                                  typename St<S>::T *__restrict__ Ac, Desc Lc, BoxGeom B, Levels V,
                                                  FreeState F, float cscale, float cinv, TR<S> *__restrict__ M,
                                                  const TR<S> *__restrict__ rc, const TR<S> *__restrict__ rf,
                                                  TR<S> inv_vol, int *__restrict__ err, int nb_reflux, int ntiles,
                                                  const double *__restrict__ part, int nparts, int nelem, double fscale,
                                                  double vscale, StepCtl *ctl, double *__restrict__ hist, int sample,
                                                  BudgetArgs G) {
  __shared__ double lds[END_LDS];
  const int b = (int)blockIdx.x;
  if (b < nb_reflux) {
    reflux_body<S, PC>(Ac, Lc, B, rc, rf, inv_vol, cscale, cinv, err, (long long)b * 256 + threadIdx.x, M);
  } else if (b == nb_reflux) {
    if (G.acc) refined_fold_body(part, nparts, nelem, fscale, vscale, G, ctl, hist, sample, lds);
    else if (sample) loads_final_body(part, nparts, nelem, fscale, fscale, vscale, ctl, hist, nullptr, lds);
    else if (threadIdx.x == 0) ctl->step += 1;
  } else if (b < nb_reflux + 1 + ntiles) {
    restrict_tile_body<S, PF, PC>(Af, Lf, finv, Ac, Lc, B, V, F, cscale, err, b - nb_reflux - 1, (TR<S> *)lds, M, cinv);
  } else {
    donor_moments_body<S, PC, 2>(Ac, Lc, B, F, cinv, M, err, (long long)(b - nb_reflux - 1 - ntiles) * 256 + threadIdx.x);
  }
}

// Initial pre-collision state of one level: rho = 1, seeded momentum, regularised
// product-form populations (no stress), written into the gather slots of a
// lattice whose next sweep is even (streamed form).
struct LevelSeed {
  double U, amp, phy, phz, ky, kz, lex, ley;  // seeded field (common frame)
  double ox, oy, oz, dx;                      // cell (i,j,k) centre = ((i+0.5) dx + o)
  double px, py;                              // plan origin added to x, y
};

template <int S>
__global__ void k_init_level(typename St<S>::T *__restrict__ A, const uint8_t *__restrict__ mask, Desc L, LevelSeed p,
                             float scale, int *__restrict__ err) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)L.nx * L.ny * L.nz) return;
  const int x = (int)(t % L.nx), y = (int)((t / L.nx) % L.ny), z = (int)(t / ((long long)L.nx * L.ny));
  const double xx = ((double)x + 0.5) * p.dx + p.ox + p.px, yy = ((double)y + 0.5) * p.dx + p.oy + p.py;
  const double zz = ((double)z + 0.5) * p.dx + p.oz;
  const double rx = xx - p.lex, ry = yy - p.ley;
  const double e = (rx - 0.5) / 0.75;
  const double env = exp(-e * e);
  double m[10] = {1.0, p.U, 0, 0, 0, 0, 0, 0, 0, 0};
  m[2] = p.amp * sin(p.ky * ry + p.phy) * cos(p.kz * zz + p.phz) * env;
  m[3] = -p.amp * p.ky / p.kz * cos(p.ky * ry + p.phy) * sin(p.kz * zz + p.phz) * env;
  if (mask[(long long)(y + 1) * (L.nx + 2) + (x + 1)] == 1) m[1] = m[2] = m[3] = 0.0;
  double g[NQ];
  int er = 0;
  populations_of(m, g, er);
  scatter_in<S, false>(A, L, x, y, z, scale, g);
  if (er) atomicOr(err, er);
}

// ------------------------------------------------------------------ fused phases
// After the coarse sweep, the coarse face flux (cells next to the box faces), the coarse
// farfield (ghost slots of the domain edge and the edge cells' own slots) and the donor
// moments of the new state (box +- 3, inside the domain) read and write disjoint slots, so
// they run as one launch over consecutive block ranges: their latencies overlap.
template <int S, bool PE>
__global__ __launch_bounds__(256) void k_after_coarse(typename St<S>::T *__restrict__ A, Desc L, BoxGeom B,
                                                      const FluxTable *__restrict__ T, TR<S> volc, float scale,
                                                      float inv, TR<S> *__restrict__ regc, LegacyLinks K, double U,
                                                      FreeState F, TR<S> *__restrict__ after, int *__restrict__ err,
                                                      int nb_flux, int nb_leg) {
  const int b = blockIdx.x;
  if (b < nb_flux) {
    flux_coarse_body<S, PE>(A, L, B, T, volc, inv, regc, (long long)b * 256 + threadIdx.x);
  } else if (b < nb_flux + nb_leg) {
    legacy_body<S, PE>(A, L, K, U, scale, inv, (long long)(b - nb_flux) * 256 + threadIdx.x);
  } else {
    donor_moments_body<S, PE>(A, L, B, F, inv, after, err, (long long)(b - nb_flux - nb_leg) * 256 + threadIdx.x);
  }
}

// After a fine sweep, the fine face flux (box faces) and the walls with their loads
// (around the body, which keeps clear of the faces) likewise share one launch.  The wall
// partials keep their per-block layout (block index within the wall range).  After the first
// substep a third range adds the coarse domain's outer-face flux of this coarse step (the
// coarse lattice is untouched between its farfield and the end of the step), with PC the
// coarse lattice's form.
template <int S, bool PE, bool PC>
__global__ __launch_bounds__(256) void k_after_fine(typename St<S>::T *__restrict__ A, Desc L, BoxGeom B,
                                                    const FluxTable *__restrict__ T, TR<S> volf, float scale, float inv,
                                                    int sub, TR<S> *__restrict__ regf, WallLinks W,
                                                    double *__restrict__ part, int nb_flux, int nb_walls,
                                                    const typename St<S>::T *__restrict__ Ac, Desc Lc, float cinv,
                                                    FreeState F, double *__restrict__ opart) {
  const int b = blockIdx.x;
  if (b < nb_flux) {
    flux_fine_body<S, PE>(A, L, B, T, volf, inv, sub, regf, (long long)b * 256 + threadIdx.x);
  } else if (b < nb_flux + nb_walls) {
    const int wb = b - nb_flux;
    walls_body<S, PE, true>(A, L, W, scale, inv, part, (long long)wb * 256 + threadIdx.x, wb);
  } else {
    const int ob = b - nb_flux - nb_walls;
    outer_flux_body<S, PC>(Ac, Lc, T, F, cinv, opart, (long long)ob * 256 + threadIdx.x, ob);
  }
}

}  // namespace ak
