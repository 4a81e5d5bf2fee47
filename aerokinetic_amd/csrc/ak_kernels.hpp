// ak_kernels.hpp -- gfx950 kernels of one lattice step (sweep, walls, farfield, reductions).
#pragma once
#include "../../include/aerokinetic.h"
#include "ak_device.hpp"

namespace ak {

// ------------------------------------------------------------------ sweep
// One thread per interior cell; blocks tile (x, y) of one z plane, or (row-linear mode)
// take bx consecutive cells of a plane in (y, x) order, so that a row width that is not a
// multiple of 64 leaves idle lanes only in a plane's last block.  The 1-D grid is remapped
// so that each XCD (blocks are dealt round-robin over 8 XCDs) owns a contiguous slab of
// tiles and the neighbour rows it gathers stay in its own L2.
struct SweepGeom {
  int bx, by;      // block tile (threads) in x, y; row-linear: bx threads, by = 1
  int gx, gy, gz;  // tiles per axis; row-linear: gx blocks per plane, gy = 1
  int nblocks;
  int lin;         // 1: row-linear mode
  // multiply-shift constants of the divisions by gx, gx*gy, bx and nx (magic_of)
  unsigned m_gx, m_gxy, m_bx, m_nx;
  int s_gx, s_gxy, s_bx, s_nx;
};

// Division by an invariant d in [1, 2^31) of any n in [0, 2^31) as one multiply and a shift:
// with l = ceil(log2 d), m = floor(2^(31+l) / d) + 1 (< 2^32) gives 2^(31+l) < m d <= 2^(31+l) + 2^l,
// so floor(n m / 2^(31+l)) = floor(n / d) for every n < 2^31 (Granlund and Montgomery, 1994).
__host__ __device__ inline void magic_of(unsigned d, unsigned &m, int &shift) {
  int l = 0;
  while ((1ull << l) < d) ++l;
  m = (unsigned)((1ull << (31 + l)) / d + 1);
  shift = 31 + l;
}
__host__ __device__ __forceinline__ unsigned magic_div(unsigned n, unsigned m, int shift) {
  return (unsigned)(((unsigned long long)n * m) >> shift);
}

// Order-independent maxima/minima as bit patterns of non-negative doubles.  Blocks
// update one of DIAG_SLOTS slots (block index modulo the slot count) so that atomics
// from hundreds of thousands of blocks never pile onto one address.
constexpr int DIAG_SLOTS = 1024;
struct DiagSlot {
  unsigned long long u2max, drmax, rhomin, nonfinite;
};
struct DiagAcc {
  DiagSlot s[DIAG_SLOTS];
};

__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int per = nb / 8, rem = nb % 8, xcd = b % 8, k = b / 8;
  return xcd < rem ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
}

// Wave maximum of finite non-negative values (callers replace non-finite inputs first): for
// FP32 an unsigned maximum of the bit patterns (the same order), one instruction per step.
template <typename C>
__device__ __forceinline__ C wave_max(C v) {
  if constexpr (sizeof(C) == 4) {
    unsigned b = __float_as_uint(v);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) b = max(b, (unsigned)__shfl_xor((int)b, o, 64));
    return __uint_as_float(b);
  } else {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
  }
}
template <typename C>
__device__ __forceinline__ C wave_min(C v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) { C t = __shfl_xor(v, o, 64); v = t < v ? t : v; }
  return v;
}

// Order-preserving keys of floating-point values: k(a) < k(b) exactly when a < b (finite
// values and infinities); the complement of a key is the key of the negated value.
__device__ __forceinline__ unsigned ord_key(float v) {
  const unsigned b = __float_as_uint(v);
  return b ^ ((unsigned)((int)b >> 31) | 0x80000000u);
}
__device__ __forceinline__ unsigned long long ord_key(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return b ^ ((unsigned long long)((long long)b >> 63) | 0x8000000000000000ull);
}
__device__ __forceinline__ float key_value(unsigned k) {
  return __uint_as_float(k & 0x80000000u ? k ^ 0x80000000u : ~k);
}
__device__ __forceinline__ double key_value(unsigned long long k) {
  return __longlong_as_double((long long)(k & 0x8000000000000000ull ? k ^ 0x8000000000000000ull : ~k));
}
__device__ __forceinline__ unsigned key_bits(float v) { return __float_as_uint(v); }
__device__ __forceinline__ unsigned long long key_bits(double v) { return (unsigned long long)__double_as_longlong(v); }

// Wave maximum of unsigned keys through DPP row operations (all 64 lanes must be active):
// two quad permutations and two row rotations leave every lane of a row of 16 with the row's
// maximum, and the four row maxima are combined from lanes 0, 16, 32 and 48 in scalar registers.
template <int CTL>
__device__ __forceinline__ unsigned dpp_u32(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTL, 0xF, 0xF, false);
}
template <int CTL>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long v) {
  const unsigned lo = dpp_u32<CTL>((unsigned)v), hi = dpp_u32<CTL>((unsigned)(v >> 32));
  return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ unsigned readlane_key(unsigned v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ unsigned long long readlane_key(unsigned long long v, int lane) {
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)v, lane), hi = __builtin_amdgcn_readlane((unsigned)(v >> 32), lane);
  return ((unsigned long long)hi << 32) | lo;
}
template <typename K>
__device__ __forceinline__ K wave_max_key(K v) {
  auto mx = [](K a, K b) { return a > b ? a : b; };
  if constexpr (sizeof(K) == 4) {
    v = mx(v, dpp_u32<0xB1>(v));   // quad_perm [1,0,3,2]
    v = mx(v, dpp_u32<0x4E>(v));   // quad_perm [2,3,0,1]
    v = mx(v, dpp_u32<0x124>(v));  // row_ror:4
    v = mx(v, dpp_u32<0x128>(v));  // row_ror:8
  } else {
    v = mx(v, dpp_u64<0xB1>(v));
    v = mx(v, dpp_u64<0x4E>(v));
    v = mx(v, dpp_u64<0x124>(v));
    v = mx(v, dpp_u64<0x128>(v));
  }
  return mx(mx(readlane_key(v, 0), readlane_key(v, 16)), mx(readlane_key(v, 32), readlane_key(v, 48)));
}

// Fused diagnostics of one wave of the sweep (every lane active): largest |u|^2 and |drho|,
// smallest rho and the non-finite flag over the lanes with `counted`, folded by lane 0 into
// diagnostics slot `slot` (a slot is only ever touched by blocks of one XCD: slot = block
// index mod 1024, a multiple of the 8 XCDs).  Each extreme is reduced as an order-preserving
// integer key; the results equal maxima and minima over the cell values themselves (the
// smallest rho is 1 + the smallest drho, since rounding of 1 + x is monotone in x).
template <typename C>
__device__ __forceinline__ void diag_wave(bool counted, C drho, C u2, DiagAcc *__restrict__ diag, int slot) {
  using K = decltype(ord_key(C(0)));
  const bool fu = isfinite(u2), fd = isfinite(drho);
  const K kd = ord_key(drho);
  const K ku = wave_max_key<K>(counted && fu ? key_bits(u2) : K(0));  // u2 >= +0: its bits are a key
  const K khi = wave_max_key<K>(counted && fd ? kd : K(0));           // key of the largest drho
  const K klo = wave_max_key<K>(counted && fd ? (K)~kd : K(0));       // complement of the smallest drho's key
  const bool bad = __any(counted && !(fu && fd));
  if ((threadIdx.x & 63) != 0) return;
  DiagSlot *sl = &diag->s[slot];
  if (ku != K(0)) {
    C u2max;
    if constexpr (sizeof(K) == 4) u2max = __uint_as_float(ku);
    else u2max = __longlong_as_double((long long)ku);
    atomicMax(&sl->u2max, (unsigned long long)__double_as_longlong((double)u2max));
  }
  if (khi != K(0)) {  // some lane had a counted finite drho (finite keys are never 0)
    const C dmax = key_value(khi), dmin = key_value((K)~klo);
    const C amax = fabs(dmax), amin = fabs(dmin);
    atomicMax(&sl->drmax, (unsigned long long)__double_as_longlong((double)(amin > amax ? amin : amax)));
    const C rmin = C(1) + dmin;
    if (rmin > C(0)) atomicMin(&sl->rhomin, (unsigned long long)__double_as_longlong((double)rmin));
    else atomicMax(&sl->nonfinite, 2ull);  // non-positive density
  }
  if (bad) atomicMax(&sl->nonfinite, 1ull);
}

// Keep a wave-uniform pointer in scalar registers (the compiler otherwise may carry the
// 27 direction bases of the in-place sweep as 64-bit vector values).
typedef __attribute__((address_space(1))) char gchar;
__device__ __forceinline__ gchar *uniform_ptr(const void *p) {
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (gchar *)(((unsigned long long)hi << 32) | lo);
}
// A wave-uniform 64-bit value in scalar registers, opaque to the optimiser (an empty asm
// that "modifies" it), so that values derived from it are recomputed where they are used
// instead of being kept alive from an earlier identical computation.
__device__ __forceinline__ long long fresh_sgpr64(long long v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)((unsigned long long)v >> 32));
  unsigned long long r = ((unsigned long long)hi << 32) | lo;
  asm volatile("" : "+s"(r));
  return (long long)r;
}
template <int S>
__device__ __forceinline__ typename St<S>::C ld_g(const gchar *p, float inv_scale) {
  typedef __attribute__((address_space(1))) const typename St<S>::T gT;
  if constexpr (S == 2) return (float)(*(const gT *)p) * inv_scale;
  else return *(const gT *)p;
}
template <int S>
__device__ __forceinline__ void st_g(gchar *p, typename St<S>::C v, float scale) {
  typedef __attribute__((address_space(1))) typename St<S>::T gT;
  if constexpr (S == 2) *(gT *)p = (_Float16)(v * scale);
  else *(gT *)p = v;
}

// Density deviation and velocity of one cell from its f_out(x, j), j = 0..26 in order
// (`load(j)`), so that every reader of a cell's moments gets bit-identical values.
template <typename C, class Load>
__device__ __forceinline__ void macro_of(Load &&load, C &drho, C (&u)[3]) {
  C r = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
  for (int j = 0; j < NQ; ++j) {
    const C v = load(j);
    r += v;
    if (cxq(j)) jx += cxq(j) * v;
    if (cyq(j)) jy += cyq(j) * v;
    if (czq(j)) jz += czq(j) * v;
  }
  drho = r;
  const C rho = C(1) + r;
  u[0] = jx / rho; u[1] = jy / rho; u[2] = jz / rho;
}

template <int S, bool PE>
__device__ __forceinline__ void macro_cell(const typename St<S>::T *A, const Desc &L, int x, int y, int z,
                                           float inv_scale, typename St<S>::C &drho, typename St<S>::C (&u)[3]) {
  macro_of([&](int j) { return ld<S>(A + loc_out<PE>(L, x, y, z, j), inv_scale); }, drho, u);
}

// Perimeter moments of the characteristic farfield: the density deviation and velocity
// (macro_of of the cell's f_out, in the collision's compute type) of every cell on the x
// faces (x = 0 and nx - 1: [z][side][c][y]) and on the y faces (y = 0 and ny - 1:
// [z][side][c][x] after the x faces), c = drho, ux, uy, uz; side = 0 for the low face.  The
// sweep writes them for every cell it updates, and the farfield reads four coalesced values
// per cell instead of 27 populations, which on an x face lie in 27 different 128-byte lines
// that no other lane of a wave uses.
__device__ __forceinline__ long long pm_x(const Desc &L, int side, int y, int z) {
  return (long long)(z * 2 + side) * 4 * L.ny + y;  // component stride L.ny
}
__device__ __forceinline__ long long pm_y(const Desc &L, int side, int x, int z) {
  return 8LL * L.ny * L.nz + (long long)(z * 2 + side) * 4 * L.nx + x;  // component stride L.nx
}
__host__ __device__ inline size_t pm_count(int nx, int ny, int nz) { return 8 * ((size_t)ny * nz + (size_t)nx * nz); }

// Per-cell copies of nine populations of each x-face cell, at [z][side][y][j] (a cell's nine
// values together, consecutive cells along y): the outgoing ones (for x = 0 the nine
// f_out(0, y, z, q) with c_x = -1, q = 0..8, j = q; for x = nx - 1 those with c_x = +1,
// q = 18..26, j = q - 18), in the storage type and bit for bit as the sweep stored them, and
// the entering ones (c_x = +1 at x = 0, j = q - 18; c_x = -1 at x = nx - 1, j = q) that the
// farfield computes for the next sweep.  The farfield's lanes run along y, so both copies
// replace nine lattice rows per face cell (one 128-byte line per value) by a few lines per
// wave, and the sweep's one x-face lane per wave reads or writes one or two lines.
__device__ __forceinline__ long long xo_at(const Desc &L, int side, int j, int y, int z) {
  return ((long long)(z * 2 + side) * L.ny + y) * 9 + j;
}
__host__ __device__ inline size_t xo_count(int ny, int nz) { return (size_t)18 * ny * nz; }

// Writes the perimeter moments of cell (x, y, z).
template <typename C>
__device__ __forceinline__ void perimeter_put(C *__restrict__ pm, const Desc &L, int x, int y, int z, C dr,
                                              const C (&u)[3]) {
  if (x == 0 || x == L.nx - 1) {
    C *p = pm + pm_x(L, x != 0, y, z);
    p[0] = dr; p[L.ny] = u[0]; p[2 * L.ny] = u[1]; p[3 * L.ny] = u[2];
  }
  if (y == 0 || y == L.ny - 1) {
    C *p = pm + pm_y(L, y != 0, x, z);
    p[0] = dr; p[L.nx] = u[0]; p[2 * L.nx] = u[1]; p[3 * L.nx] = u[2];
  }
}

// The 27 loaded populations as register values at this point (an empty asm that "modifies"
// them): the loads are issued, and waited for, ahead of the mask test that follows, instead
// of being moved behind it into the branch that uses them.
template <typename C>
__device__ __forceinline__ void loads_issued(C (&g)[NQ]) {
#pragma unroll
  for (int q = 0; q < NQ; ++q) asm volatile("" : "+v"(g[q]));
}

// One tile (block `lb` of the lattice's sweep grid) of the fused AA stream + collide.
// `slot` picks the diagnostics accumulator slot of this block; with `pm` the cells on the
// x and y faces also write their perimeter moments.
template <int S, bool ODD, bool PXY, bool DIAG>
__device__ __forceinline__ void sweep_block(typename St<S>::T *__restrict__ A, const uint8_t *__restrict__ solid,
                                            const Desc &L, const SweepGeom &G, typename St<S>::C ws,
                                            typename St<S>::C wb, float scale, float inv_scale,
                                            DiagAcc *__restrict__ diag, int lb, int slot,
                                            typename St<S>::C *__restrict__ pm = nullptr,
                                            typename St<S>::T *__restrict__ xo = nullptr,
                                            const typename St<S>::T *__restrict__ xi = nullptr) {
  using C = typename St<S>::C;
  // the block's tile coordinates are wave-uniform: kept in scalar registers, so that the
  // per-direction plane bases below are scalar arithmetic, not 64-bit vector adds
  // (divisions by multiply and shift: these index computations precede every load of the block)
  int x, y, z;
  const unsigned ub = (unsigned)__builtin_amdgcn_readfirstlane(lb);
  if (G.lin) {
    const unsigned zq = magic_div(ub, G.m_gx, G.s_gx);
    const int j0 = __builtin_amdgcn_readfirstlane((int)(ub - zq * (unsigned)G.gx) * G.bx);
    z = __builtin_amdgcn_readfirstlane((int)zq);
    const int j = j0 + (int)threadIdx.x;
    y = (int)magic_div((unsigned)j, G.m_nx, G.s_nx);
    x = j - y * L.nx;
  } else {
    const unsigned zq = magic_div(ub, G.m_gxy, G.s_gxy), r = ub - zq * (unsigned)(G.gx * G.gy);
    const unsigned tyq = magic_div(r, G.m_gx, G.s_gx);
    const int tx = __builtin_amdgcn_readfirstlane((int)(r - tyq * (unsigned)G.gx));
    const int ty = __builtin_amdgcn_readfirstlane((int)tyq);
    const unsigned ry = magic_div(threadIdx.x, G.m_bx, G.s_bx);
    x = tx * G.bx + (int)(threadIdx.x - ry * (unsigned)G.bx);
    y = ty * G.by + (int)ry;
    z = __builtin_amdgcn_readfirstlane((int)zq);
  }
  // mask: 0 fluid, 1 solid, 2 fluid outside the diagnostics, 3 exterior ring (its own kernel).
  // The mask byte is loaded beside the populations, not before them: every lane inside the
  // lattice gathers and collides, and only the cells the mask admits store (a solid or ring
  // cell's slots are read, never written, and no other cell writes them), so no wave waits a
  // memory round trip before issuing its 27 loads.
  const bool inside = x < L.nx && y < L.ny;
  uint8_t mk = 1;
  if (inside) mk = solid[(long long)(y + 1) * (L.nx + 2) + (x + 1)];
  bool active = false;
  C drho = 0, u2 = 0;
  if (inside) {
    C g[NQ];
    if constexpr (PXY) {
      const long long cell = L.base + cell_of(L, x, y, z);
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        long long src;
        if constexpr (ODD) src = cell + (long long)opp(q) * L.QS + nb_off<PXY>(L, x, y, z, -cxq(q), -cyq(q), -czq(q));
        else src = cell + (long long)q * L.QS;
        g[q] = ld<S>(A + src, inv_scale);
      }
      loads_issued(g);
      active = mk != 1 && mk != 3;
      if (active) {
        collide<C>(g, ws, wb, drho, u2);
        conserve_fp16<S>(g, scale, inv_scale);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          long long dst;
          if constexpr (ODD) dst = cell + (long long)q * L.QS + nb_off<PXY>(L, x, y, z, cxq(q), cyq(q), czq(q));
          else dst = cell + (long long)opp(q) * L.QS;
          stv<S>(A + dst, g[q], scale);
        }
      }
    } else {
      // A block covers one z plane, so every direction's shift (including the periodic
      // z wrap) is wave-uniform: a scalar base per direction plus one 32-bit offset.
      // byte offset of the cell inside a direction plane (32 bits): global loads then use a
      // scalar base plus this vector offset, with no 64-bit vector address arithmetic
      const unsigned boff = (unsigned)((y + 1) * L.P + x) * (unsigned)sizeof(typename St<S>::T);
      // the even form keeps the plane base in scalar registers explicitly (left to itself the
      // compiler keeps it, and the 27 direction bases derived from it, in vector registers)
      const long long plane = ODD ? L.base + (long long)z * L.PS : fresh_sgpr64(L.base + (long long)z * L.PS);
      const long long zp = (z == L.nz - 1) ? -(long long)(L.nz - 1) * L.PS : L.PS;
      const long long zm = (z == 0) ? (long long)(L.nz - 1) * L.PS : -L.PS;
      // FP16 storage: the loads keep the stored halves, which the collision's first pass adds in
      // mixed precision (collide_h); FP64 / FP32: the values themselves
      using RT = std::conditional_t<S == 2, _Float16, C>;
      typedef __attribute__((address_space(1))) const typename St<S>::T gT;
      RT r[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        long long d;
        if constexpr (ODD) {
          const int cz = -czq(q);
          d = (long long)opp(q) * L.QS - cxq(q) - (long long)cyq(q) * L.P + (cz == 1 ? zp : cz == -1 ? zm : 0);
        } else {
          d = (long long)q * L.QS;
        }
        const gchar *bq = uniform_ptr(A + (plane + d));
        r[q] = *(gT *)(bq + boff);
      }
      // x-face cells take their nine populations entering through that face from the farfield's
      // copy (the farfield no longer writes the lattice slots these lanes gathered above)
      // (the register index into r stays a compile-time constant, with one select per face:
      // a lane-dependent index becomes a compare-and-select over all 27 registers)
      if (xi != nullptr && __any(x == 0 || x == L.nx - 1)) {
        const bool lo = x == 0, hi = !lo && x == L.nx - 1;
        if (lo || hi) {
          RT v[9];
#pragma unroll
          for (int j = 0; j < 9; ++j) v[j] = ((gT *)xi)[xo_at(L, hi, j, y, z)];
#pragma unroll
          for (int j = 0; j < 9; ++j) {
            r[18 + j] = lo ? v[j] : r[18 + j];
            r[j] = hi ? v[j] : r[j];
          }
        }
      }
      loads_issued(r);
      active = mk != 1 && mk != 3;
      if (active) {
        C uc[3];
        // FP16: the collision returns the populations times the storage scale, and the mass- and
        // momentum-conserving rounding rounds each of them once
        if constexpr (S == 2) {
          collide_h(r, inv_scale, g, ws, wb, drho, u2, uc, scale);
        } else {
#pragma unroll
          for (int q = 0; q < NQ; ++q) g[q] = r[q];
          collide<C>(g, ws, wb, drho, u2, uc, C(1));
        }
        // cells on the x and y faces keep their density deviation and velocity (conserved by
        // the collision) for the next farfield pass: written before the stores, so that the four
        // values are not live beside the stores' (FP16: 50-51 instead of 56-57 VGPRs)
        if (pm != nullptr && (x == 0 || y == 0 || x == L.nx - 1 || y == L.ny - 1))
          perimeter_put<C>(pm, L, x, y, z, drho, uc);
        const long long splane = plane;
        // the 32-bit offset taken afresh in this block, so that each store, like each load, is a
        // scalar base plus this offset (otherwise the stores get 64-bit vector addresses, one
        // 64-bit add each, from the offset widened in the loads' block)
        unsigned soff = boff;
        asm volatile("" : "+v"(soff));
        auto dst = [&](int q) -> gchar * {
          long long d;
          if constexpr (ODD) {
            const int cz = czq(q);
            d = (long long)q * L.QS + cxq(q) + (long long)cyq(q) * L.P + (cz == 1 ? zp : cz == -1 ? zm : 0);
          } else {
            d = (long long)opp(q) * L.QS;
          }
          return uniform_ptr(A + (splane + d)) + soff;
        };
        if constexpr (S == 2) {
          // each value stored as soon as the conserving rounding has made it final; waves with an
          // x-face cell also put its nine outgoing values into the face copy
          typedef __attribute__((address_space(1))) _Float16 hT;
          auto put = [&](int q, _Float16 v) { *(hT *)dst(q) = v; };
          if (xo != nullptr && __any(x == 0 || x == L.nx - 1)) {
            const bool lo = x == 0, hi = !lo && x == L.nx - 1;
            fp16_conserving_each(g, [&](int q, _Float16 v) {
              put(q, v);
              if (q < 9) {
                if (lo) ((hT *)xo)[xo_at(L, 0, q, y, z)] = v;
              } else if (q >= 18) {
                if (hi) ((hT *)xo)[xo_at(L, 1, q - 18, y, z)] = v;
              }
            });
          } else {
            fp16_conserving_each(g, put);
          }
        } else {
#pragma unroll
          for (int q = 0; q < NQ; ++q) st_g<S>(dst(q), g[q], scale);
        }
        // x-face cells keep their outgoing populations for the next farfield pass (FP16 storage:
        // above, as they are rounded)
        if constexpr (S != 2) {
          if (xo != nullptr && (x == 0 || x == L.nx - 1)) {
            typedef __attribute__((address_space(1))) typename St<S>::T gT;
            const bool hi = x != 0;
#pragma unroll
            for (int j = 0; j < 9; ++j) ((gT *)xo)[xo_at(L, hi, j, y, z)] = hi ? g[18 + j] : g[j];
          }
        }
      }
    }
  }
  // rho and u are conserved by the collision: the pre-collision moments are the
  // post-collision state's.  Wave reductions run in the collision's compute type.
  if constexpr (DIAG) diag_wave<C>(active && mk == 0, drho, u2, diag, slot);
}

template <int S, bool ODD, bool PXY, bool DIAG>
__global__ __launch_bounds__(PXY ? 256 : 1024) void k_sweep(typename St<S>::T *__restrict__ A,
                                                            const uint8_t *__restrict__ solid, Desc L, SweepGeom G,
                                                            typename St<S>::C ws, typename St<S>::C wb, float scale,
                                                            float inv_scale, DiagAcc *__restrict__ diag,
                                                            typename St<S>::C *__restrict__ pm,
                                                            typename St<S>::T *__restrict__ xo,
                                                            const typename St<S>::T *__restrict__ xi) {
  sweep_block<S, ODD, PXY, DIAG>(A, solid, L, G, ws, wb, scale, inv_scale, diag, xcd_remap(blockIdx.x, G.nblocks),
                                 (int)(blockIdx.x % DIAG_SLOTS), PXY ? nullptr : pm, PXY ? nullptr : xo,
                                 PXY ? nullptr : xi);
}

// One lattice of a batched sweep: its arguments and the first block of its tiles in the
// concatenated tile list.  `odd` selects the AA form of this member's sweep.
struct BatchMember {
  void *A;
  const uint8_t *solid;
  Desc L;
  SweepGeom G;
  double ws, wb;
  float scale, inv_scale;
  DiagAcc *diag;
  int first_block, odd;
};

// The sweeps of several lattices of one storage type (uneven sizes) in one launch over
// their concatenated tile lists (256-thread tiles; x/y ghosts kept, z periodic).  The
// member of a tile comes from a per-block table (bmem[b] = member index); each member
// keeps its own parity, relaxation rates, scale and diagnostics.  The grid is either one
// block per tile or a persistent grid (a multiple of 8 blocks, at most the tile count) whose
// block i takes tiles i, i + grid, ...: every such index has the same residue mod 8, so the
// XCD-contiguous remap still keeps each XCD on its own slab of tiles.  Without diagnostics
// the tile body has no barrier, so blocks may loop over different numbers of tiles.
template <int S, bool DIAG>
__global__ __launch_bounds__(256) void k_sweep_batch(const BatchMember *__restrict__ M,
                                                     const uint16_t *__restrict__ bmem, int nblocks) {
  static_assert(!DIAG, "the persistent loop assumes a tile body without block barriers");
  using T = typename St<S>::T;
  using C = typename St<S>::C;
  for (int vb = (int)blockIdx.x; vb < nblocks; vb += (int)gridDim.x) {
    const int b = xcd_remap(vb, nblocks);
    const BatchMember &m = M[bmem[b]];
    const int lb = b - m.first_block;
    if (m.odd)
      sweep_block<S, true, false, DIAG>((T *)m.A, m.solid, m.L, m.G, (C)m.ws, (C)m.wb, m.scale, m.inv_scale, m.diag,
                                        lb, lb % DIAG_SLOTS);
    else
      sweep_block<S, false, false, DIAG>((T *)m.A, m.solid, m.L, m.G, (C)m.ws, (C)m.wb, m.scale, m.inv_scale, m.diag,
                                         lb, lb % DIAG_SLOTS);
  }
}

// ------------------------------------------------------------------ copy
// Streaming device copy: two 8-byte words per thread, both loads issued before the
// stores (the attainable HBM read+write bandwidth, and the calibration kernel for the
// memory counters).
__global__ __launch_bounds__(256) void k_stream_copy(const uint64_t *__restrict__ a, uint64_t *__restrict__ b,
                                                     long long n) {
  const long long i = blockIdx.x * 512LL + threadIdx.x, j = i + 256;
  uint64_t v0 = 0, v1 = 0;
  if (i < n) v0 = a[i];
  if (j < n) v1 = a[j];
  if (i < n) b[i] = v0;
  if (j < n) b[j] = v1;
}

// ------------------------------------------------------------------ walls
// Bouzidi interpolated bounce-back over the compacted fluid->solid link list, fused
// with the momentum-exchange loads.  One thread per (link, z).  Per-block sums of
// [Fx, Fy, Fz, Mz, mass] for every element are written in a fixed order; the final
// reduction sums the blocks in a fixed order, so loads are deterministic.
struct WallLinks {
  const int32_t *xy;     // x | (y << 16)
  const int8_t *dir;
  const double *q;
  const int8_t *elem;
  int n, nelem;
  double rx_off, ry_off, lever;  // lever arm = (index + 0.5 + off) * lever
};

constexpr int MAX_ELEM = 8;

template <int S, bool PE, bool LOADS>
__device__ __forceinline__ void walls_body(typename St<S>::T *__restrict__ A, Desc L, WallLinks W, float scale, float inv_scale, double *__restrict__ partials, long long t, int blk) {
  using C = typename St<S>::C;
  const long long total = (long long)W.n * L.nz;
  double f[5] = {0, 0, 0, 0, 0};
  int el = -1;
  if (t < total) {
    const int k = (int)(t / L.nz), z = (int)(t % L.nz);
    const int xy = W.xy[k];
    const int x = xy & 0xffff, y = xy >> 16;
    const int d = W.dir[k], id = opp(d);
    const C q = (C)W.q[k];
    const C fxf = ld<S>(A + loc_out<PE>(L, x, y, z, d), inv_scale);
    const C fxi = ld<S>(A + loc_out<PE>(L, x, y, z, id), inv_scale);
    C val;
    if (q >= C(0.5)) {
      val = C(0.5) * (fxf + fxi * (C(2) * q - C(1))) / q;
    } else if (q > C(0)) {
      // f_out(x - c_d, d): the second fluid node along the link
      const int zz = (int)(((z - czq(d)) % L.nz + L.nz) % L.nz);
      const C d2 = ld<S>(A + loc_out<PE>(L, x - cxq(d), y - cyq(d), zz, d), inv_scale);
      val = C(2) * q * fxf + (C(1) - C(2) * q) * d2;
    } else {
      val = fxf;
    }
    stv<S>(A + loc_in<PE>(L, x, y, z, id), val, scale);
    if constexpr (LOADS) {
      // force on the solid: c_d (outgoing + reflected); lever arm at the fluid cell centre
      const double v = (double)fxf + (double)val;
      f[0] = cxq(d) * v; f[1] = cyq(d) * v; f[2] = czq(d) * v;
      const double rx = (x + 0.5 + W.rx_off) * W.lever, ry = (y + 0.5 + W.ry_off) * W.lever;
      f[3] = rx * f[1] - ry * f[0];
      f[4] = (double)val - (double)fxf;
      el = W.elem[k];
    }
  }
  if constexpr (LOADS) {
    __shared__ double red[4][5];
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    for (int e = 0; e < W.nelem; ++e) {
      double v[5];
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        double a = el == e ? f[c] : 0.0;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) a += __shfl_xor(a, o, 64);
        v[c] = a;
      }
      if (lane == 0)
        for (int c = 0; c < 5; ++c) red[w][c] = v[c];
      __syncthreads();
      if (threadIdx.x == 0) {
        const int nw = blockDim.x / 64;
        for (int c = 0; c < 5; ++c) {
          double s = red[0][c];
          for (int k = 1; k < nw; ++k) s += red[k][c];
          partials[((long long)blk * W.nelem + e) * 5 + c] = s;
        }
      }
      __syncthreads();
    }
  }
}

template <int S, bool PE, bool LOADS>
__global__ __launch_bounds__(256) void k_walls(typename St<S>::T *__restrict__ A, Desc L, WallLinks W,
                                               float scale, float inv_scale, double *__restrict__ partials) {
  walls_body<S, PE, LOADS>(A, L, W, scale, inv_scale, partials, (long long)blockIdx.x * blockDim.x + threadIdx.x, (int)blockIdx.x);
}

// Step number and history-row cursor of a lattice, kept in device memory so that a
// captured step (HIP graph) replays with the right row and step number.  At the
// start of a step `step` is that step's number; the kernel that ends a step
// advances it (and `row` after a sampled step).
struct StepCtl {
  long long step;
  int row, width;
};

__global__ void k_set_ctl(StepCtl *ctl, long long step, int row, int width) {
  if (threadIdx.x == 0) {
    ctl->step = step;
    ctl->row = row;
    ctl->width = width;
  }
}

// End of an unsampled step.
__global__ void k_tick(StepCtl *ctl) {
  if (threadIdx.x == 0) ctl->step += 1;
}

// Lane sums of the wall-load partials (lanes 0..63 of one wave): acc[c * 64 + lane] is the
// sum of column c over the blocks lane, lane + 64, ... in that order.  The columns of one
// block are loaded together (unrolled to the largest element count).
__device__ __forceinline__ void loads_lane_sums(const double *__restrict__ partials, int nblocks, int nelem,
                                                double *__restrict__ acc, int lane) {
  constexpr int NC = MAX_ELEM * 5;
  const int nc = nelem * 5;
  double s[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) s[c] = 0.0;
  for (int b = lane; b < nblocks; b += 64) {
    const double *p = partials + (long long)b * nc;
#pragma unroll
    for (int c = 0; c < NC; ++c)
      if (c < nc) s[c] += p[c];
  }
#pragma unroll
  for (int c = 0; c < NC; ++c)
    if (c < nc) acc[c * 64 + lane] = s[c];
}

// One lane: the loads of a history row from the lane sums, added in a fixed order:
// cl = Fy fscale, cd = Fx fscale, cm = -Mz mscale, mass exchange = mass vscale, and the
// per-element forces and moment.
__device__ __forceinline__ void loads_row(const double *__restrict__ acc, int nelem, double fscale, double mscale,
                                          double vscale, double *__restrict__ row, long long step) {
  double tot[5] = {0, 0, 0, 0, 0};
  for (int e = 0; e < nelem; ++e) {
    double v[5];
    for (int c = 0; c < 5; ++c) {
      double s = 0.0;
      for (int k = 0; k < 64; ++k) s += acc[(e * 5 + c) * 64 + k];
      v[c] = s;
      tot[c] += s;
    }
    for (int c = 0; c < 4; ++c) row[AK_HIST_FIXED + e * 4 + c] = v[c];
  }
  row[0] = (double)step;
  row[1] = tot[1] * fscale;
  row[2] = tot[0] * fscale;
  row[3] = -tot[3] * mscale;
  row[8] = tot[4] * vscale;
}

// Diagnostics slots back to their start values (lanes 0..63 of one wave).
__device__ __forceinline__ void diag_clear_lanes(DiagAcc *diag, int lane) {
  for (int k = lane; k < DIAG_SLOTS; k += 64) {
    diag->s[k].u2max = 0ull;
    diag->s[k].drmax = 0ull;
    diag->s[k].rhomin = (unsigned long long)__double_as_longlong(1e300);
    diag->s[k].nonfinite = 0ull;
  }
}

// Fold the wall-load partials into the current history row (the first wave of the block
// sums, lane 0 writes).  diag != nullptr: the same lanes then clear the diagnostic slots
// and the cursor stays; diag == nullptr: this body ends the sampled step.  `acc` is
// LOADS_LDS doubles of the calling kernel's shared memory (kernels that fuse several phases
// give every phase the same array).
constexpr int LOADS_LDS = MAX_ELEM * 5 * 64;
__device__ __forceinline__ void loads_final_body(const double *__restrict__ partials, int nblocks, int nelem, double fscale,
                              double mscale, double vscale, StepCtl *ctl, double *__restrict__ hist, DiagAcc *diag,
                              double *__restrict__ acc) {
  const int lane = threadIdx.x;
  if (lane < 64) loads_lane_sums(partials, nblocks, nelem, acc, lane);
  __syncthreads();
  if (lane == 0) {
    loads_row(acc, nelem, fscale, mscale, vscale, hist + (long long)ctl->row * ctl->width, ctl->step);
    if (!diag) {
      ctl->step += 1;
      ctl->row += 1;
    }
  }
  if (!diag || lane >= 64) return;
  diag_clear_lanes(diag, lane);
}

// Diagnostics slots set for a sampled sweep to accumulate into (lattice creation; each
// sampled step's end clears them again).
__global__ void k_diag_clear(DiagAcc *diag) {
  for (int k = threadIdx.x; k < DIAG_SLOTS; k += blockDim.x) {
    diag->s[k].u2max = 0ull;
    diag->s[k].drmax = 0ull;
    diag->s[k].rhomin = (unsigned long long)__double_as_longlong(1e300);
    diag->s[k].nonfinite = 0ull;
  }
}

// Diagnostics slots reduced by one wave (lanes 0..63; every lane ends with the totals):
// largest |u|^2 and |drho|, smallest rho, non-finite flag.
struct DiagTot {
  unsigned long long u, d, r, b;
};
__device__ __forceinline__ DiagTot diag_reduce_wave(const DiagAcc *diag, int lane) {
  unsigned long long u = 0ull, d = 0ull, r = (unsigned long long)__double_as_longlong(1e300), b = 0ull;
  for (int k = lane; k < DIAG_SLOTS; k += 64) {
    const DiagSlot &sl = diag->s[k];
    u = sl.u2max > u ? sl.u2max : u;
    d = sl.drmax > d ? sl.drmax : d;
    r = sl.rhomin < r ? sl.rhomin : r;
    b = sl.nonfinite > b ? sl.nonfinite : b;
  }
  for (int o = 32; o >= 1; o >>= 1) {
    const unsigned long long u2 = __shfl_xor(u, o, 64), d2 = __shfl_xor(d, o, 64);
    const unsigned long long r2 = __shfl_xor(r, o, 64), b2 = __shfl_xor(b, o, 64);
    u = u2 > u ? u2 : u; d = d2 > d ? d2 : d; r = r2 < r ? r2 : r; b = b2 > b ? b2 : b;
  }
  return DiagTot{u, d, r, b};
}
__device__ __forceinline__ void diag_row(const DiagTot &t, double *__restrict__ row) {
  row[4] = sqrt(3.0) * sqrt(__longlong_as_double((long long)t.u));
  row[5] = __longlong_as_double((long long)t.d);
  row[6] = __longlong_as_double((long long)t.r);
  row[7] = (double)t.b;
}

// End of a sampled uniform step, one block of two waves: wave 0 reduces the sweep's
// diagnostics slots, wave 1 sums the wall-load partials (their load chains overlap); then
// lane 0 of each writes its columns of the current history row, wave 0 clears the slots
// for the next sampled sweep and thread 0 advances the cursor (read by every thread before
// the barrier).
__global__ __launch_bounds__(128) void k_step_final(const double *__restrict__ partials, int nblocks, int nelem,
                                                    double fscale, double mscale, double vscale, StepCtl *ctl,
                                                    double *__restrict__ hist, DiagAcc *diag) {
  __shared__ double acc[LOADS_LDS];
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const long long step = ctl->step;
  const int rowi = ctl->row;
  double *row = hist + (long long)rowi * ctl->width;
  DiagTot t{};
  if (w == 0) t = diag_reduce_wave(diag, lane);
  else loads_lane_sums(partials, nblocks, nelem, acc, lane);
  __syncthreads();
  if (lane == 0) {
    if (w == 0) diag_row(t, row);
    else loads_row(acc, nelem, fscale, mscale, vscale, row, step);
  }
  if (w == 0) diag_clear_lanes(diag, lane);
  if (threadIdx.x == 0) {
    ctl->step = step + 1;
    ctl->row = rowi + 1;
  }
}

// ------------------------------------------------------------------ farfield
// Characteristic farfield (isothermal Riemann invariants, NEEP reconstruction, linear
// transverse transport).  Work item = one exterior ghost cell g of one face/edge
// object; every link entering the fluid from g shares the foot g - n, hence the
// same boundary state, which is computed once per item.
struct FFItems {
  const int32_t *g;        // ghost cell: (x+1) | (y+1) << 16
  const int32_t *gz;       // its z in [0, nz)
  const int8_t *obj;       // object 0..7
  const int32_t *dmask;    // bit d set: link from g - c_d (direction d) exists
  double *incoming, *derivative;
  long long *clock;
  const double *initial;   // glancing trace rows: [side][z][x], side 0 = y low, 1 = y high
  int n;
  const void *pm;          // perimeter moments of the current state (compute type), or null
  const void *xo;          // outgoing populations of the x-face cells (storage type, xo_at), or null
  void *xi;                // their populations entering through the x faces, for the next sweep, or null
};

struct FFObj {
  int n[3];
  double nh[3], un_inf, ut_inf[3], P[3][3];
  int evolve, glancing, use_axis[3], side;
};

struct FFParams {
  FFObj o[8];
  double U, rho0;
};

template <typename C>
__device__ __forceinline__ C log_rho(C drho, C rho0) {
  // ln(rho / rho0) with rho = 1 + drho
  if constexpr (sizeof(C) == 8) return log((C(1) + drho) / rho0);
  else return log1pf(drho) - logf(rho0);
}

template <int S, class Macro>
__device__ double ff_transport(const Desc &L, const FFObj &o, const int foot[3], const int edge[2],
                               typename St<S>::C rm, const typename St<S>::C (&u)[3], double rho0, Macro &&macro) {
  using C = typename St<S>::C;
  const C cs = (C)0.57735026918962576451;
  C F = 0;
  for (int a = 0; a < 3; ++a) {
    if (!o.use_axis[a]) continue;
    const int e = a < 2 ? edge[a] : 0;
    const int off1 = e == -1 ? 1 : -1;
    const int off2 = e == 0 ? 1 : -2 * e;
    const C c0 = C(1.5) * e;
    const C c1 = e == 0 ? C(-0.5) : C(-2) * e;
    const C c2 = e == 0 ? C(0.5) : C(0.5) * e;
    int p1[3] = {foot[0], foot[1], foot[2]}, p2[3] = {foot[0], foot[1], foot[2]};
    p1[a] += off1; p2[a] += off2;
    p1[2] = ((p1[2] % L.nz) + L.nz) % L.nz;
    p2[2] = ((p2[2] % L.nz) + L.nz) % L.nz;
    C d1, d2, u1[3], u2[3];
    macro(p1[0], p1[1], p1[2], d1, u1);
    macro(p2[0], p2[1], p2[2], d2, u2);
    C pdu = 0;
    for (int b = 0; b < 3; ++b) pdu += (C)o.P[a][b] * (c0 * u[b] + c1 * u1[b] + c2 * u2[b]);
    const C rm1 = (C)o.nh[0] * u1[0] + (C)o.nh[1] * u1[1] + (C)o.nh[2] * u1[2] - cs * log_rho<C>(d1, (C)rho0);
    const C rm2 = (C)o.nh[0] * u2[0] + (C)o.nh[1] * u2[1] + (C)o.nh[2] * u2[2] - cs * log_rho<C>(d2, (C)rho0);
    F += (C)o.un_inf * pdu - (C)o.ut_inf[a] * (c0 * rm + c1 * rm1 + c2 * rm2);
  }
  return (double)F;
}

__device__ __forceinline__ void ff_unpack(int32_t p, int &gx, int &gy) {
  gx = (p & 0xffff) - 1;
  gy = ((p >> 16) & 0xffff) - 1;
}

template <int S, bool PE>
__device__ __forceinline__ void farfield_body(typename St<S>::T *__restrict__ A, const Desc &L, const FFItems &I,
                                              const FFParams &Pm, float scale, float inv_scale, long long t) {
  using C = typename St<S>::C;
  if (t >= I.n) return;
  const int k = (int)t;
  int gx, gy;
  ff_unpack(I.g[k], gx, gy);
  const int gz = I.gz[k];
  const FFObj &o = Pm.o[I.obj[k]];
  int foot[3] = {gx - o.n[0], gy - o.n[1], gz};
  const int edge[2] = {foot[0] == 0 ? -1 : (foot[0] == L.nx - 1 ? 1 : 0),
                       foot[1] == 0 ? -1 : (foot[1] == L.ny - 1 ? 1 : 0)};
  // the foot and the transverse neighbours of the transport are perimeter cells: their
  // moments come from the perimeter moments when the sweep keeps them
  const C *pm = (const C *)I.pm;
  auto macro = [&](int x, int y, int z, C &dr, C (&uu)[3]) {
    if (pm) {
      const bool xf = x == 0 || x == L.nx - 1;
      const C *p = pm + (xf ? pm_x(L, x != 0, y, z) : pm_y(L, y != 0, x, z));
      const int st = xf ? L.ny : L.nx;
      dr = p[0]; uu[0] = p[st]; uu[1] = p[2 * st]; uu[2] = p[3 * st];
    } else {
      macro_cell<S, PE>(A, L, x, y, z, inv_scale, dr, uu);
    }
  };
  C drho, u[3];
  macro(foot[0], foot[1], foot[2], drho, u);
  const C cs = (C)0.57735026918962576451;
  const C un = (C)o.nh[0] * u[0] + (C)o.nh[1] * u[1] + (C)o.nh[2] * u[2];
  const C lr = log_rho<C>(drho, (C)Pm.rho0);
  const C rm = un - cs * lr, rp = un + cs * lr;
  C incoming;
  if (o.evolve && !o.glancing) {
    const double F = ff_transport<S>(L, o, foot, edge, rm, u, Pm.rho0, macro);
    const double inc = I.incoming[k] + (F + I.derivative[k]) / 2.0;
    I.incoming[k] = inc;
    I.derivative[k] = F;
    incoming = (C)inc;
  } else if (o.glancing) {
    // one interpolation of the stored initial invariant at the total traced shift
    const long long clk = I.clock[k];
    const double target = (double)foot[0] - Pm.U * (double)clk;
    const int lower = (int)floor(target);
    const double frac = target - lower;
    const double *row = I.initial + ((long long)o.side * L.nz + foot[2]) * L.nx;
    double s = 0.0;
    for (int bit = 0; bit < 2; ++bit) {
      const int kx = lower + bit;
      const double wgt = bit ? frac : 1.0 - frac;
      const double v = (kx >= 0 && kx < L.nx) ? row[kx] : 0.0;
      s += wgt * v;
    }
    incoming = (C)(o.un_inf + s);
    I.clock[k] = clk + 1;
  } else {
    incoming = (C)o.un_inf;
  }
  const C unb = (rp + incoming) * C(0.5);
  const C rho = C(1) + drho;
  const C rhob = (C)Pm.rho0 * exp((rp - incoming) / (C(2) * cs));
  C ub[3];
  const C vinf[3] = {(C)Pm.U, C(0), C(0)};
  for (int a = 0; a < 3; ++a)
    ub[a] = (C)o.nh[a] * unb + (unb >= C(0) ? u[a] - (C)o.nh[a] * un : vinf[a] - (C)o.nh[a] * (C)o.un_inf);
  const int mask = I.dmask[k];
#pragma unroll
  for (int d = 0; d < NQ; ++d) {
    if (!((mask >> d) & 1)) continue;
    const int id = opp(d);
    // an x-face foot's populations leaving through that face come from the sweep's face copy
    const int xside = foot[0] == 0 && cxq(id) == -1 ? 0 : (foot[0] == L.nx - 1 && cxq(id) == 1 ? 1 : -1);
    const C fo = I.xo != nullptr && xside >= 0
                     ? ld<S>((const typename St<S>::T *)I.xo +
                                 xo_at(L, xside, xside ? id - 18 : id, foot[1], foot[2]), inv_scale)
                     : ld<S>(A + loc_out<PE>(L, foot[0], foot[1], foot[2], id), inv_scale);
    const C val = fo + (equilibrium<C>(id, rhob, ub[0], ub[1], ub[2]) - equilibrium<C>(id, rho, u[0], u[1], u[2]));
    // fluid cell x = g - c_d reads direction id; what enters an x-face cell through that face
    // goes to the per-cell copy (xo_at layout) that the next sweep reads, not to nine lattice rows
    const int xz = ((gz - czq(d)) % L.nz + L.nz) % L.nz;
    const int tx = gx - cxq(d);
    const int tside = tx == 0 && cxq(id) == 1 ? 0 : (tx == L.nx - 1 && cxq(id) == -1 ? 1 : -1);
    if (I.xi != nullptr && tside >= 0)
      stv<S>((typename St<S>::T *)I.xi + xo_at(L, tside, tside ? id : id - 18, gy - cyq(d), xz), val, scale);
    else
      stv<S>(A + loc_in<PE>(L, tx, gy - cyq(d), xz, id), val, scale);
  }
}

template <int S, bool PE>
__global__ __launch_bounds__(128) void k_farfield(typename St<S>::T *__restrict__ A, Desc L, FFItems I,
                                                  FFParams Pm, float scale, float inv_scale) {
  farfield_body<S, PE>(A, L, I, Pm, scale, inv_scale, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Perimeter moments of a state in form PE from the lattice (start of a run call; every
// sweep then keeps them).  Thread = (y, side, z) of the x faces, then (x, side, z) of the
// y faces.
template <int S, bool PE>
__global__ void k_perimeter_fill(const typename St<S>::T *__restrict__ A, Desc L, typename St<S>::C *__restrict__ pm,
                                 float inv_scale) {
  using C = typename St<S>::C;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nxf = 2LL * L.ny * L.nz;
  int x, y, z;
  bool xface;
  if (t < nxf) {
    y = (int)(t % L.ny);
    x = (int)((t / L.ny) % 2) ? L.nx - 1 : 0;
    z = (int)(t / (2LL * L.ny));
    xface = true;
  } else {
    t -= nxf;
    if (t >= 2LL * L.nx * L.nz) return;
    x = (int)(t % L.nx);
    y = (int)((t / L.nx) % 2) ? L.ny - 1 : 0;
    z = (int)(t / (2LL * L.nx));
    xface = false;
  }
  C dr, u[3];
  macro_cell<S, PE>(A, L, x, y, z, inv_scale, dr, u);
  C *p = pm + (xface ? pm_x(L, x != 0, y, z) : pm_y(L, y != 0, x, z));
  const int st = xface ? L.ny : L.nx;
  p[0] = dr; p[st] = u[0]; p[2 * st] = u[1]; p[3 * st] = u[2];
}

// The x-face copy (xo_at) of a state in form PE from the lattice: thread = (y, side, z).  Used
// when the state was set from outside or loaded (every sweep then keeps the copy).
template <int S, bool PE>
__global__ void k_xface_fill(const typename St<S>::T *__restrict__ A, Desc L, typename St<S>::T *__restrict__ xo) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2LL * L.ny * L.nz) return;
  const int y = (int)(t % L.ny), side = (int)((t / L.ny) % 2), z = (int)(t / (2LL * L.ny));
  const int x = side ? L.nx - 1 : 0;
#pragma unroll
  for (int j = 0; j < 9; ++j) xo[xo_at(L, side, j, y, z)] = A[loc_out<PE>(L, x, y, z, side ? 18 + j : j)];
}

// Initial farfield state from the initial populations (post-even form): incoming =
// R- - F and derivative = F on evolving objects (a backward Taylor seed), and the
// trace rows of glancing objects.
template <int S>
__global__ void k_farfield_init(const typename St<S>::T *__restrict__ A, Desc L, FFItems I, FFParams Pm,
                                float inv_scale) {
  using C = typename St<S>::C;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= I.n) return;
  int gx, gy;
  ff_unpack(I.g[k], gx, gy);
  const int gz = I.gz[k];
  const FFObj &o = Pm.o[I.obj[k]];
  int foot[3] = {gx - o.n[0], gy - o.n[1], gz};
  const int edge[2] = {foot[0] == 0 ? -1 : (foot[0] == L.nx - 1 ? 1 : 0),
                       foot[1] == 0 ? -1 : (foot[1] == L.ny - 1 ? 1 : 0)};
  if (o.glancing) { I.clock[k] = 0; return; }
  if (!o.evolve) { I.incoming[k] = o.un_inf; I.derivative[k] = 0.0; return; }
  C drho, u[3];
  macro_cell<S, true>(A, L, foot[0], foot[1], foot[2], inv_scale, drho, u);
  const C cs = (C)0.57735026918962576451;
  const C rm = (C)o.nh[0] * u[0] + (C)o.nh[1] * u[1] + (C)o.nh[2] * u[2] - cs * log_rho<C>(drho, (C)Pm.rho0);
  const double F = ff_transport<S>(L, o, foot, edge, rm, u, Pm.rho0, [&](int x, int y, int z, C &dr, C (&uu)[3]) {
    macro_cell<S, true>(A, L, x, y, z, inv_scale, dr, uu);
  });
  I.incoming[k] = (double)rm - F;
  I.derivative[k] = F;
}

// Initial outgoing invariant of the two y boundary rows (glancing objects):
// rows[side][z][x] = u.n - cs ln(rho/rho0) - un_inf with n = (0, -1, 0) / (0, 1, 0).
template <int S>
__global__ void k_glancing_rows(const typename St<S>::T *__restrict__ A, Desc L, double *rows, double rho0,
                                float inv_scale) {
  using C = typename St<S>::C;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2LL * L.nz * L.nx) return;
  const int x = (int)(t % L.nx), z = (int)((t / L.nx) % L.nz), side = (int)(t / ((long long)L.nx * L.nz));
  const int y = side == 0 ? 0 : L.ny - 1;
  C drho, u[3];
  macro_cell<S, true>(A, L, x, y, z, inv_scale, drho, u);
  const double ny_ = side == 0 ? -1.0 : 1.0;
  rows[t] = ny_ * (double)u[1] - 0.57735026918962576451 * (double)log_rho<C>(drho, (C)rho0);
}

// ------------------------------------------------------------------ legacy farfield
// Velocity bounce-back on x-, y-, y+ (density of the fluid cell) and the extrapolation
// outflow on x+ (c-weighted blend of the last outgoing population of the tangential
// neighbour and the previous boundary value).  Work item = one link.
struct LegacyLinks {
  const int32_t *x;   // fluid cell packed like FFItems::g: (x+1) | (y+1) << 16
  const int32_t *z;
  const int8_t *dir;
  const int8_t *kind; // 0 velocity bounce-back, 1 outflow
  double *pdf, *pdf_nd;
  int n;
};

template <int S, bool PE>
__device__ __forceinline__ typename St<S>::C ubb_value(const typename St<S>::T *A, const Desc &L, int x, int y,
                                                      int z, int d, double U, float inv_scale) {
  using C = typename St<S>::C;
  C r = 0;
#pragma unroll
  for (int j = 0; j < NQ; ++j) r += ld<S>(A + loc_out<PE>(L, x, y, z, j), inv_scale);
  const C rho = C(1) + r;
  const C fo = ld<S>(A + loc_out<PE>(L, x, y, z, d), inv_scale);
  return fo - C(6) * (C)wq(d) * (C)(cxq(d) * U) * rho;
}

template <int S, bool PE>
__device__ __forceinline__ void legacy_body(typename St<S>::T *__restrict__ A, Desc L, LegacyLinks K, double U, float scale, float inv_scale, long long t) {
  using C = typename St<S>::C;
  const int k = (int)t;
  if (k >= K.n) return;
  int x, y;
  ff_unpack(K.x[k], x, y);
  const int z = K.z[k];
  const int d = K.dir[k], id = opp(d);
  C val;
  if (K.kind[k] == 0) {
    val = ubb_value<S, PE>(A, L, x, y, z, d, U, inv_scale);
  } else {
    const double c = 0.57735026918962576451;
    const double v = K.pdf[k] * c + (1.0 - c) * K.pdf_nd[k];
    val = (C)v;
    // new history: f_out(t, id) of the tangential neighbour t = x + c_d - n
    const int ty = y + cyq(d), tz = ((z + czq(d)) % L.nz + L.nz) % L.nz;
    double nv;
    if (ty < 0 || ty >= L.ny) {
      // t is a velocity-bounce-back ghost: its slot holds the bounce-back value of
      // link (t - c_d, d), written in this same boundary pass
      nv = (double)ubb_value<S, PE>(A, L, x - 1, ty - cyq(d), z, d, U, inv_scale);
    } else {
      nv = (double)ld<S>(A + loc_out<PE>(L, x, ty, tz, id), inv_scale);
    }
    K.pdf[k] = nv;
    K.pdf_nd[k] = v;
  }
  stv<S>(A + loc_in<PE>(L, x, y, z, id), val, scale);
}

template <int S, bool PE>
__global__ __launch_bounds__(128) void k_legacy(typename St<S>::T *__restrict__ A, Desc L, LegacyLinks K,
                                                double U, float scale, float inv_scale) {
  legacy_body<S, PE>(A, L, K, U, scale, inv_scale, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// The walls and the edge phase as one launch over consecutive block ranges, so that their
// latencies overlap: the walls (blocks below nb_wall, so the wall partials keep their
// per-block layout), then the legacy links (EDGE 0) or the characteristic farfield items
// (EDGE 1).  With every wall link at least three cells from the x and y edges the walls
// read and write slots of cells next to the body only, and the edge phase those of the
// cells next to the edges: no slot written by one is touched by the other.
template <int S, bool PE, bool LOADS, int EDGE>
__global__ __launch_bounds__(256) void k_walls_edge(typename St<S>::T *__restrict__ A, Desc L, WallLinks W,
                                                    float scale, float inv_scale, double *__restrict__ partials,
                                                    int nb_wall, LegacyLinks K, double U, FFItems I, FFParams Pm) {
  const int b = blockIdx.x;
  if (b < nb_wall) {
    walls_body<S, PE, LOADS>(A, L, W, scale, inv_scale, partials, (long long)b * 256 + threadIdx.x, b);
  } else if constexpr (EDGE == 0) {
    legacy_body<S, PE>(A, L, K, U, scale, inv_scale, (long long)(b - nb_wall) * 256 + threadIdx.x);
  } else {
    farfield_body<S, PE>(A, L, I, Pm, scale, inv_scale, (long long)(b - nb_wall) * 256 + threadIdx.x);
  }
}

// ------------------------------------------------------------------ init / exchange
struct SeedParams {
  double U, amp, phy, phz, ky, kz, lex, ley, ox, oy, dx;
};

// Seeded initial field, product-form equilibrium with rho = 1, written in post-even
// form.  Evaluated in FP64 for every storage precision.
template <int S>
__global__ void k_init_seeded(typename St<S>::T *__restrict__ A, const uint8_t *__restrict__ solid, Desc L,
                              SeedParams p, float scale) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)L.nx * L.ny * L.nz) return;
  const int x = (int)(t % L.nx), y = (int)((t / L.nx) % L.ny), z = (int)(t / ((long long)L.nx * L.ny));
  const double xx = ((double)x + 0.5) * p.dx + p.ox, yy = ((double)y + 0.5) * p.dx + p.oy;
  const double zz = ((double)z + 0.5) * p.dx;
  const double rx = xx - p.lex, ry = yy - p.ley;
  const double e = (rx - 0.5) / 0.75;
  const double env = exp(-e * e);
  double ux = p.U;
  double uy = p.amp * sin(p.ky * ry + p.phy) * cos(p.kz * zz + p.phz) * env;
  double uz = -p.amp * p.ky / p.kz * cos(p.ky * ry + p.phy) * sin(p.kz * zz + p.phz) * env;
  const uint8_t mk = solid[(long long)(y + 1) * (L.nx + 2) + (x + 1)];
  if (mk == 3) { ux = p.U; uy = uz = 0.0; }  // exterior ring: free stream
  else if (mk) ux = uy = uz = 0.0;
  const long long c = L.base + cell_of(L, x, y, z);
#pragma unroll
  for (int i = 0; i < NQ; ++i) {
    const double g = equilibrium<double>(i, 1.0, ux, uy, uz) - wq(i);
    stv<S>(A + c + (long long)opp(i) * L.QS, (typename St<S>::C)g, scale);
  }
}

template <int S, bool PE, bool PXY>
__global__ void k_export(const typename St<S>::T *__restrict__ A, Desc L, double *__restrict__ out,
                         float inv_scale) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)L.nx * L.ny * L.nz) return;
  const int z = (int)(t % L.nz), y = (int)((t / L.nz) % L.ny), x = (int)(t / ((long long)L.nz * L.ny));
  const long long c = L.base + cell_of(L, x, y, z);
  for (int j = 0; j < NQ; ++j) {
    long long src;
    if constexpr (PE) src = c + (long long)opp(j) * L.QS;
    else src = c + (long long)j * L.QS + nb_off<PXY>(L, x, y, z, cxq(j), cyq(j), czq(j));
    out[t * NQ + j] = (double)ld<S>(A + src, inv_scale);
  }
}

// Incoming (pre-collision) populations f_in(x, i) as the next sweep would gather them.
template <int S, bool PE>
__global__ void k_export_in(const typename St<S>::T *__restrict__ A, Desc L, double *__restrict__ out,
                            float inv_scale) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)L.nx * L.ny * L.nz) return;
  const int z = (int)(t % L.nz), y = (int)((t / L.nz) % L.ny), x = (int)(t / ((long long)L.nz * L.ny));
  for (int j = 0; j < NQ; ++j) out[t * NQ + j] = (double)ld<S>(A + loc_in<PE>(L, x, y, z, j), inv_scale);
}

template <int S, bool PE, bool PXY>
__global__ void k_import(typename St<S>::T *__restrict__ A, Desc L, const double *__restrict__ in, float scale) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)L.nx * L.ny * L.nz) return;
  const int z = (int)(t % L.nz), y = (int)((t / L.nz) % L.ny), x = (int)(t / ((long long)L.nz * L.ny));
  const long long c = L.base + cell_of(L, x, y, z);
  for (int j = 0; j < NQ; ++j) {
    long long dst;
    if constexpr (PE) dst = c + (long long)opp(j) * L.QS;
    else dst = c + (long long)j * L.QS + nb_off<PXY>(L, x, y, z, cxq(j), cyq(j), czq(j));
    stv<S>(A + dst, (typename St<S>::C)in[t * NQ + j], scale);
  }
}

// Read-only state diagnostics: per-block [sum rho, max |u|^2, max |rho-1|, min rho, bad]
// over fluid cells (the guards also over the exterior ring, the density sum over the
// physical fluid only); k_diag_final folds the blocks in a fixed order.  The grid strides over
// the lattice, so the partials need at most DIAG_MAX_BLOCKS x 5 doubles whatever its size.
// IN reads the incoming (pre-collision) populations, the stored state of a two-level case's
// levels.
constexpr int DIAG_MAX_BLOCKS = 4096;
template <int S, bool PE, bool PXY, bool IN>
__global__ __launch_bounds__(256) void k_diag_state(const typename St<S>::T *__restrict__ A,
                                                    const uint8_t *__restrict__ solid, Desc L,
                                                    float inv_scale, double *__restrict__ part) {
  using C = typename St<S>::C;
  const long long n = (long long)L.nx * L.ny * L.nz;
  const long long stride = (long long)gridDim.x * blockDim.x;
  double sr = 0.0, vu = 0.0, vd = 0.0, vr = 1e300;
  int bad = 0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
    const int x = (int)(t % L.nx), y = (int)((t / L.nx) % L.ny), z = (int)(t / ((long long)L.nx * L.ny));
    const uint8_t mk = solid[(long long)(y + 1) * (L.nx + 2) + (x + 1)];
    if (mk != 0 && mk != 3) continue;
    const long long c = L.base + cell_of(L, x, y, z);
    C r = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
      long long src;
      if constexpr (IN) src = loc_in<PE>(L, x, y, z, j);
      else if constexpr (PE) src = c + (long long)opp(j) * L.QS;
      else src = c + (long long)j * L.QS + nb_off<PXY>(L, x, y, z, cxq(j), cyq(j), czq(j));
      const C v = ld<S>(A + src, inv_scale);
      r += v;
      if (cxq(j)) jx += cxq(j) * v;
      if (cyq(j)) jy += cyq(j) * v;
      if (czq(j)) jz += czq(j) * v;
    }
    const double rho = 1.0 + (double)r;
    const double ux = (double)jx / rho, uy = (double)jy / rho, uz = (double)jz / rho;
    const double u2 = ux * ux + uy * uy + uz * uz;
    if (!(isfinite(rho) && isfinite(u2))) {
      bad = 1;
    } else {
      if (mk == 0) sr += rho;
      vu = u2 > vu ? u2 : vu;
      const double d = fabs((double)r);
      vd = d > vd ? d : vd;
      vr = rho < vr ? rho : vr;
    }
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    sr += __shfl_xor(sr, o, 64);
    double a = __shfl_xor(vu, o, 64); vu = a > vu ? a : vu;
    a = __shfl_xor(vd, o, 64); vd = a > vd ? a : vd;
    a = __shfl_xor(vr, o, 64); vr = a < vr ? a : vr;
  }
  bad = __any(bad) ? 1 : 0;
  __shared__ double sh[4][5];
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0) { sh[w][0] = sr; sh[w][1] = vu; sh[w][2] = vd; sh[w][3] = vr; sh[w][4] = bad; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double o0 = sh[0][0], o1 = sh[0][1], o2 = sh[0][2], o3 = sh[0][3], o4 = sh[0][4];
    for (int k = 1; k < 4; ++k) {
      o0 += sh[k][0]; o1 = sh[k][1] > o1 ? sh[k][1] : o1; o2 = sh[k][2] > o2 ? sh[k][2] : o2;
      o3 = sh[k][3] < o3 ? sh[k][3] : o3; o4 = sh[k][4] > o4 ? sh[k][4] : o4;
    }
    double *p = part + (long long)blockIdx.x * 5;
    p[0] = o0; p[1] = o1; p[2] = o2; p[3] = o3; p[4] = o4;
  }
}

// Fold of the diagnostics partials (one wave, fixed order: lane-strided blocks, then lanes in
// order): out = {local Mach max, max |rho-1|, min rho, non-finite flag, sum of rho}.
__global__ void k_diag_final(const double *__restrict__ part, int nblocks, double *__restrict__ out) {
  __shared__ double acc[5][64];
  const int lane = threadIdx.x;
  double s = 0.0, u = 0.0, d = 0.0, r = 1e300, b = 0.0;
  for (int k = lane; k < nblocks; k += 64) {
    const double *p = part + (long long)k * 5;
    s += p[0];
    u = p[1] > u ? p[1] : u;
    d = p[2] > d ? p[2] : d;
    r = p[3] < r ? p[3] : r;
    b = p[4] > b ? p[4] : b;
  }
  acc[0][lane] = s; acc[1][lane] = u; acc[2][lane] = d; acc[3][lane] = r; acc[4][lane] = b;
  __syncthreads();
  if (lane == 0) {
    for (int k = 1; k < 64; ++k) {
      s += acc[0][k];
      u = acc[1][k] > u ? acc[1][k] : u;
      d = acc[2][k] > d ? acc[2][k] : d;
      r = acc[3][k] < r ? acc[3][k] : r;
      b = acc[4][k] > b ? acc[4][k] : b;
    }
    out[0] = sqrt(3.0) * sqrt(u);
    out[1] = d;
    out[2] = r;
    out[3] = b;
    out[4] = s;
  }
}

// rho and u of one z plane (interior x, y), x-major output: rho[x*ny + y], u[(x*ny + y)*3 + a];
// IN as for k_diag_state
template <int S, bool PE, bool PXY, bool IN>
__global__ void k_plane(const typename St<S>::T *__restrict__ A, Desc L, int z, float inv_scale,
                        double *__restrict__ rho_out, double *__restrict__ u_out) {
  using C = typename St<S>::C;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)L.nx * L.ny) return;
  const int y = (int)(t % L.ny), x = (int)(t / L.ny);
  const long long c = L.base + cell_of(L, x, y, z);
  C r = 0, jx = 0, jy = 0, jz = 0;
  for (int j = 0; j < NQ; ++j) {
    long long src;
    if constexpr (IN) src = loc_in<PE>(L, x, y, z, j);
    else if constexpr (PE) src = c + (long long)opp(j) * L.QS;
    else src = c + (long long)j * L.QS + nb_off<PXY>(L, x, y, z, cxq(j), cyq(j), czq(j));
    const C v = ld<S>(A + src, inv_scale);
    r += v; jx += cxq(j) * v; jy += cyq(j) * v; jz += czq(j) * v;
  }
  const double rho = 1.0 + (double)r;
  rho_out[t] = rho;
  u_out[3 * t] = (double)jx / rho; u_out[3 * t + 1] = (double)jy / rho; u_out[3 * t + 2] = (double)jz / rho;
}

// Stationarity residual of the conserved fields q = (rho, j / U) over owned fluid cells
// (mask 0): d = (q - q_prev) / dt.  Per-block sums of d^2 and maxima of |d| per channel and
// the cell count (fixed-order reductions), then q_prev = q.  IN selects the incoming
// (pre-collision) populations of a refined level, otherwise the stored post-collision ones.
template <int S, bool PE, bool PXY, bool IN>
__global__ __launch_bounds__(256) void k_residual(const typename St<S>::T *__restrict__ A,
                                                  const uint8_t *__restrict__ solid, Desc L, float inv_scale,
                                                  double inv_u, double inv_dt, int reset, double *__restrict__ prev,
                                                  double *__restrict__ part) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // sum d^2 [4], max |d| [4], cells
  if (t < (long long)L.nx * L.ny * L.nz) {
    const int z = (int)(t % L.nz), y = (int)((t / L.nz) % L.ny), x = (int)(t / ((long long)L.nz * L.ny));
    if (solid[(long long)(y + 1) * (L.nx + 2) + (x + 1)] == 0) {
      const long long c = L.base + cell_of(L, x, y, z);
      double r = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
      for (int j = 0; j < NQ; ++j) {
        long long src;
        if constexpr (IN) src = loc_in<PE>(L, x, y, z, j);
        else if constexpr (PE) src = c + (long long)opp(j) * L.QS;
        else src = c + (long long)j * L.QS + nb_off<PXY>(L, x, y, z, cxq(j), cyq(j), czq(j));
        const double g = (double)ld<S>(A + src, inv_scale);
        r += g; jx += cxq(j) * g; jy += cyq(j) * g; jz += czq(j) * g;
      }
      const double q[4] = {1.0 + r, jx * inv_u, jy * inv_u, jz * inv_u};
      double *p = prev + t * 4;
      for (int k = 0; k < 4; ++k) {
        if (!reset) {
          const double d = (q[k] - p[k]) * inv_dt;
          v[k] = d * d;
          v[4 + k] = fabs(d);
        }
        p[k] = q[k];
      }
      v[8] = 1.0;
    }
  }
  for (int o = 32; o >= 1; o >>= 1)
    for (int k = 0; k < 9; ++k) {
      const double w = __shfl_xor(v[k], o, 64);
      v[k] = (k >= 4 && k < 8) ? (w > v[k] ? w : v[k]) : v[k] + w;
    }
  __shared__ double red[4][9];
  const int w = threadIdx.x / 64;
  if (threadIdx.x % 64 == 0)
    for (int k = 0; k < 9; ++k) red[w][k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < 9; ++k) {
      double a = red[0][k];
      for (int i = 1; i < (int)(blockDim.x / 64); ++i) a = (k >= 4 && k < 8) ? (red[i][k] > a ? red[i][k] : a) : a + red[i][k];
      part[(long long)blockIdx.x * 9 + k] = a;
    }
  }
}

// Fixed-order fold of the residual block partials: out = {sum d^2 [4], max |d| [4], cells}.
__global__ void k_residual_final(const double *__restrict__ part, int nblocks, double *__restrict__ out) {
  __shared__ double acc[9][64];
  const int lane = threadIdx.x;  // 64 threads
  for (int k = 0; k < 9; ++k) {
    double a = 0.0;
    for (int b = lane; b < nblocks; b += 64) {
      const double x = part[(long long)b * 9 + k];
      a = (k >= 4 && k < 8) ? (x > a ? x : a) : a + x;
    }
    acc[k][lane] = a;
  }
  __syncthreads();
  if (lane == 0)
    for (int k = 0; k < 9; ++k) {
      double a = acc[k][0];
      for (int i = 1; i < 64; ++i) a = (k >= 4 && k < 8) ? (acc[k][i] > a ? acc[k][i] : a) : a + acc[k][i];
      out[k] = a;
    }
}

template <typename T>
__global__ void k_fill(T *p, long long n, T v) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = t; i < n; i += stride) p[i] = v;
}

}  // namespace ak
