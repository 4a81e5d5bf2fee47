// ak_device.hpp -- device-side building blocks of the D3Q27 cumulant step on gfx950.
//
// Population storage (one array per lattice, AA pattern, in place):
//   element (q, x, y, z) at  base + q*QS + z*PS + (y+1)*P + x,
//   x in [-1, nx], y in [-1, ny] (one ghost cell each side), z in [0, nz) periodic.
// The x = -1 ghost of a row is the last slot of the previous memory row (a row holds
// >= nx+2 slots), so interior rows start on a 128-byte boundary.  The 27 direction rows of
// each y row are interleaved (QS = row pitch, P = 27 row pitches) unless a z plane of that
// layout exceeds 2 GiB; such lattices keep one plane per direction (QS = nz*PS).  Kernels
// only use the strides.  Every wavefront of the sweep reads and writes 64 consecutive x of
// one direction.
//
// AA pattern.  After an even sweep (and initially) the state is in "post-even"
// form: f_out(x, i) is stored at (x, opposite(i)).  An odd sweep gathers its
// incoming populations f_in(x, i) = f_out(x - c_i, i) from (x - c_i, opposite(i))
// and scatters f_out(x, i) to (x + c_i, i) ("streamed" form); an even sweep reads
// (x, i) and writes (x, opposite(i)).  Boundary kernels run between sweeps and
// write the incoming value of a boundary link exactly where the next sweep reads
// it, so the sweep itself has no boundary logic.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ak {

constexpr int NQ = 27;
__host__ __device__ constexpr int cxq(int i) { return i / 9 - 1; }
__host__ __device__ constexpr int cyq(int i) { return (i / 3) % 3 - 1; }
__host__ __device__ constexpr int czq(int i) { return i % 3 - 1; }
__host__ __device__ constexpr int opp(int i) { return 26 - i; }
__host__ __device__ constexpr double wq(int i) {
  return (cxq(i) * cxq(i) + cyq(i) * cyq(i) + czq(i) * czq(i)) == 0 ? 8.0 / 27.0
       : (cxq(i) * cxq(i) + cyq(i) * cyq(i) + czq(i) * czq(i)) == 1 ? 2.0 / 27.0
       : (cxq(i) * cxq(i) + cyq(i) * cyq(i) + czq(i) * czq(i)) == 2 ? 1.0 / 54.0
                                                                      : 1.0 / 216.0;
}

struct Desc {
  int nx, ny, nz, pad_;
  long long P, PS, QS, base;
};

// Storage traits: T = element type in HBM, C = compute type of the collision.
template <int S> struct St;
template <> struct St<0> { using T = double;   using C = double; };
template <> struct St<1> { using T = float;    using C = float; };
template <> struct St<2> { using T = _Float16; using C = float; };

template <int S>
__device__ __forceinline__ typename St<S>::C ld(const typename St<S>::T *p, float inv_scale) {
  if constexpr (S == 2) return (float)(*p) * inv_scale;
  else return *p;
}
template <int S>
__device__ __forceinline__ void stv(typename St<S>::T *p, typename St<S>::C v, float scale) {
  if constexpr (S == 2) *p = (_Float16)(v * scale);
  else *p = v;
}

__device__ __forceinline__ long long cell_of(const Desc &L, int x, int y, int z) {
  return (long long)z * L.PS + (long long)(y + 1) * L.P + x;
}

// Offset from cell (x,y,z) to its neighbour (x+cx, y+cy, z+cz); z wraps.  With PXY
// x and y wrap too (fully periodic boxes).
template <bool PXY>
__device__ __forceinline__ long long nb_off(const Desc &L, int x, int y, int z, int cx, int cy, int cz) {
  long long o;
  if constexpr (PXY) {
    int xn = x + cx, yn = y + cy;
    xn = xn < 0 ? xn + L.nx : (xn >= L.nx ? xn - L.nx : xn);
    yn = yn < 0 ? yn + L.ny : (yn >= L.ny ? yn - L.ny : yn);
    o = (long long)(xn - x) + (long long)(yn - y) * L.P;
  } else {
    o = (long long)cx + (long long)cy * L.P;
  }
  if (cz == 1) o += (z == L.nz - 1) ? -(long long)(L.nz - 1) * L.PS : L.PS;
  else if (cz == -1) o += (z == 0) ? (long long)(L.nz - 1) * L.PS : -L.PS;
  return o;
}

// Location of the post-collision population f_out(cell, j) in the current form.
template <bool PE>
__device__ __forceinline__ long long loc_out(const Desc &L, int x, int y, int z, int j) {
  const long long c = L.base + cell_of(L, x, y, z);
  if constexpr (PE) return c + (long long)opp(j) * L.QS;
  else return c + (long long)j * L.QS + nb_off<false>(L, x, y, z, cxq(j), cyq(j), czq(j));
}
// Location the next sweep reads as the incoming population i of fluid cell (x,y,z).
template <bool PE>
__device__ __forceinline__ long long loc_in(const Desc &L, int x, int y, int z, int i) {
  const long long c = L.base + cell_of(L, x, y, z);
  if constexpr (PE) return c + (long long)opp(i) * L.QS + nb_off<false>(L, x, y, z, -cxq(i), -cyq(i), -czq(i));
  else return c + (long long)i * L.QS;
}

// ---------------------------------------------------------------------------
// Compressible D3Q27 cumulant collision in shifted-population form.
//
// Input g = f - w (pre-collision), output g' = f' - w.  Rates: ws on the five
// deviatoric second-order cumulants, wb on the trace, 1 on every higher group,
// so the post-collision cumulants of order >= 3 vanish and the post-collision
// central moments are the Gaussian closure of the relaxed second moments.
// Everything is carried as deviations from the rest state (central moments
// of f minus those of w about the same velocity), so no O(1) quantity is
// subtracted from another: the arithmetic keeps the precision of g itself.
// Returns drho and |u|^2 for the fused diagnostics, and u (conserved by the collision).  The
// output populations are multiplied by out_scale (a power of two: the FP16 storage scale, or 1).
//
// The forward transform needs the raw moments up to second order only, by three
// one-dimensional passes.  The first pass (z lines of g[3l..3l+2] = (m, o, p)) gives
// Z0 = (p + m) + o, Z1 = p - m, Z2 = p + m.
template <typename C>
__device__ __forceinline__ void zlines(const C (&g)[NQ], C (&Z0)[9], C (&Z1)[9], C (&Z2)[9]) {
#pragma unroll
  for (int l = 0; l < 9; ++l) {
    const C m = g[3 * l + 0], o = g[3 * l + 1], p = g[3 * l + 2];
    const C s = p + m;
    Z0[l] = s + o;
    Z1[l] = p - m;
    Z2[l] = s;
  }
}

// FP32 sums of FP16 values in one instruction: v_fma_mix_f32 converts its FP16 operands
// exactly and rounds a * 1.0 + b once, which is bit for bit the FP32 sum of the converted
// values (a conversion of each operand and an add, three instructions, otherwise).
__device__ __forceinline__ float add_hh(_Float16 a, _Float16 b) {
  float r;
  asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,1]" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ float sub_hh(_Float16 a, _Float16 b) {
  float r;
  asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel_hi:[1,0,1]" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ float add_hf(_Float16 a, float b) {
  float r;
  asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ float sub_hf(_Float16 a, float b) {
  float r;
  asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
// The first pass straight from the stored FP16 values (still multiplied by the storage
// scale): every sum, like every later sum of the forward transform, is then exactly the
// scale times the sum of the unscaled values (power-of-two factor, far from overflow and
// underflow), so one multiply of each of the 10 raw moments by the inverse scale gives bit for
// bit what scaling the 27 populations first would.
__device__ __forceinline__ void zlines_h(const _Float16 (&h)[NQ], float (&Z0)[9], float (&Z1)[9], float (&Z2)[9]) {
#pragma unroll
  for (int l = 0; l < 9; ++l) {
    const _Float16 m = h[3 * l + 0], o = h[3 * l + 1], p = h[3 * l + 2];
    const float s = add_hh(p, m);
    Z0[l] = add_hf(o, s);
    Z1[l] = sub_hh(p, m);
    Z2[l] = s;
  }
}

// The collision from the first pass.  With SCALED the raw moments are multiplied by in_scale.
template <typename C, bool SCALED>
__device__ __forceinline__ void collide_z(const C (&Z0)[9], const C (&Z1)[9], const C (&Z2)[9], C in_scale,
                                          C (&g)[NQ], C ws, C wb, C &drho_out, C &u2_out, C (&u_out)[3],
                                          C out_scale) {
  C Y00[3], Y10[3], Y20[3], Y01[3], Y11[3], Y02[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int b = 3 * a;
    const C s0 = Z0[b + 2] + Z0[b + 0];
    Y00[a] = s0 + Z0[b + 1];
    Y10[a] = Z0[b + 2] - Z0[b + 0];
    Y20[a] = s0;
    Y01[a] = Z1[b + 0] + Z1[b + 1] + Z1[b + 2];
    Y11[a] = Z1[b + 2] - Z1[b + 0];
    Y02[a] = Z2[b + 0] + Z2[b + 1] + Z2[b + 2];
  }
  C drho = Y00[0] + Y00[1] + Y00[2];
  C jx = Y00[2] - Y00[0], sxx = Y00[2] + Y00[0];
  C jy = Y10[0] + Y10[1] + Y10[2], sxy = Y10[2] - Y10[0];
  C syy = Y20[0] + Y20[1] + Y20[2];
  C jz = Y01[0] + Y01[1] + Y01[2], sxz = Y01[2] - Y01[0];
  C syz = Y11[0] + Y11[1] + Y11[2];
  C szz = Y02[0] + Y02[1] + Y02[2];
  if constexpr (SCALED) {
    drho *= in_scale; jx *= in_scale; jy *= in_scale; jz *= in_scale;
    sxx *= in_scale; sxy *= in_scale; syy *= in_scale; sxz *= in_scale; syz *= in_scale; szz *= in_scale;
  }

  const C third = C(1) / C(3);
  // FP32: the hardware reciprocal (within 1 ulp) instead of the correctly rounded division
  // sequence (about ten instructions); FP64 keeps the division
  C ir;
  if constexpr (sizeof(C) == 4) ir = __builtin_amdgcn_rcpf(C(1) + drho);
  else ir = C(1) / (C(1) + drho);
  const C ux = jx * ir, uy = jy * ir, uz = jz * ir;
  // second central moments minus rho/3 (A) and off-diagonals (B)
  const C d3 = drho * third;
  const C Ax = sxx - d3 - jx * ux, Ay = syy - d3 - jy * uy, Az = szz - d3 - jz * uz;
  const C Bxy = sxy - jx * uy, Bxz = sxz - jx * uz, Byz = syz - jy * uz;
  // relaxation
  const C om = C(1) - ws;
  const C sA = (C(1) - wb) * (Ax + Ay + Az);
  const C dxy = om * (Ax - Ay), dxz = om * (Ax - Az);
  const C ax = (sA + dxy + dxz) * third;
  const C ay = (sA - C(2) * dxy + dxz) * third;
  const C az = (sA + dxy - C(2) * dxz) * third;
  const C bxy = om * Bxy, bxz = om * Bxz, byz = om * Byz;

  const C ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz;
  const C sx = third + ux2, sy = third + uy2, sz = third + uz2;
  const C ex = ax - ux2, ey = ay - uy2, ez = az - uz2;
  const C d9 = drho * (C(1) / C(9));
  // deviation post-collision central moments K[a][b][c], index 9a+3b+c
  C K[NQ];
  K[0] = drho;
  K[9] = ux; K[3] = uy; K[1] = uz;
  K[18] = d3 + ex; K[6] = d3 + ey; K[2] = d3 + ez;
  K[12] = bxy - ux * uy; K[10] = bxz - ux * uz; K[4] = byz - uy * uz;
  K[21] = sx * uy; K[19] = sx * uz; K[15] = sy * ux; K[7] = sy * uz; K[11] = sz * ux; K[5] = sz * uy;
  K[13] = ux * uy * uz;
  K[24] = d9 + (ex + ey) * third + (ax * ay + C(2) * bxy * bxy) * ir - ux2 * uy2;
  K[20] = d9 + (ex + ez) * third + (ax * az + C(2) * bxz * bxz) * ir - ux2 * uz2;
  K[8] = d9 + (ey + ez) * third + (ay * az + C(2) * byz * byz) * ir - uy2 * uz2;
  K[22] = K[4] * third + (ax * byz + C(2) * bxy * bxz) * ir - ux2 * uy * uz;
  K[16] = K[10] * third + (ay * bxz + C(2) * bxy * byz) * ir - uy2 * ux * uz;
  K[14] = K[12] * third + (az * bxy + C(2) * bxz * byz) * ir - uz2 * ux * uy;
  K[25] = sx * sy * uz; K[23] = sx * sz * uy; K[17] = sy * sz * ux;
  {
    const C ir2 = ir * ir;
    const C b2xy = bxy * bxy, b2xz = bxz * bxz, b2yz = byz * byz;
    K[26] = drho * (C(1) / C(27)) + (ax + ay + az) * (C(1) / C(9))
          + (ax * ay + ax * az + ay * az) * ir * third + ax * ay * az * ir2
          + (C(2) / C(3)) * (b2yz + b2xz + b2xy) * ir
          + C(2) * (ax * b2yz + ay * b2xz + az * b2xy) * ir2 + C(8) * bxy * bxz * byz * ir2
          - ((ux2 + uy2 + uz2) * (C(1) / C(9)) + (ux2 * uy2 + ux2 * uz2 + uy2 * uz2) * third + ux2 * uy2 * uz2);
  }
  // back-transform: per axis, central moments (k0,k1,k2) about u -> populations (-1,0,+1).  The
  // halves of the two outer populations of each line are not applied in the passes: every
  // value of one line carries the same power-of-two factor, so a population with n nonzero
  // velocity components comes out 2^n times too large, and one multiply by out_scale * 2^-n
  // at the end gives exactly what halving in every pass would (power-of-two scaling is exact).
  const C fac[4] = {out_scale, out_scale * C(0.5), out_scale * C(0.25), out_scale * C(0.125)};
  // one line's two raw moments from its central ones, m1 = k1 + u k0 and m2 = k2 + 2 u k1 + u^2 k0:
  // FP64 keeps the nested form; FP32 takes three fused multiply-adds with u^2 and 2u formed once
  // per axis (one instruction fewer per line, 27 per cell)
  auto raw = [](C u, C uu, C k0, C k1, C k2, C &m1, C &m2) {
    if constexpr (sizeof(C) == 4) {
      m1 = __builtin_fmaf(u, k0, k1);
      m2 = __builtin_fmaf(uu, k0, __builtin_fmaf(C(2) * u, k1, k2));
    } else {
      m1 = k1 + u * k0;
      m2 = k2 + u * (C(2) * k1 + u * k0);
    }
  };
#pragma unroll
  for (int l = 0; l < 9; ++l) {  // z lines: index 3l + c
    const C k0 = K[3 * l], k1 = K[3 * l + 1], k2 = K[3 * l + 2];
    C m1, m2;
    raw(uz, uz2, k0, k1, k2, m1, m2);
    K[3 * l + 1] = k0 - m2;
    K[3 * l + 2] = m2 + m1;
    K[3 * l + 0] = m2 - m1;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) {  // y lines: index 9a + 3b + c
      const int o = 9 * a + c;
      const C k0 = K[o], k1 = K[o + 3], k2 = K[o + 6];
      C m1, m2;
      raw(uy, uy2, k0, k1, k2, m1, m2);
      K[o + 3] = k0 - m2;
      K[o + 6] = m2 + m1;
      K[o + 0] = m2 - m1;
    }
#pragma unroll
  for (int o = 0; o < 9; ++o) {  // x lines: index 9a + o
    const C k0 = K[o], k1 = K[o + 9], k2 = K[o + 18];
    C m1, m2;
    raw(ux, ux2, k0, k1, k2, m1, m2);
    const int n = (o / 3 != 1) + (o % 3 != 1);  // nonzero y and z components of this line
    g[o + 9] = (k0 - m2) * fac[n];
    g[o + 18] = (m2 + m1) * fac[n + 1];
    g[o + 0] = (m2 - m1) * fac[n + 1];
  }
  drho_out = drho;
  u2_out = ux2 + uy2 + uz2;
  u_out[0] = ux; u_out[1] = uy; u_out[2] = uz;
}
template <typename C>
__device__ __forceinline__ void collide(C (&g)[NQ], C ws, C wb, C &drho_out, C &u2_out, C (&u_out)[3],
                                        C out_scale = C(1)) {
  C Z0[9], Z1[9], Z2[9];
  zlines<C>(g, Z0, Z1, Z2);
  collide_z<C, false>(Z0, Z1, Z2, C(1), g, ws, wb, drho_out, u2_out, u_out, out_scale);
}
// The collision of FP16-stored populations h (the shifted populations times the storage
// scale, 1 / inv_scale), bit for bit collide() of the populations loaded as FP32 and
// multiplied by inv_scale; the result goes to g.
__device__ __forceinline__ void collide_h(const _Float16 (&h)[NQ], float inv_scale, float (&g)[NQ], float ws,
                                          float wb, float &drho_out, float &u2_out, float (&u_out)[3],
                                          float out_scale) {
  float Z0[9], Z1[9], Z2[9];
  zlines_h(h, Z0, Z1, Z2);
  collide_z<float, true>(Z0, Z1, Z2, inv_scale, g, ws, wb, drho_out, u2_out, u_out, out_scale);
}
template <typename C>
__device__ __forceinline__ void collide(C (&g)[NQ], C ws, C wb, C &drho_out, C &u2_out) {
  C u[3];
  collide<C>(g, ws, wb, drho_out, u2_out, u);
}


// FP16 storage of a cell's 27 post-collision populations that keeps its mass and momentum.
// Rounding each shifted population to FP16 on its own changes the cell's sum and first
// moments by the sum of 27 rounding errors, a few units in the last place of the largest
// populations (~1e-5 at Ma 0.1); in a smooth field the errors repeat from cell to cell and
// step to step, a spurious source of mass and momentum.  Here the rounding deficits of the 19
// non-corner directions are moved onto the 8 corner directions (c = (+-1, +-1, +-1), the
// smallest populations, whose FP16 steps are the finest) as e_k = -(dm + c_k . dj) / 8, which
// restores the sum and the first moments exactly (sum_k c_k = 0, sum_k c_k c_k^T = 8 I); only
// the corners' own rounding remains.  The 19 others are stored exactly as plain rounding
// would store them.  No-op for FP64 and FP32 storage.
template <int S>
__device__ __forceinline__ void conserve_fp16(typename St<S>::C (&g)[NQ], float scale, float inv_scale) {
  if constexpr (S == 2) {
    // the deficits summed as fp16_conserving_each sums them
    auto deficit = [&](int i) { return (float)(_Float16)(g[i] * scale) * inv_scale - g[i]; };
    float dm = 0.f, dx = 0.f, dy = 0.f, dz = 0.f;
    constexpr int full[5] = {1, 3, 4, 5, 7}, ends[4] = {0, 2, 6, 8};
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const int o = full[k];
      const float a = deficit(o), b = deficit(o + 9), c = deficit(o + 18);
      const float line = a + b + c;
      dm += line;
      dx += c - a;
      if (cyq(o)) dy += (float)cyq(o) * line;
      if (czq(o)) dz += (float)czq(o) * line;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int o = ends[k];
      const float b = deficit(o + 9);
      dm += b;
      dy += (float)cyq(o) * b;
      dz += (float)czq(o) * b;
    }
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      if (cxq(i) == 0 || cyq(i) == 0 || czq(i) == 0) continue;
      g[i] -= 0.125f * (dm + (float)cxq(i) * dx + (float)cyq(i) * dy + (float)czq(i) * dz);
    }
  }
}

// The FP16 values conserve_fp16 followed by the storage rounding stores, computed from the
// scaled values s = g * scale (the collision's output with out_scale = scale; scale is a power
// of two, so each step gives the same result in either unit and each direction is rounded once).
// store(i, h) receives direction i's value as soon as it is final, the 19 non-corner
// directions first, then the corners.
template <class Store>
__device__ __forceinline__ void fp16_conserving_each(const float (&s)[NQ], Store &&store) {
  // rounding deficit (float)h - s of a non-corner direction, its value stored
  auto deficit = [&](int i) {
    const _Float16 hi = (_Float16)s[i];
    store(i, hi);
    return sub_hf(hi, s[i]);
  };
  // the deficits summed per x line (directions o, o + 9, o + 18 at c_x = -1, 0, +1): the five
  // lines without a corner first, then the middle values of the four lines that end in corners
  float dm = 0.f, dx = 0.f, dy = 0.f, dz = 0.f;
  constexpr int full[5] = {1, 3, 4, 5, 7}, ends[4] = {0, 2, 6, 8};
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const int o = full[k];
    const float a = deficit(o), b = deficit(o + 9), c = deficit(o + 18);
    const float line = a + b + c;
    dm += line;
    dx += c - a;
    if (cyq(o)) dy += (float)cyq(o) * line;
    if (czq(o)) dz += (float)czq(o) * line;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int o = ends[k];
    const float b = deficit(o + 9);
    dm += b;
    dy += (float)cyq(o) * b;
    dz += (float)czq(o) * b;
  }
#pragma unroll
  for (int i = 0; i < NQ; ++i) {
    if (cxq(i) == 0 || cyq(i) == 0 || czq(i) == 0) continue;
    // rounded to FP32 first, then to FP16, in every kernel: the empty asm keeps the compiler from
    // fusing the subtraction and the conversion into one mixed-precision instruction (a single
    // rounding), which it does or not depending on the surrounding code
    float c = s[i] - 0.125f * (dm + (float)cxq(i) * dx + (float)cyq(i) * dy + (float)czq(i) * dz);
    asm volatile("" : "+v"(c));
    store(i, (_Float16)c);
  }
}

// Product-form equilibrium of direction i (absolute, not shifted).
template <typename C>
__device__ __forceinline__ C phi1(int c, C q) {
  return c == 0 ? C(2) / C(3) - q * q : (C(1) / C(3) + q * q + C(c) * q) * C(0.5);
}
template <typename C>
__device__ __forceinline__ C equilibrium(int i, C r, C ux, C uy, C uz) {
  return r * phi1<C>(cxq(i), ux) * phi1<C>(cyq(i), uy) * phi1<C>(czq(i), uz);
}

}  // namespace ak
