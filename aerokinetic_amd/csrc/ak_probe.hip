// ak_probe.hip -- memory-pattern probe of the in-place AA sweep (measurement tool, not part
// of the step).
//
// One launch moves the sweep's exact population traffic with no collision: every cell loads
// its 27 values and stores them back, in the even form (read (x, q), write (x, opp q)) or the
// odd form (read (x - c_q, opp q), write (x + c_q, q)), over a fully periodic box (every
// neighbour index wraps, so every access stays inside the buffer).  Each thread reads exactly
// the slots it writes, as in the sweep, so a launch has no races.  Layouts:
//   0 rows       interleaved direction rows [z][y][q][x] (the sweep's layout), row pitch a
//                multiple of 128 bytes
//   1 planes     one plane per direction [q][z][y][x], same pitch
//   2 xblocks64  blocks of 64 cells along x, directions inside: [z][y][x/64][q][64]
//   3 blocks1024 blocks of 1024 consecutive cells in (z, y, x) order: [n/1024][q][1024]
// Threads: 64 x 4 tiles of one z plane with the sweep's XCD-contiguous block order, or (cells
// per lane k = 2 or 4, even form, layouts 0-2) 64 lanes x k consecutive cells as one k-wide
// access per direction.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <vector>
#include <cstdint>
#include <cstdio>

namespace {

__host__ __device__ constexpr int pcx(int i) { return i / 9 - 1; }
__host__ __device__ constexpr int pcy(int i) { return (i / 3) % 3 - 1; }
__host__ __device__ constexpr int pcz(int i) { return i % 3 - 1; }
__host__ __device__ constexpr int popp(int i) { return 26 - i; }

struct Box {
  int nx, ny, nz;
  long long R;  // row pitch (elements), layouts 0 and 1
  int gx, gy, nblocks;
};

__device__ __forceinline__ int wrapi(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

template <int LAYOUT>
__device__ __forceinline__ long long addr(const Box &B, int x, int y, int z, int q) {
  if constexpr (LAYOUT == 0) return (((long long)z * B.ny + y) * 27 + q) * B.R + x;
  else if constexpr (LAYOUT == 1) return (((long long)q * B.nz + z) * B.ny + y) * B.R + x;
  else if constexpr (LAYOUT == 2) {
    const int nxb = B.nx / 64;
    return ((((long long)z * B.ny + y) * nxb + x / 64) * 27 + q) * 64 + (x % 64);
  } else {
    const long long n = ((long long)z * B.ny + y) * B.nx + x;
    return ((n / 1024) * 27 + q) * 1024 + (n % 1024);
  }
}

__device__ __forceinline__ int xcd_order(int b, int nb) {
  const int per = nb / 8, rem = nb % 8, xcd = b % 8, k = b / 8;
  return xcd < rem ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
}

template <typename T, int LAYOUT, bool ODD>
__global__ __launch_bounds__(256) void kp_pattern(T *__restrict__ A, Box B) {
  const int b = xcd_order((int)blockIdx.x, B.nblocks);
  const int tx = b % B.gx, ty = (b / B.gx) % B.gy, z = b / (B.gx * B.gy);
  const int x = tx * 64 + (int)(threadIdx.x % 64), y = ty * 4 + (int)(threadIdx.x / 64);
  if (x >= B.nx || y >= B.ny) return;
  T v[27];
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    if constexpr (ODD)
      v[q] = A[addr<LAYOUT>(B, wrapi(x - pcx(q), B.nx), wrapi(y - pcy(q), B.ny), wrapi(z - pcz(q), B.nz), popp(q))];
    else
      v[q] = A[addr<LAYOUT>(B, x, y, z, q)];
  }
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    if constexpr (ODD)
      A[addr<LAYOUT>(B, wrapi(x + pcx(q), B.nx), wrapi(y + pcy(q), B.ny), wrapi(z + pcz(q), B.nz), q)] = v[q];
    else
      A[addr<LAYOUT>(B, x, y, z, popp(q))] = v[q];
  }
}

// Even form with K consecutive cells per lane: one K-element access per direction (W is an
// integer type of K elements).  The cell groups never straddle a row or a 64-cell block.
template <typename W, int LAYOUT, int K>
__global__ __launch_bounds__(256) void kp_pattern_wide(W *__restrict__ A, Box B) {
  const int b = xcd_order((int)blockIdx.x, B.nblocks);
  const int tx = b % B.gx, ty = (b / B.gx) % B.gy, z = b / (B.gx * B.gy);
  const int x = (tx * 64 + (int)(threadIdx.x % 64)) * K, y = ty * 4 + (int)(threadIdx.x / 64);
  if (x >= B.nx || y >= B.ny) return;
  W v[27];
#pragma unroll
  for (int q = 0; q < 27; ++q) v[q] = A[addr<LAYOUT>(B, x, y, z, q) / K];
#pragma unroll
  for (int q = 0; q < 27; ++q) A[addr<LAYOUT>(B, x, y, z, popp(q)) / K] = v[q];
}

template <typename T, int LAYOUT>
hipError_t launch_one(T *A, const Box &B, bool odd, hipStream_t st) {
  if (odd) hipLaunchKernelGGL((kp_pattern<T, LAYOUT, true>), dim3(B.nblocks), dim3(256), 0, st, A, B);
  else hipLaunchKernelGGL((kp_pattern<T, LAYOUT, false>), dim3(B.nblocks), dim3(256), 0, st, A, B);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_layout(void *A, const Box &B, int layout, bool odd, hipStream_t st) {
  switch (layout) {
    case 0: return launch_one<T, 0>((T *)A, B, odd, st);
    case 1: return launch_one<T, 1>((T *)A, B, odd, st);
    case 2: return launch_one<T, 2>((T *)A, B, odd, st);
    default: return launch_one<T, 3>((T *)A, B, odd, st);
  }
}

template <typename W, int K>
hipError_t launch_wide(void *A, const Box &B, int layout, hipStream_t st) {
  switch (layout) {
    case 0: hipLaunchKernelGGL((kp_pattern_wide<W, 0, K>), dim3(B.nblocks), dim3(256), 0, st, (W *)A, B); break;
    case 1: hipLaunchKernelGGL((kp_pattern_wide<W, 1, K>), dim3(B.nblocks), dim3(256), 0, st, (W *)A, B); break;
    default: hipLaunchKernelGGL((kp_pattern_wide<W, 2, K>), dim3(B.nblocks), dim3(256), 0, st, (W *)A, B); break;
  }
  return hipGetLastError();
}

// The rows layout with the sweep's own row pitch (nx + 2 x-ghosts, rounded up to a 128-byte
// line) and any tile of 256 threads (bx along x, 256 / bx rows along y), in the sweep's
// XCD-contiguous block order.  Same traffic as kp_pattern<T, 0, ODD>.
struct Box2 {
  int nx, ny, nz, bx, by, gx, gy, nblocks;
  long long R;
  int chunk;  // block order: 0 = one contiguous slab per XCD, -1 = hardware order, C > 0 = chunks of C
};
// Tile of hardware block b: with chunk C > 0 the tiles are dealt to the 8 XCDs (b % 8) in
// chunks of C consecutive tiles, XCD k taking chunks k, k + 8, ...; a tail shorter than 8 C
// tiles keeps the contiguous-slab order among its own tiles.
__device__ __forceinline__ int tile_order(int b, int nb, int chunk) {
  if (chunk < 0) return b;
  if (chunk == 0) return xcd_order(b, nb);
  const int span = 8 * chunk, full = nb / span * span;
  if (b >= full) return full + xcd_order(b - full, nb - full);
  const int xcd = b % 8, k = b / 8;
  return (k / chunk) * span + xcd * chunk + k % chunk;
}
template <typename T, bool ODD>
__global__ __launch_bounds__(256) void kp_rows_tile(T *__restrict__ A, Box2 B) {
  const int b = tile_order((int)blockIdx.x, B.nblocks, B.chunk);
  const int tx = b % B.gx, ty = (b / B.gx) % B.gy, z = b / (B.gx * B.gy);
  const int x = tx * B.bx + (int)(threadIdx.x % B.bx), y = ty * B.by + (int)(threadIdx.x / B.bx);
  if (x >= B.nx || y >= B.ny) return;
  auto at = [&](int xx, int yy, int zz, int q) { return (((long long)zz * B.ny + yy) * 27 + q) * B.R + xx; };
  T v[27];
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    if constexpr (ODD)
      v[q] = A[at(wrapi(x - pcx(q), B.nx), wrapi(y - pcy(q), B.ny), wrapi(z - pcz(q), B.nz), popp(q))];
    else
      v[q] = A[at(x, y, z, q)];
  }
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    if constexpr (ODD)
      A[at(wrapi(x + pcx(q), B.nx), wrapi(y + pcy(q), B.ny), wrapi(z + pcz(q), B.nz), q)] = v[q];
    else
      A[at(x, y, z, popp(q))] = v[q];
  }
}

template <typename T>
hipError_t launch_rows_tile(void *A, const Box2 &B, bool odd, size_t lds, hipStream_t st) {
  if (odd) hipLaunchKernelGGL((kp_rows_tile<T, true>), dim3(B.nblocks), dim3(256), lds, st, (T *)A, B);
  else hipLaunchKernelGGL((kp_rows_tile<T, false>), dim3(B.nblocks), dim3(256), lds, st, (T *)A, B);
  return hipGetLastError();
}

// Streaming read of one region by the blocks of one XCD only (hardware block b runs on XCD
// b % 8): every other block returns at once.  The xor of what a lane read goes to `sink`
// only when it equals a value no zero-filled region produces, so the loads stay live.
__global__ __launch_bounds__(256) void kp_xcd_read(const uint4 *__restrict__ p, size_t n16, int xcd,
                                                   unsigned *__restrict__ sink) {
  if ((int)(blockIdx.x % 8) != xcd) return;
  const size_t j = blockIdx.x / 8, nact = gridDim.x / 8;
  unsigned acc = 0;
  for (size_t i = j * 256 + threadIdx.x; i < n16; i += nact * 256) {
    const uint4 v = p[i];
    acc ^= v.x ^ v.y ^ v.z ^ v.w;
  }
  if (acc == 0x9e3779b9u) sink[0] = acc;
}

char g_err[256] = "";

}  // namespace

extern "C" {

const char *akp_last_error() { return g_err; }

// Mean milliseconds per launch of one pattern over an nx x ny x nz box of elem_bytes-byte
// elements (2, 4 or 8), `reps` timed launches after two untimed ones, on a buffer of its own.
// cells_per_lane 1 (both forms, every layout) or 2 / 4 (even form, layouts 0-2, nx a multiple
// of 64 x cells_per_lane).  out = {ms per launch, bytes per launch}.  Returns 0, or -1 with
// akp_last_error() set.
int akp_pattern(int layout, int elem_bytes, int odd, int cells_per_lane, int nx, int ny, int nz, int reps,
                double *out) {
  if (layout < 0 || layout > 3 || (elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8) || nx < 2 || ny < 2 ||
      nz < 2 || reps < 1 || !out) {
    snprintf(g_err, sizeof g_err, "bad probe arguments");
    return -1;
  }
  const int k = cells_per_lane;
  if (!(k == 1 || ((k == 2 || k == 4) && !odd && layout <= 2 && nx % (64 * k) == 0 && elem_bytes * k <= 8))) {
    snprintf(g_err, sizeof g_err, "cells_per_lane %d needs the even form, layouts 0-2, nx a multiple of %d", k, 64 * k);
    return -1;
  }
  if (layout == 2 && nx % 64) {
    snprintf(g_err, sizeof g_err, "layout 2 needs nx a multiple of 64");
    return -1;
  }
  Box B{};
  B.nx = nx; B.ny = ny; B.nz = nz;
  const int per_line = 128 / elem_bytes;
  B.R = ((long long)nx + per_line - 1) / per_line * per_line;
  B.gx = (nx / k + 63) / 64;
  B.gy = (ny + 3) / 4;
  B.nblocks = B.gx * B.gy * nz;
  long long elems;
  if (layout <= 1) elems = B.R * ny * nz * 27;
  else if (layout == 2) elems = (long long)nx * ny * nz * 27;
  else elems = ((long long)nx * ny * nz + 1023) / 1024 * 1024 * 27;
  const size_t bytes = (size_t)elems * elem_bytes;
  void *A = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&A, bytes);
  if (e == hipSuccess) e = hipMemset(A, 0, bytes);
  if (e == hipSuccess) e = hipStreamCreate(&st);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  auto launch = [&]() -> hipError_t {
    if (k == 1) {
      switch (elem_bytes) {
        case 2: return launch_layout<uint16_t>(A, B, layout, odd != 0, st);
        case 4: return launch_layout<uint32_t>(A, B, layout, odd != 0, st);
        default: return launch_layout<uint64_t>(A, B, layout, odd != 0, st);
      }
    }
    if (elem_bytes * k == 4) return launch_wide<uint32_t, 2>(A, B, layout, st);  // 2 x FP16
    if (k == 2) return launch_wide<uint64_t, 2>(A, B, layout, st);              // 2 x FP32
    return launch_wide<uint64_t, 4>(A, B, layout, st);                          // 4 x FP16
  };
  for (int i = 0; i < 2 && e == hipSuccess; ++i) e = launch();
  if (e == hipSuccess) e = hipEventRecord(e0, st);
  for (int i = 0; i < reps && e == hipSuccess; ++i) e = launch();
  if (e == hipSuccess) e = hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (A) (void)hipFree(A);
  if (e != hipSuccess) {
    snprintf(g_err, sizeof g_err, "%s", hipGetErrorString(e));
    return -1;
  }
  out[0] = ms / reps;
  out[1] = 2.0 * 27.0 * elem_bytes * (double)nx * ny * nz;
  return 0;
}

// Placement and occupancy: `nbuf` population buffers of an nx x ny x nz box (rows layout, the
// sweep's row pitch), all allocated and held at once, each timed over `reps` launches of the
// even and of the odd form with tiles of bx x (256 / bx) threads and at most `waves` resident
// waves per SIMD (0: no cap; the cap reserves LDS per block, which the kernel does not use).
// out[2k] / out[2k + 1] = ms per even / odd launch on buffer k; out[2 nbuf] = bytes per launch.
// Returns the number of buffers timed (allocation stops early when memory runs out), or -1.
int akp_placements(int elem_bytes, int nx, int ny, int nz, int bx, int waves, int nbuf, int reps, double *out) {
  if ((elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8) || nx < 2 || ny < 2 || nz < 2 || reps < 1 || !out ||
      nbuf < 1 || nbuf > 16 || (bx != 64 && bx != 128 && bx != 256) || waves < 0 || waves > 8 ||
      (waves > 0 && waves < 3)) {
    snprintf(g_err, sizeof g_err, "bad placement-probe arguments");
    return -1;
  }
  Box2 B{};
  B.nx = nx; B.ny = ny; B.nz = nz; B.bx = bx; B.by = 256 / bx;
  const int per_line = 128 / elem_bytes;
  B.R = ((long long)nx + 2 + per_line - 1) / per_line * per_line;
  B.gx = (nx + bx - 1) / bx;
  B.gy = (ny + B.by - 1) / B.by;
  B.nblocks = B.gx * B.gy * nz;
  const size_t bytes = (size_t)B.R * ny * nz * 27 * elem_bytes;
  int lds_per_cu = 0;
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&lds_per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess)
    lds_per_cu = 0;
  const size_t lds = waves > 0 && lds_per_cu > 0 ? (size_t)(lds_per_cu / waves) & ~(size_t)255 : 0;
  void *held[16] = {};
  int got = 0;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipStreamCreate(&st);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  for (int k = 0; k < nbuf && e == hipSuccess; ++k) {
    if (hipMalloc(&held[k], bytes) != hipSuccess) {
      (void)hipGetLastError();  // out of memory: time the buffers held so far
      held[k] = nullptr;
      break;
    }
    e = hipMemset(held[k], 0, bytes);
    for (int odd = 0; odd < 2 && e == hipSuccess; ++odd) {
      auto launch = [&]() -> hipError_t {
        switch (elem_bytes) {
          case 2: return launch_rows_tile<uint16_t>(held[k], B, odd != 0, lds, st);
          case 4: return launch_rows_tile<uint32_t>(held[k], B, odd != 0, lds, st);
          default: return launch_rows_tile<uint64_t>(held[k], B, odd != 0, lds, st);
        }
      };
      for (int i = 0; i < 2 && e == hipSuccess; ++i) e = launch();
      if (e == hipSuccess) e = hipEventRecord(e0, st);
      for (int i = 0; i < reps && e == hipSuccess; ++i) e = launch();
      if (e == hipSuccess) e = hipEventRecord(e1, st);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      out[2 * k + odd] = ms / reps;
    }
    if (e == hipSuccess) got = k + 1;
  }
  for (int k = 0; k < 16; ++k)
    if (held[k]) (void)hipFree(held[k]);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (e != hipSuccess) {
    snprintf(g_err, sizeof g_err, "%s", hipGetErrorString(e));
    return -1;
  }
  out[2 * nbuf] = 2.0 * 27.0 * elem_bytes * (double)nx * ny * nz;
  return got;
}


// The same pattern on `nbuf` buffers (held at once, each allocated `max offset` bytes larger),
// each timed with the box starting at every one of `noff` byte offsets into the buffer
// (multiples of 256): does a buffer's rate depend on where in its pages the box starts?
// out[(k * noff + j) * 2 + odd] = ms per launch (buffer k, offset j); returns the number of
// buffers timed, or -1 with akp_last_error() set.
int akp_offsets(int elem_bytes, int nx, int ny, int nz, int bx, int nbuf, int noff, const long long *offsets,
                int reps, double *out) {
  if ((elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8) || nx < 2 || ny < 2 || nz < 2 || reps < 1 || !out ||
      !offsets || nbuf < 1 || nbuf > 16 || noff < 1 || noff > 64 || (bx != 64 && bx != 128 && bx != 256)) {
    snprintf(g_err, sizeof g_err, "bad offset-probe arguments");
    return -1;
  }
  long long max_off = 0;
  for (int j = 0; j < noff; ++j) {
    if (offsets[j] < 0 || offsets[j] % 256) {
      snprintf(g_err, sizeof g_err, "offsets must be non-negative multiples of 256");
      return -1;
    }
    max_off = offsets[j] > max_off ? offsets[j] : max_off;
  }
  Box2 B{};
  B.nx = nx; B.ny = ny; B.nz = nz; B.bx = bx; B.by = 256 / bx;
  const int per_line = 128 / elem_bytes;
  B.R = ((long long)nx + 2 + per_line - 1) / per_line * per_line;
  B.gx = (nx + bx - 1) / bx;
  B.gy = (ny + B.by - 1) / B.by;
  B.nblocks = B.gx * B.gy * nz;
  const size_t bytes = (size_t)B.R * ny * nz * 27 * elem_bytes + (size_t)max_off;
  void *held[16] = {};
  int got = 0;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipStreamCreate(&st);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  for (int k = 0; k < nbuf && e == hipSuccess; ++k) {
    if (hipMalloc(&held[k], bytes) != hipSuccess) {
      (void)hipGetLastError();  // out of memory: time the buffers held so far
      held[k] = nullptr;
      break;
    }
    e = hipMemset(held[k], 0, bytes);
    for (int j = 0; j < noff && e == hipSuccess; ++j) {
      void *A = (char *)held[k] + offsets[j];
      for (int odd = 0; odd < 2 && e == hipSuccess; ++odd) {
        auto launch = [&]() -> hipError_t {
          switch (elem_bytes) {
            case 2: return launch_rows_tile<uint16_t>(A, B, odd != 0, 0, st);
            case 4: return launch_rows_tile<uint32_t>(A, B, odd != 0, 0, st);
            default: return launch_rows_tile<uint64_t>(A, B, odd != 0, 0, st);
          }
        };
        for (int i = 0; i < 2 && e == hipSuccess; ++i) e = launch();
        if (e == hipSuccess) e = hipEventRecord(e0, st);
        for (int i = 0; i < reps && e == hipSuccess; ++i) e = launch();
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        out[((size_t)k * noff + j) * 2 + odd] = ms / reps;
      }
    }
    if (e == hipSuccess) got = k + 1;
  }
  for (int k = 0; k < 16; ++k)
    if (held[k]) (void)hipFree(held[k]);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (e != hipSuccess) {
    snprintf(g_err, sizeof g_err, "%s", hipGetErrorString(e));
    return -1;
  }
  return got;
}


// The same pattern on `nbuf` buffers made with the virtual-memory interface: each buffer is
// physical memory of `chunk_bytes`-sized pieces (rounded to the allocation granularity; 0 =
// one piece) mapped into a reserved address range, and the same physical pieces mapped a
// second time into another range.  Separates what the rate follows: the physical memory (both
// ranges alike) or the addresses it is reached through.  out[4k + {0,1}] = ms per launch
// (even, odd) through the first range of buffer k, out[4k + {2,3}] through the second;
// out[4 nbuf] = the granularity.  Returns the number of buffers timed, or -1.
int akp_vmm(int elem_bytes, int nx, int ny, int nz, int bx, int nbuf, long long chunk_bytes, int ranges, int reps,
            double *out) {
  if ((elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8) || nx < 2 || ny < 2 || nz < 2 || reps < 1 || !out ||
      nbuf < 1 || nbuf > 16 || chunk_bytes < 0 || (bx != 64 && bx != 128 && bx != 256) || ranges < 1 ||
      ranges > 2) {
    snprintf(g_err, sizeof g_err, "bad vmm-probe arguments");
    return -1;
  }
  Box2 B{};
  B.nx = nx; B.ny = ny; B.nz = nz; B.bx = bx; B.by = 256 / bx;
  const int per_line = 128 / elem_bytes;
  B.R = ((long long)nx + 2 + per_line - 1) / per_line * per_line;
  B.gx = (nx + bx - 1) / bx;
  B.gy = (ny + B.by - 1) / B.by;
  B.nblocks = B.gx * B.gy * nz;
  const size_t need = (size_t)B.R * ny * nz * 27 * elem_bytes;
  int dev = 0;
  (void)hipGetDevice(&dev);
  hipMemAllocationProp prop = {};
  prop.type = hipMemAllocationTypePinned;
  prop.location.type = hipMemLocationTypeDevice;
  prop.location.id = dev;
  size_t gran = 0;
  hipError_t e = hipMemGetAllocationGranularity(&gran, &prop, hipMemAllocationGranularityRecommended);
  if (e != hipSuccess || gran == 0) {
    snprintf(g_err, sizeof g_err, "granularity: %s", hipGetErrorString(e));
    return -1;
  }
  const size_t chunk = chunk_bytes == 0 ? (need + gran - 1) / gran * gran
                                        : ((size_t)chunk_bytes + gran - 1) / gran * gran;
  const size_t size = (need + chunk - 1) / chunk * chunk;
  const size_t npieces = size / chunk;
  struct Buf { void *va[2]; std::vector<hipMemGenericAllocationHandle_t> h; size_t mapped[2]; };
  std::vector<Buf> bufs;
  int got = 0;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  e = hipStreamCreate(&st);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  for (int k = 0; k < nbuf && e == hipSuccess; ++k) {
    bufs.push_back(Buf{{nullptr, nullptr}, {}, {0, 0}});
    Buf &b = bufs.back();
    bool ok = true;
    hipError_t ve = hipSuccess;
    const char *what = "";
    for (size_t i = 0; i < npieces && ok; ++i) {
      hipMemGenericAllocationHandle_t h{};
      if ((ve = hipMemCreate(&h, chunk, &prop, 0)) != hipSuccess) { ok = false; what = "hipMemCreate"; }
      else b.h.push_back(h);
    }
    for (int r = 0; r < ranges && ok; ++r) {
      if ((ve = hipMemAddressReserve(&b.va[r], size, 0, nullptr, 0)) != hipSuccess) {
        b.va[r] = nullptr; ok = false; what = "hipMemAddressReserve"; break;
      }
      for (size_t i = 0; i < npieces && ok; ++i) {
        if ((ve = hipMemMap((char *)b.va[r] + i * chunk, chunk, 0, b.h[i], 0)) != hipSuccess) { ok = false; what = "hipMemMap"; }
        else b.mapped[r] = i + 1;
      }
      if (!ok) break;
      hipMemAccessDesc acc = {};
      acc.location = prop.location;
      acc.flags = hipMemAccessFlagsProtReadWrite;
      if ((ve = hipMemSetAccess(b.va[r], size, &acc, 1)) != hipSuccess) { ok = false; what = "hipMemSetAccess"; }
    }
    if (!ok) {  // out of memory or no second mapping: time the buffers made so far
      snprintf(g_err, sizeof g_err, "buffer %d: %s failed after %zu pieces: %s", k, what, b.h.size(),
               hipGetErrorString(ve));
      (void)hipGetLastError();
      break;
    }
    e = hipMemset(b.va[0], 0, need);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    for (int r = 0; r < ranges && e == hipSuccess; ++r)
      for (int odd = 0; odd < 2 && e == hipSuccess; ++odd) {
        void *A = b.va[r];
        auto launch = [&]() -> hipError_t {
          switch (elem_bytes) {
            case 2: return launch_rows_tile<uint16_t>(A, B, odd != 0, 0, st);
            case 4: return launch_rows_tile<uint32_t>(A, B, odd != 0, 0, st);
            default: return launch_rows_tile<uint64_t>(A, B, odd != 0, 0, st);
          }
        };
        for (int i = 0; i < 2 && e == hipSuccess; ++i) e = launch();
        if (e == hipSuccess) e = hipEventRecord(e0, st);
        for (int i = 0; i < reps && e == hipSuccess; ++i) e = launch();
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        out[4 * k + 2 * r + odd] = ms / reps;
      }
    if (e == hipSuccess) got = k + 1;
  }
  (void)hipDeviceSynchronize();
  for (Buf &b : bufs) {
    for (int r = 0; r < 2; ++r) {
      for (size_t i = 0; i < b.mapped[r]; ++i) (void)hipMemUnmap((char *)b.va[r] + i * chunk, chunk);
      if (b.va[r]) (void)hipMemAddressFree(b.va[r], size);
    }
    for (auto h : b.h) (void)hipMemRelease(h);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (e != hipSuccess) {
    snprintf(g_err, sizeof g_err, "%s", hipGetErrorString(e));
    return -1;
  }
  out[4 * nbuf] = (double)gran;
  return got;
}


// Screening physical memory piece by piece: `npieces` pieces of `piece_bytes` (virtual-memory
// interface), each timed alone with an nx x ny x nz_piece box in it; then one buffer mapped from
// the fastest pieces and one from the slowest (as many as an nx x ny x nz_full box needs, when
// twice that many were made), each timed with the full box.  Does a piece's own rate predict
// the rate of a buffer made of such pieces?  out[2i + {0,1}] = ms per launch (even, odd) of
// piece i; out[2 npieces + {0,1}] = fastest-piece buffer, out[2 npieces + {2,3}] = slowest-piece
// buffer (0 when not made).  Returns the number of pieces made, or -1.
int akp_screen(int elem_bytes, int nx, int ny, int nz_piece, int nz_full, int bx, int npieces, long long piece_bytes,
               int reps, double *out) {
  if ((elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8) || nx < 2 || ny < 2 || nz_piece < 2 || nz_full < 2 ||
      reps < 1 || !out || npieces < 1 || npieces > 512 || piece_bytes <= 0 || (bx != 64 && bx != 128 && bx != 256)) {
    snprintf(g_err, sizeof g_err, "bad screen-probe arguments");
    return -1;
  }
  auto box = [&](int nz) {
    Box2 B{};
    B.nx = nx; B.ny = ny; B.nz = nz; B.bx = bx; B.by = 256 / bx;
    const int per_line = 128 / elem_bytes;
    B.R = ((long long)nx + 2 + per_line - 1) / per_line * per_line;
    B.gx = (nx + bx - 1) / bx;
    B.gy = (ny + B.by - 1) / B.by;
    B.nblocks = B.gx * B.gy * nz;
    return B;
  };
  const Box2 Bp = box(nz_piece), Bf = box(nz_full);
  const size_t need_p = (size_t)Bp.R * ny * nz_piece * 27 * elem_bytes;
  const size_t need_f = (size_t)Bf.R * ny * nz_full * 27 * elem_bytes;
  int dev = 0;
  (void)hipGetDevice(&dev);
  hipMemAllocationProp prop = {};
  prop.type = hipMemAllocationTypePinned;
  prop.location.type = hipMemLocationTypeDevice;
  prop.location.id = dev;
  size_t gran = 0;
  hipError_t e = hipMemGetAllocationGranularity(&gran, &prop, hipMemAllocationGranularityRecommended);
  if (e != hipSuccess || gran == 0) {
    snprintf(g_err, sizeof g_err, "granularity: %s", hipGetErrorString(e));
    return -1;
  }
  const size_t piece = ((size_t)piece_bytes + gran - 1) / gran * gran;
  if (need_p > piece) {
    snprintf(g_err, sizeof g_err, "the piece box (%zu bytes) does not fit a piece (%zu)", need_p, piece);
    return -1;
  }
  const size_t m = (need_f + piece - 1) / piece;  // pieces per full buffer
  std::vector<hipMemGenericAllocationHandle_t> h;
  hipMemAccessDesc acc = {};
  acc.location = prop.location;
  acc.flags = hipMemAccessFlagsProtReadWrite;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  e = hipStreamCreate(&st);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  // times one box on a mapped range; false on a launch or timing error (kept in e)
  auto time_box = [&](void *A, const Box2 &B, double *ms2) {
    for (int odd = 0; odd < 2 && e == hipSuccess; ++odd) {
      auto launch = [&]() -> hipError_t {
        switch (elem_bytes) {
          case 2: return launch_rows_tile<uint16_t>(A, B, odd != 0, 0, st);
          case 4: return launch_rows_tile<uint32_t>(A, B, odd != 0, 0, st);
          default: return launch_rows_tile<uint64_t>(A, B, odd != 0, 0, st);
        }
      };
      for (int i = 0; i < 2 && e == hipSuccess; ++i) e = launch();
      if (e == hipSuccess) e = hipEventRecord(e0, st);
      for (int i = 0; i < reps && e == hipSuccess; ++i) e = launch();
      if (e == hipSuccess) e = hipEventRecord(e1, st);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      ms2[odd] = ms / reps;
    }
    return e == hipSuccess;
  };
  // maps pieces idx[0..n) in order into a fresh range; nullptr on failure (nothing left mapped)
  char why[160] = "";
  auto map_range = [&](const std::vector<size_t> &idx, size_t *mapped) -> void * {
    void *va = nullptr;
    *mapped = 0;
    hipError_t r = hipMemAddressReserve(&va, idx.size() * piece, 0, nullptr, 0);
    if (r != hipSuccess) {
      snprintf(why, sizeof why, "hipMemAddressReserve(%zu): %s", idx.size() * piece, hipGetErrorString(r));
      return nullptr;
    }
    for (size_t i = 0; i < idx.size(); ++i) {
      if ((r = hipMemMap((char *)va + i * piece, piece, 0, h[idx[i]], 0)) != hipSuccess) {
        snprintf(why, sizeof why, "hipMemMap(piece %zu): %s", i, hipGetErrorString(r));
        break;
      }
      *mapped = i + 1;
    }
    if (*mapped == idx.size()) {
      if ((r = hipMemSetAccess(va, idx.size() * piece, &acc, 1)) == hipSuccess) return va;
      snprintf(why, sizeof why, "hipMemSetAccess: %s", hipGetErrorString(r));
    }
    for (size_t i = 0; i < *mapped; ++i) (void)hipMemUnmap((char *)va + i * piece, piece);
    (void)hipMemAddressFree(va, idx.size() * piece);
    *mapped = 0;
    return nullptr;
  };
  auto unmap_range = [&](void *va, size_t n) {
    for (size_t i = 0; i < n; ++i) (void)hipMemUnmap((char *)va + i * piece, piece);
    (void)hipMemAddressFree(va, n * piece);
  };
  for (int i = 0; i < npieces; ++i) {
    hipMemGenericAllocationHandle_t hh{};
    if (hipMemCreate(&hh, piece, &prop, 0) != hipSuccess) {
      (void)hipGetLastError();
      break;
    }
    h.push_back(hh);
  }
  const int made = (int)h.size();
  out[2 * npieces + 0] = out[2 * npieces + 1] = out[2 * npieces + 2] = out[2 * npieces + 3] = 0;
  // every piece in one range, each timed on its own sub-range
  std::vector<size_t> all(made);
  for (int i = 0; i < made; ++i) all[i] = i;
  size_t mapped_all = 0;
  void *va_all = made > 0 && e == hipSuccess ? map_range(all, &mapped_all) : nullptr;
  if (!va_all) e = hipErrorOutOfMemory;
  if (e == hipSuccess) e = hipMemset(va_all, 0, (size_t)made * piece);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  for (int i = 0; i < made && e == hipSuccess; ++i) time_box((char *)va_all + (size_t)i * piece, Bp, out + 2 * i);
  (void)hipDeviceSynchronize();
  if (e == hipSuccess && (size_t)made >= 2 * m) {
    std::vector<size_t> order(all);
    std::sort(order.begin(), order.end(),
              [&](size_t a, size_t b) { return out[2 * a] + out[2 * a + 1] < out[2 * b] + out[2 * b + 1]; });
    for (int which = 0; which < 2 && e == hipSuccess; ++which) {
      std::vector<size_t> idx(order.begin() + (which ? made - m : 0), order.begin() + (which ? made : m));
      size_t mapped = 0;
      void *va = map_range(idx, &mapped);
      if (!va) { e = hipErrorOutOfMemory; break; }
      e = hipMemset(va, 0, need_f);
      if (e == hipSuccess) e = hipDeviceSynchronize();
      if (e == hipSuccess) time_box(va, Bf, out + 2 * npieces + 2 * which);
      (void)hipDeviceSynchronize();
      unmap_range(va, mapped);
    }
  }
  if (va_all) unmap_range(va_all, mapped_all);
  for (auto hh : h) (void)hipMemRelease(hh);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (e != hipSuccess) {
    snprintf(g_err, sizeof g_err, "%s (%s; %d pieces made)", hipGetErrorString(e), why, made);
    return -1;
  }
  return made;
}


// The pattern (256-thread tiles bx x 256/bx) on `nbuf` buffers held at once, each timed under
// every block order of `orders` (see tile_order).  out[(k * norders + j) * 2 + odd] = ms per
// launch.  Returns the number of buffers timed, or -1.
int akp_orders(int elem_bytes, int nx, int ny, int nz, int bx, int nbuf, int norders, const int *orders, int reps,
               double *out) {
  if ((elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8) || nx < 2 || ny < 2 || nz < 2 || reps < 1 || !out ||
      !orders || nbuf < 1 || nbuf > 16 || norders < 1 || norders > 64 || (bx != 64 && bx != 128 && bx != 256)) {
    snprintf(g_err, sizeof g_err, "bad order-probe arguments");
    return -1;
  }
  Box2 B{};
  B.nx = nx; B.ny = ny; B.nz = nz; B.bx = bx; B.by = 256 / bx;
  const int per_line = 128 / elem_bytes;
  B.R = ((long long)nx + 2 + per_line - 1) / per_line * per_line;
  B.gx = (nx + bx - 1) / bx;
  B.gy = (ny + B.by - 1) / B.by;
  B.nblocks = B.gx * B.gy * nz;
  const size_t bytes = (size_t)B.R * ny * nz * 27 * elem_bytes;
  void *held[16] = {};
  int got = 0;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipStreamCreate(&st);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  for (int k = 0; k < nbuf && e == hipSuccess; ++k) {
    if (hipMalloc(&held[k], bytes) != hipSuccess) {
      (void)hipGetLastError();
      held[k] = nullptr;
      break;
    }
    e = hipMemset(held[k], 0, bytes);
    for (int j = 0; j < norders && e == hipSuccess; ++j) {
      Box2 Bj = B;
      Bj.chunk = orders[j];
      for (int odd = 0; odd < 2 && e == hipSuccess; ++odd) {
        auto launch = [&]() -> hipError_t {
          switch (elem_bytes) {
            case 2: return launch_rows_tile<uint16_t>(held[k], Bj, odd != 0, 0, st);
            case 4: return launch_rows_tile<uint32_t>(held[k], Bj, odd != 0, 0, st);
            default: return launch_rows_tile<uint64_t>(held[k], Bj, odd != 0, 0, st);
          }
        };
        for (int i = 0; i < 2 && e == hipSuccess; ++i) e = launch();
        if (e == hipSuccess) e = hipEventRecord(e0, st);
        for (int i = 0; i < reps && e == hipSuccess; ++i) e = launch();
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        out[((size_t)k * norders + j) * 2 + odd] = ms / reps;
      }
    }
    if (e == hipSuccess) got = k + 1;
  }
  for (int k = 0; k < 16; ++k)
    if (held[k]) (void)hipFree(held[k]);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (e != hipSuccess) {
    snprintf(g_err, sizeof g_err, "%s", hipGetErrorString(e));
    return -1;
  }
  return got;
}


// Read bandwidth of each XCD alone from each `region_bytes` region of one buffer of
// `nregions` regions: out[r * 8 + k] = GB/s of XCD k reading region r (mean of `reps` launches
// after one untimed).  Returns the number of regions, or -1.
int akp_affinity(long long region_bytes, int nregions, int blocks_per_xcd, int reps, double *out) {
  if (region_bytes < (1 << 20) || region_bytes % 16 || nregions < 1 || nregions > 512 || blocks_per_xcd < 1 ||
      blocks_per_xcd > 65536 || reps < 1 || !out) {
    snprintf(g_err, sizeof g_err, "bad affinity-probe arguments");
    return -1;
  }
  const size_t bytes = (size_t)region_bytes * nregions;
  void *buf = nullptr;
  unsigned *sink = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&buf, bytes);
  if (e == hipSuccess) e = hipMalloc((void **)&sink, 256);
  if (e == hipSuccess) e = hipMemset(buf, 0, bytes);
  if (e == hipSuccess) e = hipStreamCreate(&st);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  const size_t n16 = (size_t)region_bytes / 16;
  for (int r = 0; r < nregions && e == hipSuccess; ++r)
    for (int k = 0; k < 8 && e == hipSuccess; ++k) {
      const uint4 *p = (const uint4 *)((char *)buf + (size_t)r * region_bytes);
      hipLaunchKernelGGL(kp_xcd_read, dim3(8 * blocks_per_xcd), dim3(256), 0, st, p, n16, k, sink);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipEventRecord(e0, st);
      for (int i = 0; i < reps && e == hipSuccess; ++i) {
        hipLaunchKernelGGL(kp_xcd_read, dim3(8 * blocks_per_xcd), dim3(256), 0, st, p, n16, k, sink);
        e = hipGetLastError();
      }
      if (e == hipSuccess) e = hipEventRecord(e1, st);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      out[r * 8 + k] = (double)region_bytes / (ms / reps * 1e-3) / 1e9;
    }
  if (buf) (void)hipFree(buf);
  if (sink) (void)hipFree(sink);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (e != hipSuccess) {
    snprintf(g_err, sizeof g_err, "%s", hipGetErrorString(e));
    return -1;
  }
  return nregions;
}

}  // extern "C"
