// ak_noalloc.hpp -- diagnostics without any allocation (included by ak_capi.hip).
//
// ak_diagnostics_noalloc computes what ak_diagnostics computes, with the same kernels and the
// same reduction order, but its block partials and result live in a file-scope device array
// instead of a buffer allocated per call.  For callers that cannot allocate (code under HIP
// graph capture, tight serving loops).  The array's address on each device is looked up once,
// on the first call with that device current, and kept; host code never passes the array's
// name to a kernel.  Calls that overlap on one GPU share the array and race.
#include <mutex>

// partials of DIAG_MAX_BLOCKS blocks x 5 doubles, then the 5 results (+ padding): 160 KiB
constexpr size_t NOALLOC_SCRATCH_DOUBLES = (size_t)ak::DIAG_MAX_BLOCKS * 5 + 8;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;

std::mutex g_scratch_mutex;
double *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on the current device `dev` (looked up once per device).
int scratch_on(Lattice *s, double **out) {
  if (s->dev < 0 || s->dev >= NOALLOC_MAX_DEVICES) return fail(s, "device %d beyond %d", s->dev, NOALLOC_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[s->dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(s, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[s->dev] = (double *)p;
  }
  *out = g_scratch_ptr[s->dev];
  return 0;
}

}  // namespace

extern "C" {

int ak_diagnostics_noalloc(void *handle, double *out) {
  Lattice *s = H(handle);
  if (check_ready(s)) return -1;  // makes the lattice's device current
  if (!out) return fail(s, "null output");
  const unsigned nb = diag_blocks(s);
  if ((size_t)nb * 5 + 5 > NOALLOC_SCRATCH_DOUBLES)
    return fail(s, "diagnostics of %u blocks exceed the %zu-double scratch array", nb, NOALLOC_SCRATCH_DOUBLES);
  double *d = nullptr;
  if (scratch_on(s, &d)) return -1;
  // every partial and the result are written in full by this call: nothing to clear
  AK_DISPATCH(s, diag_any, s, d, nb);
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, s->stream, (const double *)d, (int)nb, d + (size_t)nb * 5);
  hipError_t e = hipMemcpyAsync(out, d + (size_t)nb * 5, 5 * sizeof(double), hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return fail(s, "diagnostics: %s", hipGetErrorString(e));
  return 0;
}

}  // extern "C"
