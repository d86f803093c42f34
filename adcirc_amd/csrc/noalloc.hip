// adcgpu_jcg_noalloc: the drop-in JCG for callers that cannot allocate
// memory (code under graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  Every device buffer of the solve lives in one static
// device array, `scratch`; the host side allocates nothing either: the
// ELLPACK compaction that adcgpu_jcg does on the host runs here in a kernel.
//
// Each GPU holds its own copy of `scratch` at its own address, obtained once
// per device with hipGetSymbolAddress; host code uses that address, never
// the array's name (a host-side stand-in).  The array keeps what the last
// call left, including ITPACK's saved state, which carries from one call to
// the next on a device like the reference's SAVE/COMMON data.  Calls that
// overlap on one GPU race on the array.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../include/adcgpu.h"
#include "jcg_host.h"
#include "kernels.h"

namespace adcgpu {

constexpr int NA_MAX_N = ADCGPU_JCG_NOALLOC_MAX_N;
constexpr int NA_MAX_NZ = ADCGPU_JCG_NOALLOC_MAX_NZ;
constexpr int NA_MAX_NCH = (NA_MAX_N + RED_CHUNK - 1) / RED_CHUNK;

// Byte offsets of the sections of `scratch` (each 256-byte aligned).
constexpr size_t na_align(size_t v) { return (v + 255) & ~size_t(255); }
constexpr size_t NA_IN_COEF = 0;                                        // [maxnz][n] f64
constexpr size_t NA_IN_JCOEF = na_align(NA_IN_COEF + 8ull * NA_MAX_N * NA_MAX_NZ);  // i32
constexpr size_t NA_COEF = na_align(NA_IN_JCOEF + 4ull * NA_MAX_N * NA_MAX_NZ);     // f64
constexpr size_t NA_CS = na_align(NA_COEF + 8ull * NA_MAX_N * NA_MAX_NZ);           // f64
constexpr size_t NA_COL = na_align(NA_CS + 8ull * NA_MAX_N * NA_MAX_NZ);            // i32
constexpr size_t NA_NNEIGH = na_align(NA_COL + 4ull * NA_MAX_N * NA_MAX_NZ);        // i32
constexpr size_t NA_RHS = na_align(NA_NNEIGH + 4ull * NA_MAX_N);
constexpr size_t NA_U = na_align(NA_RHS + 8ull * NA_MAX_N);
constexpr size_t NA_D = na_align(NA_U + 8ull * NA_MAX_N);
constexpr size_t NA_W = na_align(NA_D + 8ull * NA_MAX_N);
constexpr size_t NA_PART = na_align(NA_W + 8ull * NA_MAX_N);                // [2][nch][4] f64
constexpr size_t NA_SAVED = na_align(NA_PART + 8ull * 2 * NA_MAX_NCH * 4);  // [12] f64
constexpr size_t NA_BAR = na_align(NA_SAVED + 8ull * 12);                   // barrier words
constexpr size_t NA_FLAGS = na_align(NA_BAR + 4ull * JCG_BAR_WORDS);        // [F_COUNT] i32
constexpr size_t NA_CHECK = na_align(NA_FLAGS + 4ull * F_COUNT);            // [1] i32
constexpr size_t NA_BYTES = na_align(NA_CHECK + 4);
static_assert(NA_BYTES == ADCGPU_JCG_NOALLOC_SCRATCH_BYTES,
              "the scratch size declared in adcgpu.h must match its layout");

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

// First failing row and check of the system, as the host drop-in finds them
// row by row: per row, no diagonal (402) before a column index above n
// (argument error) before a zero diagonal (401) before a negative one
// (unsupported).  `check` holds the smallest row*4 + kind seen.
enum { NA_NODIAG = 0, NA_BADCOL = 1, NA_ZERODIAG = 2, NA_NEGDIAG = 3 };

// One row per thread: the ELLPACK row of the caller's (n, maxnz) arrays,
// copied with leading dimension n, compacted to the solver's layout with the
// diagonal first and the padding (jcoef <= 0) dropped, columns 0-based.
__global__ __launch_bounds__(256) void k_na_compact(int n, int maxnz, const double* in_coef,
                                                    const int* in_jcoef, double* coef, int* col,
                                                    int* nneigh, int* check) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int dj = -1;
  for (int j = 0; j < maxnz; ++j)
    if (in_jcoef[(size_t)j * n + i] == i + 1) dj = j;
  if (dj < 0) {
    atomicMin(check, i * 4 + NA_NODIAG);
    return;
  }
  int k = 0;
  coef[i] = in_coef[(size_t)dj * n + i];
  col[i] = i;
  ++k;
  for (int j = 0; j < maxnz; ++j) {
    const int v = in_jcoef[(size_t)j * n + i];
    if (v > n) {
      atomicMin(check, i * 4 + NA_BADCOL);
      return;
    }
    if (j != dj && v > 0) {
      coef[(size_t)k * n + i] = in_coef[(size_t)j * n + i];
      col[(size_t)k * n + i] = v - 1;
      ++k;
    }
  }
  nneigh[i] = k;
  const double d = coef[i];
  if (d == 0.0) atomicMin(check, i * 4 + NA_ZERODIAG);
  else if (d < 0.0) atomicMin(check, i * 4 + NA_NEGDIAG);
}

// The solved (unscaled) rows back to the caller's positions.
__global__ __launch_bounds__(256) void k_na_scatter(int n, int maxnz, double* in_coef,
                                                    const int* in_jcoef, const double* coef) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int dj = -1;
  for (int j = 0; j < maxnz; ++j)
    if (in_jcoef[(size_t)j * n + i] == i + 1) dj = j;
  in_coef[(size_t)dj * n + i] = coef[i];
  int k = 1;
  for (int j = 0; j < maxnz; ++j) {
    const int v = in_jcoef[(size_t)j * n + i];
    if (j != dj && v > 0) in_coef[(size_t)j * n + i] = coef[(size_t)(k++) * n + i];
  }
}

__global__ void k_na_clear(int* flags, int* check, unsigned* bar, int bar_words) {
  for (int k = threadIdx.x; k < F_COUNT; k += blockDim.x) flags[k] = 0;
  if (threadIdx.x == 0) *check = 0x7fffffff;
  for (int k = threadIdx.x; k < bar_words; k += blockDim.x) bar[k] = 0u;
}

}  // namespace adcgpu

using namespace adcgpu;

namespace {

constexpr int NA_DEVICES = 64;
// per device: the address of its `scratch`, the resident JCG block limit and
// the wall-clock rate, set once with that device current
unsigned char* g_na_ptr[NA_DEVICES] = {nullptr};
int g_na_cap[NA_DEVICES] = {0};
int g_na_khz[NA_DEVICES] = {0};
int na_fail(int code, const char* what, hipError_t e = hipSuccess) {
  char msg[256];
  std::snprintf(msg, sizeof(msg), "adcgpu_jcg_noalloc: %s%s%s", what,
                e != hipSuccess ? ": " : "", e != hipSuccess ? hipGetErrorString(e) : "");
  set_last_error(msg);
  return code;
}

}  // namespace

#define NACHK(expr)                                              \
  do {                                                           \
    hipError_t e_ = (expr);                                      \
    if (e_ != hipSuccess) return na_fail(ADCGPU_E_HIP, #expr, e_); \
  } while (0)

extern "C" {

int adcgpu_jcg_noalloc(int n, int ndim, int maxnz, int32_t* jcoef, double* coef, double* rhs,
                       double* u, int32_t* iparm, double* rparm) {
  if (n <= 0) return 11;
  if (!jcoef || !coef || !rhs || !u || !iparm || !rparm || ndim < n || maxnz < 1)
    return na_fail(ADCGPU_E_ARG, "invalid argument");
  if (n > NA_MAX_N || maxnz > NA_MAX_NZ)
    return na_fail(ADCGPU_E_ARG, "system larger than the static scratch holds "
                                 "(n <= ADCGPU_JCG_NOALLOC_MAX_N, maxnz <= "
                                 "ADCGPU_JCG_NOALLOC_MAX_NZ)");
  if (iparm[0] <= 0) return 0;
  if (iparm[5] != 1 || iparm[8] != -2 || iparm[9] != 0)
    return na_fail(ADCGPU_E_UNSUPPORTED, "only adaptive JCG without red-black or row removal");
  const int itmax = iparm[0];
  if (itmax > JCG_MAX_ITMAX) return na_fail(ADCGPU_E_ARG, "ITMAX out of range");
  int dev = 0;
  NACHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NA_DEVICES) return na_fail(ADCGPU_E_ARG, "device ordinal");
  if (!g_na_ptr[dev]) {
    // once per device, with the device current
    void* p = nullptr;
    NACHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    int cap = 0, khz = 0;
    NACHK(jcg_resident_cap(dev, &cap));
    NACHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    g_na_cap[dev] = cap;
    g_na_khz[dev] = khz > 1000 ? khz : 1000;
    g_na_ptr[dev] = static_cast<unsigned char*>(p);
    // ITPACK's saved state starts cleared on this device (a fill kernel, so
    // that the entry point records no memset)
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, 0,
                       reinterpret_cast<int*>(g_na_ptr[dev] + NA_SAVED), 0, 24);
    NACHK(hipGetLastError());
  }
  const int grid = jcg_grid_for(n);
  if (grid > g_na_cap[dev]) return na_fail(ADCGPU_E_CAPACITY, "grid larger than the device holds");
  unsigned char* base = g_na_ptr[dev];
  double* in_coef = reinterpret_cast<double*>(base + NA_IN_COEF);
  int* in_jcoef = reinterpret_cast<int*>(base + NA_IN_JCOEF);
  double* dcoef = reinterpret_cast<double*>(base + NA_COEF);
  double* dcs = reinterpret_cast<double*>(base + NA_CS);
  int* dcol = reinterpret_cast<int*>(base + NA_COL);
  int* dnn = reinterpret_cast<int*>(base + NA_NNEIGH);
  double* drhs = reinterpret_cast<double*>(base + NA_RHS);
  double* du = reinterpret_cast<double*>(base + NA_U);
  double* dd = reinterpret_cast<double*>(base + NA_D);
  double* dw = reinterpret_cast<double*>(base + NA_W);
  double* dpart = reinterpret_cast<double*>(base + NA_PART);
  double* dsaved = reinterpret_cast<double*>(base + NA_SAVED);
  unsigned* dbar = reinterpret_cast<unsigned*>(base + NA_BAR);
  int* dflags = reinterpret_cast<int*>(base + NA_FLAGS);
  int* dcheck = reinterpret_cast<int*>(base + NA_CHECK);
  const hipStream_t st = 0;
  const int bar_words = jcg_barrier_words_for(grid);
  hipLaunchKernelGGL(k_na_clear, dim3(1), dim3(256), 0, st, dflags, dcheck, dbar, bar_words);
  // the caller's n rows of each column (leading dimension ndim)
  NACHK(hipMemcpy2DAsync(in_coef, (size_t)n * 8, coef, (size_t)ndim * 8, (size_t)n * 8, maxnz,
                         hipMemcpyHostToDevice, st));
  NACHK(hipMemcpy2DAsync(in_jcoef, (size_t)n * 4, jcoef, (size_t)ndim * 4, (size_t)n * 4, maxnz,
                         hipMemcpyHostToDevice, st));
  NACHK(hipMemcpyAsync(drhs, rhs, (size_t)n * 8, hipMemcpyHostToDevice, st));
  NACHK(hipMemcpyAsync(du, u, (size_t)n * 8, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_na_compact, dim3((n + 255) / 256), dim3(256), 0, st, n, maxnz,
                     (const double*)in_coef, (const int*)in_jcoef, dcoef, dcol, dnn, dcheck);
  NACHK(hipGetLastError());
  int chk = 0;
  NACHK(hipMemcpyAsync(&chk, dcheck, sizeof(int), hipMemcpyDeviceToHost, st));
  NACHK(hipStreamSynchronize(st));
  if (chk != 0x7fffffff) {
    switch (chk & 3) {
      case NA_NODIAG: return 402;
      case NA_BADCOL: return na_fail(ADCGPU_E_ARG, "jcoef out of range");
      case NA_ZERODIAG: return 401;
      default: return na_fail(ADCGPU_E_UNSUPPORTED, "negative diagonal (nonsymmetric scaling)");
    }
  }
  JcgArgs a;
  a.n = n;
  jcg_tolerances(a, itmax, rparm[0]);
  a.coef = dcoef;
  a.cs = dcs;
  a.col = dcol;
  a.nneigh = dnn;
  a.rhs = drhs;
  a.u = du;
  a.d = dd;
  a.w = dw;
  a.part = dpart;
  a.saved = dsaved;
  a.prescaled = 0;
  a.defer_unscal = 0;
  a.barrier = dbar;
  a.halo = -1;  // the caller's matrix: a grid barrier after each update
  a.hrows = -1;
  a.barrier_ticks = (long long)g_na_khz[dev] * 500;  // 0.5 s
  a.flags = dflags;
  a.eta2 = nullptr;
  a.eta1 = nullptr;
  a.nodecode = nullptr;
  a.u2 = nullptr;
  a.d2 = nullptr;
  void* args[] = {&a};
  NACHK(hipLaunchCooperativeKernel((const void*)k_jcg<false, false>, dim3(grid), dim3(JCG_BLOCK), args,
                                   0, st));
  int f[F_COUNT];
  double saved[8];
  NACHK(hipMemcpyAsync(f, dflags, sizeof(f), hipMemcpyDeviceToHost, st));
  NACHK(hipMemcpyAsync(saved, dsaved, sizeof(saved), hipMemcpyDeviceToHost, st));
  NACHK(hipStreamSynchronize(st));
  if (f[F_ERROR] & ERR_BARRIER) return na_fail(ADCGPU_E_HIP, "JCG wait timed out");
  if (f[F_ERROR] & ERR_DIAG) return na_fail(ADCGPU_E_DIAG, "diagonal <= 0");
  hipLaunchKernelGGL(k_na_scatter, dim3((n + 255) / 256), dim3(256), 0, st, n, maxnz, in_coef,
                     (const int*)in_jcoef, (const double*)dcoef);
  NACHK(hipGetLastError());
  NACHK(hipMemcpy2DAsync(coef, (size_t)ndim * 8, in_coef, (size_t)n * 8, (size_t)n * 8, maxnz,
                         hipMemcpyDeviceToHost, st));
  NACHK(hipMemcpyAsync(rhs, drhs, (size_t)n * 8, hipMemcpyDeviceToHost, st));
  NACHK(hipMemcpyAsync(u, du, (size_t)n * 8, hipMemcpyDeviceToHost, st));
  NACHK(hipStreamSynchronize(st));
  const int in = f[F_NUMITR];
  const int ier = f[F_IER];
  iparm[7] = 4 * n + 4 * in;
  if (ier == 13 && iparm[2] == 0) rparm[0] = saved[5];
  if (iparm[2] == 0) {
    iparm[0] = in;
    iparm[8] = -2;
    rparm[1] = saved[4];
    rparm[2] = rparm[2] == 0.0 ? -1.0 : rparm[2];
    rparm[10] = saved[5] > 0.0 ? -std::log10(std::fabs(saved[5])) : -std::log10(std::ldexp(1.0, -52));
  }
  return ier;
}

}  // extern "C"
