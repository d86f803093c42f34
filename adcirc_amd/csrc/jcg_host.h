// Host-side set-up of the persistent JCG launch (k_jcg), shared by the
// context API and the allocation-free drop-in.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "kernels.h"

namespace adcgpu {

// Records the message adcgpu_last_error returns (a fixed per-thread buffer).
void set_last_error(const char* msg);

// Rows each side of a block's rows whose pseudo-residual is staged in LDS for
// a matrix of bandwidth bw (the widest |column - row|), or -1 when they do
// not fit; and the dynamic LDS bytes of a launch.
inline int jcg_stage_rows(int bw) {
  const size_t bytes = 8 * ((size_t)JCG_CPB * RED_CHUNK + 2 * (size_t)bw);
  return bytes <= (size_t)JCG_STAGE_MAX_BYTES ? bw : -1;
}
inline size_t jcg_stage_bytes(int hrows) {
  return hrows < 0 ? 0 : 8 * ((size_t)JCG_CPB * RED_CHUNK + 2 * (size_t)hrows);
}

// Barrier words of a grid: the grid barrier's counters, then one halo flag
// per block.
inline int jcg_barrier_words_for(int grid) {
  return ((grid + JCG_BAR_GROUP - 1) / JCG_BAR_GROUP + 2) * JCG_BAR_STRIDE +
         grid * JCG_HALO_STRIDE;
}

// Blocks of k_jcg that are resident on one CU at once.  The occupancy query
// counts wave slots per CU, but the dispatcher places the waves of a
// workgroup on the four SIMDs in turn, so a block of W waves takes ceil(W/4)
// slots of one SIMD (measured on gfx950: 5-wave blocks with 164 VGPRs run one
// per CU although the query reports two; 4- and 8-wave blocks run as many as
// it reports).  A cooperative grid larger than what is really resident never
// passes its first grid barrier, so the smaller of the two counts is used.
// The device's capacity is that per-CU count times the CUs, and at most what
// the barrier words hold.
inline hipError_t jcg_resident_cap_of(const void* fn, int* per_cu_out) {
  int per_cu = 0;
  // the largest staged pseudo-residual a launch may ask for
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     JCG_STAGE_MAX_BYTES);
  if (e != hipSuccess) return e;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, JCG_BLOCK, JCG_STAGE_MAX_BYTES);
  if (e != hipSuccess) return e;
  hipFuncAttributes fa;
  if ((e = hipFuncGetAttributes(&fa, fn)) != hipSuccess) return e;
  const int vgpr = std::max((fa.numRegs + 7) & ~7, 8);  // allocation granule 8
  const int per_simd = std::min(8, 512 / vgpr);
  const int waves = (JCG_BLOCK + 63) / 64;
  *per_cu_out = std::min(per_cu, per_simd / ((waves + 3) / 4));
  return hipSuccess;
}

// Every variant of the kernel that is launched must fit: one start or two,
// and one start with the staged pseudo-residual; single and batched.
inline hipError_t jcg_resident_cap(int device, int* cap) {
  const void* fns[] = {(const void*)k_jcg<false, false>,      (const void*)k_jcg<true, false>,
                       (const void*)k_jcg<false, true>,       (const void*)k_jcg<true, true>,
                       (const void*)k_jcg_batch<false, false>, (const void*)k_jcg_batch<true, false>,
                       (const void*)k_jcg_batch<false, true>, (const void*)k_jcg_batch<true, true>,
                       (const void*)k_jcg_mem};
  int per_cu = 1 << 30;
  hipError_t e = hipSuccess;
  for (const void* fn : fns) {
    int p = 0;
    if ((e = jcg_resident_cap_of(fn, &p)) != hipSuccess) return e;
    per_cu = std::min(per_cu, p);
  }
  int cus = 0;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device)) !=
      hipSuccess)
    return e;
  int bar_cap = 0;
  while (jcg_barrier_words_for(bar_cap + 1) <= JCG_BAR_WORDS) ++bar_cap;
  *cap = std::min(per_cu * cus, bar_cap);
  return hipSuccess;
}

// Blocks of one launch: JCG_CPB chunks of 2048 rows per block plus the
// eigenvalue block.
inline int jcg_grid_for(int n) {
  const int nch = (n + RED_CHUNK - 1) / RED_CHUNK;
  return (nch + JCG_CPB - 1) / JCG_CPB + 1;
}

// ITPACK's tolerances of one solve: zeta raised to 512*srelpr when smaller
// (src/itpackv.F:223-232); eigvss's nsig = max(int(-log10|zeta|), 4) and
// zbrent's t = 10**(-nsig); maxfn = max(itmax, 50).
inline void jcg_tolerances(JcgArgs& a, int itmax, double zeta) {
  a.itmax = itmax;
  a.maxfn = std::max(itmax, 50);
  const double srelpr = std::ldexp(1.0, -52);
  a.zeta = zeta >= 512.0 * srelpr ? zeta : 512.0 * srelpr;
  const int itmp = (int)(-std::log10(std::fabs(a.zeta)));
  const int nsig = std::max(itmp, 4);
  double p10 = 1.0;
  for (int i = 0; i < nsig; ++i) p10 *= 10.0;
  a.tz = 1.0 / p10;
}

}  // namespace adcgpu
