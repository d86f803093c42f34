// Fixed-order reductions shared by the device kernels.  Every inner product
// of the hot path is formed with one blocked tree: 2048-row chunks of 512
// threads, 4 rows per thread summed in row order (rows t + 512 r), an xor
// butterfly over the 64 lanes of a wave, and ((w0+w1)+(w2+w3)) +
// ((w4+w5)+(w6+w7)) over the eight waves; the chunk values are then summed by
// thread t of 256 over chunks t, t+256, ..., the butterfly, and (w0+w1)+(w2+w3).
// The tree is a property of the data layout, not of the launch shape.
#pragma once
#include <hip/hip_runtime.h>

namespace adcgpu {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, 64);
  return v;
}

// Cross-chunk stage: reduces one value per thread of the first 256 threads
// of the block; result valid in thread 0.  `lds` holds 4 doubles.
__device__ __forceinline__ double block_tree(double v, double* lds) {
  const int t = threadIdx.x;
  double w = 0.0;
  if (t < 256) w = wave_sum(v);
  if (t < 256 && (t & 63) == 0) lds[t >> 6] = w;
  __syncthreads();
  double r = 0.0;
  if (t == 0) r = (lds[0] + lds[1]) + (lds[2] + lds[3]);
  __syncthreads();
  return r;
}

// Cross-chunk stage for three values at once (the same tree as block_tree
// applied to each); results valid in thread 0.  `lds` holds 12 doubles.
__device__ __forceinline__ void block_tree3(double& a, double& b, double& c, double* lds) {
  const int t = threadIdx.x;
  if (t < 256) {
    const double wa = wave_sum(a), wb = wave_sum(b), wc = wave_sum(c);
    if ((t & 63) == 0) {
      lds[t >> 6] = wa;
      lds[4 + (t >> 6)] = wb;
      lds[8 + (t >> 6)] = wc;
    }
  }
  __syncthreads();
  if (t == 0) {
    a = (lds[0] + lds[1]) + (lds[2] + lds[3]);
    b = (lds[4] + lds[5]) + (lds[6] + lds[7]);
    c = (lds[8] + lds[9]) + (lds[10] + lds[11]);
  }
}

// Chunk stage: reduces one value per thread of a 512-thread group (threads
// g*512 .. g*512+511 of the block, g = group); result valid in the group's
// first thread.  `lds` holds 8 doubles per group; every thread of the block
// must call it.
__device__ __forceinline__ double chunk_tree(double v, double* lds) {
  const int t = threadIdx.x;
  const double w = wave_sum(v);
  if ((t & 63) == 0) lds[t >> 6] = w;
  __syncthreads();
  double r = 0.0;
  if ((t & 511) == 0) {
    const double* x = lds + 8 * (t >> 9);
    r = ((x[0] + x[1]) + (x[2] + x[3])) + ((x[4] + x[5]) + (x[6] + x[7]));
  }
  __syncthreads();
  return r;
}

// XCD-aware tile order.  The dispatcher hands workgroup b to XCD b mod 8,
// so consecutive blocks land in different L2 caches; renumbering blocks so
// that each XCD works through one contiguous range of tiles keeps the halo of
// neighbouring tiles (shared nodes, gathered pseudo-residuals) in one L2.
// A permutation of 0..nb-1; only locality depends on it.
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int x = b % NUM_XCD, k = b / NUM_XCD;
  const int q = nb / NUM_XCD, r = nb % NUM_XCD;
  return x * q + (x < r ? x : r) + k;
}

}  // namespace adcgpu
