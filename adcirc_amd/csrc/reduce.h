// Fixed-order reductions shared by the device kernels.  Every inner product
// of the hot path is formed with one blocked tree: 2048-row chunks, 8 rows per
// thread summed in row order, an xor butterfly over the 64 lanes of a wave,
// and (w0+w1)+(w2+w3) over the four waves; chunk values are then summed by
// thread t over chunks t, t+256, ... and reduced with the same tree.  The
// tree is a property of the data layout, not of the launch shape.
#pragma once
#include <hip/hip_runtime.h>

namespace adcgpu {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, 64);
  return v;
}

// Reduces one value per thread of the first 256 threads of the block with the
// fixed tree; result valid in thread 0.  `lds` holds 4 doubles.
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

}  // namespace adcgpu
