// The learned JCG warm start on the gfx950 matrix cores.
//
// A per-node MLP, 48 inputs -> 128 -> 128 -> 1 with GELU, BF16 weights and
// FP32 accumulation, predicts the correction to the reference's JCG initial
// guess (the last increment ETAS, i.e. the second-order extrapolation
// 2*eta^k - eta^(k-1) of the new elevation; src/gwce.F:2002-2004) from the
// node's stencil (itself and up to 12 neighbours in NEITAB order): the last
// two increments and the elevation differences, plus depth, wind stress,
// pressure change, friction and the wet/dry flags.  JCG then runs to the
// unchanged stopping test (src/itpackv.F:809-890) from whichever of the two
// starts has the smaller initial pseudo-residual (k_jcg).
//
// The inference kernel chains the layers in registers.  With the nodes as
// the N dimension of v_mfma_f32_32x32x16_bf16 (lane l holds node l % 32 and
// half h = l / 32 of the K slice), layer 1's accumulator registers hold, for
// lane l and register r of hidden tile t, hidden unit 32t + 8(r/4) + 4h + r%4
// of node l % 32; after bias and GELU they are exactly the B operand of layer
// 2 if layer 2's K order is that permutation, which the packed W2 follows.
// The output layer is a dot product over each lane's 64 hidden values and
// one exchange between the two halves.
#include <hip/hip_runtime.h>

#include "mlp.h"
#include "reduce.h"

namespace adcgpu {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr float GELU_C0 = 0.7978845608028654f;  // sqrt(2/pi)
constexpr float GELU_C1 = 0.044715f;
constexpr float LOG2E = 1.4426950408889634f;

// GELU, tanh form (Hendrycks & Gimpel 2016): 0.5 x (1 + tanh(c0 (x + c1 x^3))),
// tanh(v) = 1 - 2 / (1 + exp(2v)) with the hardware exp2 and reciprocal.
__device__ __forceinline__ float ml_tanh(float v) {
  const float e = __builtin_amdgcn_exp2f(2.0f * LOG2E * v);
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + e);
}
// The same function as x / (1 + exp(-2 c0 (x + c1 x^3))), since
// 0.5 (1 + tanh(z)) = 1 / (1 + exp(-2z)): six VALU operations and two
// transcendentals per value, the inference kernel's VALU floor.
__device__ __forceinline__ float ml_gelu(float x) {
  const float x2 = x * x;
  const float p = __builtin_fmaf(GELU_C1 * x2, x, x);
  const float e = __builtin_amdgcn_exp2f((-2.0f * GELU_C0 * LOG2E) * p);
  return x * __builtin_amdgcn_rcpf(1.0f + e);
}
__device__ __forceinline__ float ml_gelu_grad(float x) {
  const float t = ml_tanh(GELU_C0 * (x + GELU_C1 * x * x * x));
  return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * GELU_C0 * (1.0f + 3.0f * GELU_C1 * x * x);
}

__device__ __forceinline__ double ml_clamp(double v, double lim) {
  return v < -lim ? -lim : (v > lim ? lim : v);
}

// Local feature q (0..23) of a lane of half h is input k = 16 (q/8) + 8h + q%8.
__device__ __forceinline__ int ml_k(int h, int q) { return 16 * (q >> 3) + 8 * h + (q & 7); }

// The 24 features of node i that half h of its lane pair holds, and the
// node's scale s (the stencil RMS of the last increment and of its change).
// Half 0: stencil nodes 0..6 (the node and its first six neighbours), three
// features each (ETAS/s, (ETAS - ETAS_prev)/s, (eta_j - eta_i)/(64 s)), then
// depth, wind stress x and y.  Half 1: stencil nodes 7..12, then the pressure
// change, the node's wet flag, its neighbour count, log10 of s, the wet
// fraction of its neighbours and TK*DT.  Padding slots are zeros.  Both
// halves of every lane pair must call it together (one exchange across the
// halves); lanes past NP compute zeros.
__device__ void ml_features(const MlIn& in, int i, int h, bool valid, float f[24], double& s_out) {
  const int lo = h == 0 ? 0 : 7, hi = h == 0 ? 7 : ML_S;
  int node[7];
  double es[7], esp[7];
  double part = 0.0;
  int nn = 1;
  if (valid) nn = in.nneigh[i];
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    const int j = lo + q;
    int v = -1;
    if (valid && j < hi) v = (j == 0) ? i : (j < nn ? in.col[(size_t)j * in.np + i] : -1);
    node[q] = v;
    es[q] = 0.0;
    esp[q] = 0.0;
    if (v >= 0) {
      es[q] = in.etas[v];
      esp[q] = in.etasp[v];
    }
    const double d2 = es[q] - esp[q];
    part = part + (es[q] * es[q] + d2 * d2);
  }
  const double tot = part + __shfl_xor(part, 32, 64);
  const double s = __builtin_sqrt(tot / (double)ML_S) + 1.0e-12;
  s_out = s;
  const double rs = 1.0 / s;
  const double ei = valid ? in.eta[i] : 0.0;
  for (int q = 0; q < 24; ++q) f[q] = 0.0f;
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    if (h == 1 && q == 6) break;
    const int v = node[q];
    if (v < 0) continue;
    f[3 * q] = (float)(es[q] * rs);
    f[3 * q + 1] = (float)((es[q] - esp[q]) * rs);
    f[3 * q + 2] = (float)ml_clamp((in.eta[v] - ei) * rs * (1.0 / 64.0), 8.0);
  }
  if (!valid) return;
  if (h == 0) {
    f[21] = (float)tanh(in.dp[i] / 50.0);
    f[22] = (float)ml_clamp(in.wsx2[i] * 1000.0, 8.0);
    f[23] = (float)ml_clamp(in.wsy2[i] * 1000.0, 8.0);
  } else {
    int wet = 0;
    for (int j = 1; j < nn; ++j) wet += in.nodecode[in.col[(size_t)j * in.np + i]];
    f[18] = (float)ml_clamp((in.pr2[i] - in.pr1[i]) * rs, 8.0);
    f[19] = (float)in.nodecode[i];
    f[20] = (float)(nn - 1) / 12.0f;
    f[21] = (float)(log10(s) / 10.0);
    f[22] = nn > 1 ? (float)wet / (float)(nn - 1) : 0.0f;
    f[23] = (float)(in.tk[i] * in.dt);
  }
}

// ---------------------------------------------------------------- packing
// master: FP32 parameters (ML_OFF_*).  One thread per packed element.
__global__ __launch_bounds__(256) void k_ml_pack(const float* __restrict__ master, MlWeights w) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const float* W1 = master + ML_OFF_W1;  // [48][128]
  const float* W2 = master + ML_OFF_W2;  // [128 in][128 out]
  // w1p[t][s][lane][u] = W1[16s + 8(lane/32) + u][32t + lane%32]
  if (g < 4 * 3 * 64 * 8) {
    const int u = g & 7, lane = (g >> 3) & 63, s = (g >> 9) % 3, t = g / (3 * 512);
    w.w1p[g] = (bf16)W1[(16 * s + 8 * (lane >> 5) + u) * ML_H + 32 * t + (lane & 31)];
  }
  // w2p[T][s2][lane][u] = W2[j][32T + lane%32], j = the layer-1 hidden unit
  // that lane half h holds in B-register u of K step s2 = 2t + jj
  if (g < 4 * 8 * 64 * 8) {
    const int u = g & 7, lane = (g >> 3) & 63, s2 = (g >> 9) & 7, T = g >> 12;
    const int t = s2 >> 1, jj = s2 & 1, h = lane >> 5;
    const int j = 32 * t + 16 * jj + 8 * (u >> 2) + 4 * h + (u & 3);
    w.w2p[g] = (bf16)W2[j * ML_H + 32 * T + (lane & 31)];
  }
  // training copies: w1t [128][48], w2t [out][in], w2n [in][out]
  if (g < ML_W1) {
    const int o = g / ML_F, k = g % ML_F;
    w.w1t[g] = (bf16)W1[k * ML_H + o];
  }
  if (g < ML_W2) {
    const int o = g / ML_H, j = g % ML_H;
    w.w2t[g] = (bf16)W2[j * ML_H + o];
    w.w2n[g] = (bf16)W2[g];
  }
  // bias / output weight of half h, tile t, register r: unit 32t + 8(r/4) + 4h + r%4
  if (g < 2 * 4 * 16) {
    const int r = g & 15, t = (g >> 4) & 3, h = g >> 6;
    const int o = 32 * t + 8 * (r >> 2) + 4 * h + (r & 3);
    w.b1p[g] = master[ML_OFF_B1 + o];
    w.b2p[g] = master[ML_OFF_B2 + o];
    w.w3p[g] = (float)(bf16)master[ML_OFF_W3 + o];
  }
  if (g == 0) *w.b3 = master[ML_OFF_B3];
}

// ---------------------------------------------------------------- inference
// Two passes.  The stencil gathers of the inputs are latency-bound and the
// network is matrix-core and VALU work with a large register footprint, so
// they run as separate kernels: k_ml_feat at full occupancy, one wave per
// tile of 32 nodes, stores every node's inputs in the operand layout of the
// first layer (fx[(tile * 3 + s3) * 64 + lane] = the 8 inputs K = 16 s3 +
// 8 (lane / 32) .. +7 of node 32 tile + lane % 32) and its scale s; nodes that
// take no correction (dry, or on the elevation boundary) get zero inputs
// without gathers.  k_ml_mlp reads them back coalesced (96 bytes per node).
__device__ __forceinline__ bool ml_use(const MlIn& in, int i) {
  return in.nodecode[i] == 1 && in.bnd_index[i] < 0;
}

__global__ __launch_bounds__(256) void k_ml_feat(MlIn in, bf16x8* __restrict__ fx,
                                                 double* __restrict__ sv) {
  const int lane = threadIdx.x & 63, h = lane >> 5;
  const int tile = xcd_tile(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  if (tile >= (in.np + 31) >> 5) return;  // the whole wave
  const int i = tile * 32 + (lane & 31);
  const bool valid = i < in.np;
  float f[24];
  double s;
  ml_features(in, i, h, valid && ml_use(in, i), f, s);
  bf16x8 x[3];
#pragma unroll
  for (int q = 0; q < 24; ++q) x[q >> 3][q & 7] = (bf16)f[q];
#pragma unroll
  for (int s3 = 0; s3 < 3; ++s3) fx[((size_t)tile * 3 + s3) * 64 + lane] = x[s3];
  if (valid && h == 0) sv[i] = s;
}

// The network over the stored inputs.  Four waves per block, each taking 32
// nodes at a time; a tile with no node that takes a correction is skipped.
// The packed weights sit in LDS.  Output: the correction s * y for wet nodes
// off the elevation boundary, zero elsewhere.
__global__ __launch_bounds__(256) void k_ml_mlp(MlIn in, MlWeights w,
                                                const bf16x8* __restrict__ fx,
                                                const double* __restrict__ sv,
                                                double* __restrict__ corr) {
  __shared__ bf16x8 w2s[4 * 8 * 64];
  __shared__ bf16x8 w1s[4 * 3 * 64];
  __shared__ float bs[3][128];
  for (int k = threadIdx.x; k < 4 * 8 * 64; k += blockDim.x)
    w2s[k] = reinterpret_cast<const bf16x8*>(w.w2p)[k];
  for (int k = threadIdx.x; k < 4 * 3 * 64; k += blockDim.x)
    w1s[k] = reinterpret_cast<const bf16x8*>(w.w1p)[k];
  for (int k = threadIdx.x; k < 128; k += blockDim.x) {
    bs[0][k] = w.b1p[k];
    bs[1][k] = w.b2p[k];
    bs[2][k] = w.w3p[k];
  }
  const int lane = threadIdx.x & 63, h = lane >> 5;
  const float b3 = *w.b3;
  __syncthreads();
  const int waves = gridDim.x * (blockDim.x >> 6);
  const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int ntile = (in.np + 31) >> 5;
  for (int tile = wave; tile < ntile; tile += waves) {
    const int i = tile * 32 + (lane & 31);
    const bool valid = i < in.np;
    const bool use = valid && ml_use(in, i);
    if (__ballot(use) == 0) {
      if (valid && h == 0) corr[i] = 0.0;
      continue;
    }
    bf16x8 x[3];
#pragma unroll
    for (int s3 = 0; s3 < 3; ++s3) x[s3] = fx[((size_t)tile * 3 + s3) * 64 + lane];
    // layer 1: 4 hidden tiles x 3 K steps, each tile activated at once
    bf16x8 a1[8];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f32x16 acc = f32x16{};
#pragma unroll
      for (int s3 = 0; s3 < 3; ++s3)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w1s[(t * 3 + s3) * 64 + lane], x[s3], acc,
                                                      0, 0, 0);
#pragma unroll
      for (int r = 0; r < 16; ++r)
        a1[2 * t + (r >> 3)][r & 7] = (bf16)ml_gelu(acc[r] + bs[0][(h * 4 + t) * 16 + r]);
    }
    // layer 2: 4 output tiles x 8 K steps, K in layer 1's register order
    float part = 0.0f;
#pragma unroll 1
    for (int T = 0; T < 4; ++T) {
      f32x16 c = f32x16{};
#pragma unroll
      for (int s2 = 0; s2 < 8; ++s2)
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w2s[(T * 8 + s2) * 64 + lane], a1[s2], c, 0, 0,
                                                    0);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int pk = (h * 4 + T) * 16 + r;
        part = part + bs[2][pk] * ml_gelu(c[r] + bs[1][pk]);
      }
    }
    const float y = part + __shfl_xor(part, 32, 64) + b3;
    if (valid && h == 0) corr[i] = use ? sv[i] * (double)y : 0.0;
  }
}

// ---------------------------------------------------------------- training data
// Sample q of this step is node (it * 7919 + q * stride) mod NP: a spread,
// deterministic set.  Nodes that are dry or on the elevation boundary are
// kept with node = -1 (no weight).  Same lane layout as k_ml_feat.
__global__ __launch_bounds__(256) void k_ml_collect(MlIn in, MlSamples smp, int slot0, int count,
                                                    int it) {
  const int lane = threadIdx.x & 63, h = lane >> 5;
  const int q = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 32 + (lane & 31);
  const bool active = q < count;
  const long long stride = in.np / (count > 0 ? count : 1) + 1;
  // it < 0: every node in order (slot q = node q), for inspection
  const int i = !active ? 0 : (it < 0 ? q : (int)(((long long)it * 7919 + (long long)q * stride) %
                                                  in.np));
  float f[24];
  double s;
  ml_features(in, i, h, active, f, s);
  if (!active) return;
  const int slot = (slot0 + q) % smp.cap;
#pragma unroll
  for (int k = 0; k < 24; ++k) smp.x[(size_t)slot * ML_F + ml_k(h, k)] = (bf16)f[k];
  if (h == 0) {
    const bool use = in.nodecode[i] == 1 && in.bnd_index[i] < 0;
    smp.s[slot] = (float)s;
    smp.e0[slot] = in.etas[i];
    smp.node[slot] = use ? i : -1;
  }
}

__global__ __launch_bounds__(256) void k_ml_target(MlSamples smp, int slot0, int count,
                                                   const double* __restrict__ etas_new) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= count) return;
  const int slot = (slot0 + q) % smp.cap;
  const int i = smp.node[slot];
  smp.y[slot] = i >= 0 ? (float)((etas_new[i] - smp.e0[slot]) / (double)smp.s[slot]) : 0.0f;
}

// ---------------------------------------------------------------- training GEMM
// C[M][N] = A[M][K] * B[N][K]^T (bf16 in, f32 out), K a multiple of 16.  One
// wave per 32 x 32 tile of C; operand fragments loaded straight from memory
// (lane l: row l%32 of the tile, K elements 8(l/32) .. +7); rows and columns
// past M, N read as zero and are not stored.
__global__ __launch_bounds__(64) void k_ml_gemm_nt(int M, int N, int K, const bf16* __restrict__ A,
                                                   int lda, const bf16* __restrict__ B, int ldb,
                                                   float* __restrict__ C, int ldc) {
  const int lane = threadIdx.x & 63, h = lane >> 5;
  const int m0 = blockIdx.x * 32, n0 = blockIdx.y * 32;
  const int am = m0 + (lane & 31), bn = n0 + (lane & 31);
  f32x16 acc = f32x16{};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x8 a = bf16x8{}, b = bf16x8{};
    if (am < M) a = *reinterpret_cast<const bf16x8*>(A + (size_t)am * lda + k0 + 8 * h);
    if (bn < N) b = *reinterpret_cast<const bf16x8*>(B + (size_t)bn * ldb + k0 + 8 * h);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
  // D register r of lane l: row 8(r/4) + 4h + r%4, column l%32
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = m0 + 8 * (r >> 2) + 4 * h + (r & 3), colx = n0 + (lane & 31);
    if (row < M && colx < N) C[(size_t)row * ldc + colx] = acc[r];
  }
}

// ---------------------------------------------------------------- training passes
// One minibatch of ML_BATCH samples (slots of the sample ring; -1 pads):
//   Z1 = X W1 + b1, A1 = gelu(Z1) (bf16), Z2 = A1 W2 + b2, A2 = gelu(Z2),
//   y = A2 w3 + b3, loss = sum_b w_b (y_b - t_b)^2 / n,
// and the gradients of the loss by back-propagation.  The matrix products
// are k_ml_gemm_nt; everything else is here.  Per-block partial sums are
// reduced in a fixed order by k_ml_adam, so training is deterministic.

__device__ __forceinline__ float wave_sumf(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, 64);
  return v;
}

// Per-sample loss weights: s^2 / mean(s^2) over the samples that count
// (node >= 0), zero for the others.  One block.
__global__ __launch_bounds__(1024) void k_ml_weights(MlSamples smp, int nslots,
                                                     float* __restrict__ wt,
                                                     float* __restrict__ meanw) {
  __shared__ double lds[32];
  __shared__ double mean_s;
  double acc = 0.0, cnt = 0.0;
  for (int k = threadIdx.x; k < nslots; k += blockDim.x)
    if (smp.node[k] >= 0) {
      const double s = smp.s[k];
      acc = acc + s * s;
      cnt = cnt + 1.0;
    }
  acc = wave_sum(acc);
  cnt = wave_sum(cnt);
  if ((threadIdx.x & 63) == 0) {
    lds[threadIdx.x >> 6] = acc;
    lds[16 + (threadIdx.x >> 6)] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, n = 0.0;
    for (int k = 0; k < 16; ++k) {
      a = a + lds[k];
      n = n + lds[16 + k];
    }
    mean_s = n > 0.0 ? a / n : 1.0;
    *meanw = (float)mean_s;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nslots; k += blockDim.x) {
    const double s = smp.s[k];
    wt[k] = smp.node[k] >= 0 ? (float)(s * s / mean_s) : 0.0f;
  }
}

// Gathers the minibatch: X [B][48], X^T [48][B], targets and weights.
__global__ __launch_bounds__(256) void k_ml_batch(MlSamples smp, const float* __restrict__ wt,
                                                  const int* __restrict__ perm, bf16* __restrict__ x,
                                                  bf16* __restrict__ xt, float* __restrict__ y,
                                                  float* __restrict__ w) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ML_BATCH * ML_F) return;
  const int b = g / ML_F, k = g % ML_F;
  const int slot = perm[b];
  const bf16 v = slot >= 0 ? smp.x[(size_t)slot * ML_F + k] : (bf16)0.0f;
  x[g] = v;
  xt[(size_t)k * ML_BATCH + b] = v;
  if (k == 0) {
    y[b] = slot >= 0 ? smp.y[slot] : 0.0f;
    w[b] = slot >= 0 ? wt[slot] : 0.0f;
  }
}

// Z1 += b1 (kept for the backward pass); A1 = bf16(gelu(Z1)) and its transpose.
__global__ __launch_bounds__(256) void k_ml_act1(float* __restrict__ z1,
                                                 const float* __restrict__ master,
                                                 bf16* __restrict__ a1, bf16* __restrict__ a1t) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ML_BATCH * ML_H) return;
  const int b = g / ML_H, j = g % ML_H;
  const float z = z1[g] + master[ML_OFF_B1 + j];
  z1[g] = z;
  const bf16 a = (bf16)ml_gelu(z);
  a1[g] = a;
  a1t[(size_t)j * ML_BATCH + b] = a;
}

// Output layer and its backward pass, 32 samples per 128-thread block
// (thread o = hidden unit o): y, dL/dy, dZ2 (bf16, and its transpose), the
// block's partial sums of dw3 (0..127), db2 (128..255), db3 (256) and loss.
constexpr int ML_SPB = 32;                          // samples per block
constexpr int ML_NBLK = ML_BATCH / ML_SPB;          // blocks per minibatch
constexpr int ML_PO = 2 * ML_H + 1;                 // partials per block
__global__ __launch_bounds__(ML_H) void k_ml_out(const float* __restrict__ z2,
                                                 const float* __restrict__ master,
                                                 const float* __restrict__ y,
                                                 const float* __restrict__ w, float inv_n,
                                                 bf16* __restrict__ dz2, bf16* __restrict__ dz2t,
                                                 float* __restrict__ part,
                                                 float* __restrict__ loss) {
  __shared__ float red[2];
  const int o = threadIdx.x, blk = blockIdx.x;
  const float b2 = master[ML_OFF_B2 + o];
  const float w3 = (float)(bf16)master[ML_OFF_W3 + o];
  const float b3 = master[ML_OFF_B3];
  float gw3 = 0.0f, gb2 = 0.0f, gb3 = 0.0f, l = 0.0f;
  for (int q = 0; q < ML_SPB; ++q) {
    const int b = blk * ML_SPB + q;
    const float z = z2[(size_t)b * ML_H + o] + b2;
    const float a2 = ml_gelu(z);
    const float p = wave_sumf(w3 * a2);
    if ((o & 63) == 0) red[o >> 6] = p;
    __syncthreads();
    const float yb = red[0] + red[1] + b3;
    __syncthreads();
    const float err = yb - y[b];
    const float dy = 2.0f * w[b] * err * inv_n;
    const float dz = dy * w3 * ml_gelu_grad(z);
    const bf16 dzb = (bf16)dz;
    dz2[(size_t)b * ML_H + o] = dzb;
    dz2t[(size_t)o * ML_BATCH + b] = dzb;
    gw3 = gw3 + dy * a2;
    gb2 = gb2 + dz;
    gb3 = gb3 + dy;
    l = l + w[b] * err * err * inv_n;
  }
  part[(size_t)blk * ML_PO + o] = gw3;
  part[(size_t)blk * ML_PO + ML_H + o] = gb2;
  if (o == 0) {
    part[(size_t)blk * ML_PO + 2 * ML_H] = gb3;
    loss[blk] = l;
  }
}

// dZ1 = dA1 * gelu'(Z1) (transposed, bf16) and the block's partial db1.
__global__ __launch_bounds__(ML_H) void k_ml_dz1(const float* __restrict__ da1,
                                                 const float* __restrict__ z1,
                                                 bf16* __restrict__ dz1t,
                                                 float* __restrict__ part) {
  const int j = threadIdx.x, blk = blockIdx.x;
  float gb1 = 0.0f;
  for (int q = 0; q < ML_SPB; ++q) {
    const int b = blk * ML_SPB + q;
    const float dz = da1[(size_t)b * ML_H + j] * ml_gelu_grad(z1[(size_t)b * ML_H + j]);
    dz1t[(size_t)j * ML_BATCH + b] = (bf16)dz;
    gb1 = gb1 + dz;
  }
  part[(size_t)blk * ML_H + j] = gb1;
}

// The gradient of every parameter (fixed-order sums of the block partials),
// written to gout when given; then, with apply, one Adam step (Kingma & Ba
// 2015: beta1 0.9, beta2 0.999, eps 1e-8) on the FP32 master parameters.
__global__ __launch_bounds__(256) void k_ml_adam(float* __restrict__ master, float* __restrict__ m1,
                                                 float* __restrict__ m2,
                                                 const float* __restrict__ gw1,
                                                 const float* __restrict__ gw2,
                                                 const float* __restrict__ po,
                                                 const float* __restrict__ pz,
                                                 const float* __restrict__ loss,
                                                 float* __restrict__ lossacc,
                                                 float* __restrict__ gout, int apply, int step,
                                                 float lr) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ML_NPARAM) return;
  float g = 0.0f;
  if (p < ML_OFF_B1) {
    g = gw1[p];
  } else if (p < ML_OFF_W2) {
    for (int b = 0; b < ML_NBLK; ++b) g = g + pz[(size_t)b * ML_H + (p - ML_OFF_B1)];
  } else if (p < ML_OFF_B2) {
    g = gw2[p - ML_OFF_W2];
  } else if (p < ML_OFF_W3) {
    for (int b = 0; b < ML_NBLK; ++b) g = g + po[(size_t)b * ML_PO + ML_H + (p - ML_OFF_B2)];
  } else if (p < ML_OFF_B3) {
    for (int b = 0; b < ML_NBLK; ++b) g = g + po[(size_t)b * ML_PO + (p - ML_OFF_W3)];
  } else {
    for (int b = 0; b < ML_NBLK; ++b) g = g + po[(size_t)b * ML_PO + 2 * ML_H];
  }
  if (gout) gout[p] = g;
  if (p == 0 && lossacc) {
    float l = 0.0f;
    for (int b = 0; b < ML_NBLK; ++b) l = l + loss[b];
    *lossacc = *lossacc + l;
  }
  if (!apply) return;
  const float b1 = 0.9f, b2 = 0.999f;
  const float m = b1 * m1[p] + (1.0f - b1) * g;
  const float v = b2 * m2[p] + (1.0f - b2) * g * g;
  m1[p] = m;
  m2[p] = v;
  const float mh = m / (1.0f - powf(b1, (float)step));
  const float vh = v / (1.0f - powf(b2, (float)step));
  master[p] = master[p] - lr * mh / (sqrtf(vh) + 1.0e-8f);
}

}  // namespace adcgpu
