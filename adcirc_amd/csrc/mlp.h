// The learned JCG warm start: a per-node MLP (48 -> 128 -> 128 -> 1, GELU)
// that predicts the correction to the reference's initial guess, with BF16
// weights and FP32 accumulation on the gfx950 matrix cores.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

namespace adcgpu {

constexpr int ML_F = 48;        // inputs
constexpr int ML_H = 128;       // hidden width
constexpr int ML_NB = 12;       // stencil neighbours (NEITAB slots 2..13)
constexpr int ML_S = ML_NB + 1; // stencil nodes, the node itself first
constexpr int ML_BATCH = 4096;  // training minibatch (samples)
constexpr int ML_W1 = ML_F * ML_H, ML_W2 = ML_H * ML_H;
// master parameters, FP32, in this order: W1[48][128], b1[128], W2[128][128],
// b2[128], w3[128], b3
constexpr int ML_OFF_W1 = 0, ML_OFF_B1 = ML_W1, ML_OFF_W2 = ML_OFF_B1 + ML_H,
              ML_OFF_B2 = ML_OFF_W2 + ML_W2, ML_OFF_W3 = ML_OFF_B2 + ML_H,
              ML_OFF_B3 = ML_OFF_W3 + ML_H, ML_NPARAM = ML_OFF_B3 + 1;

typedef __bf16 bf16;

// Weights in the layouts the kernels read (written by k_ml_pack):
//  - w1p, w2p: the matrix-core operand fragments of the inference kernel,
//    lane by lane (see k_ml_mlp);
//  - b1p, b2p, w3p: biases and output weights in the same lane order;
//  - w1t [128][48], w2t [128 out][128 in], w2n [128 in][128 out]: row-major
//    copies for the training GEMMs (K contiguous).
struct MlWeights {
  bf16 *w1p, *w2p, *w1t, *w2t, *w2n;
  float *b1p, *b2p, *w3p;  // [2 halves][4 tiles][16]
  float* b3;
};

// The step's inputs of the features of node i.
struct MlIn {
  int np;
  const int* nneigh;
  const int* col;        // [mnei*np] 0-based, slot 0 = the node
  const int* bnd_index;  // elevation-boundary rows get no correction
  const double* eta;     // eta^k (before this step's solve)
  const double* etas;    // ETAS^k, the last increment
  const double* etasp;   // ETAS^(k-1)
  const double* dp;
  const double* wsx2;    // this step's wind stress / rho0
  const double* wsy2;
  const double* pr1;     // pressure head, last and this step
  const double* pr2;
  const double* tk;
  const int* nodecode;
  double dt;
};

// Training samples, a ring of `cap` node snapshots: features (bf16, the
// kernels' k order), the node's scale s, its ETAS^k, its node id (-1: no
// sample), then the target (ETAS^(k+1) - ETAS^k) / s once the step is solved.
struct MlSamples {
  bf16* x;        // [cap][48]
  float* s;       // [cap]
  double* e0;     // [cap]
  int* node;      // [cap]
  float* y;       // [cap]
  int cap;
};

__global__ void k_ml_pack(const float* master, MlWeights w);
// inference: the inputs of every node (fx: 3 x 64 fragments of 8 bf16 per
// tile of 32 nodes; sv: the nodes' scales), then the network over them
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
__global__ void k_ml_feat(MlIn in, bf16x8* fx, double* sv);
__global__ void k_ml_mlp(MlIn in, MlWeights w, const bf16x8* fx, const double* sv, double* corr);
__global__ void k_ml_collect(MlIn in, MlSamples smp, int slot0, int count, int it);
__global__ void k_ml_target(MlSamples smp, int slot0, int count, const double* etas_new);
// training (minibatches of ML_BATCH samples; see mlp.hip)
constexpr int ML_TRAIN_BLOCKS = ML_BATCH / 32;  // blocks of k_ml_out / k_ml_dz1
constexpr int ML_PART_OUT = 2 * ML_H + 1;        // their partial sums per block
__global__ void k_ml_gemm_nt(int M, int N, int K, const bf16* A, int lda, const bf16* B, int ldb,
                             float* C, int ldc);
__global__ void k_ml_weights(MlSamples smp, int nslots, float* wt, float* meanw);
__global__ void k_ml_batch(MlSamples smp, const float* wt, const int* perm, bf16* x, bf16* xt,
                           float* y, float* w);
__global__ void k_ml_act1(float* z1, const float* master, bf16* a1, bf16* a1t);
__global__ void k_ml_out(const float* z2, const float* master, const float* y, const float* w,
                         float inv_n, bf16* dz2, bf16* dz2t, float* part, float* loss);
__global__ void k_ml_dz1(const float* da1, const float* z1, bf16* dz1t, float* part);
__global__ void k_ml_adam(float* master, float* m1, float* m2, const float* gw1, const float* gw2,
                          const float* po, const float* pz, const float* loss, float* lossacc,
                          float* gout, int apply, int step, float lr);

}  // namespace adcgpu
