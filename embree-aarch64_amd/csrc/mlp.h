// Learned split cost: a small MLP that predicts, from a box and the number
// of primitives under it, log2 of the ratio between the probability that a
// ray of the target distribution visits a node with that box and the box's
// half area relative to the scene's (what uniformly distributed rays would
// give). The binned-SAH builder uses halfArea * 2^prediction in place of the
// half area in its split cost.
//
// Network: 8 inputs -> 32 ReLU -> 32 ReLU -> 1, FP16 weights and
// activations with FP32 accumulation (v_dot2_f32_f16), FP32 biases. Training
// keeps FP32 master weights and runs the forward pass in the same FP16 form
// as inference.
#pragma once
#include "gpu_util.h"

namespace rtg {

constexpr int MLP_IN = 8;
constexpr int MLP_H = 32;
// FP32 parameter vector: W1[H][IN] | b1[H] | W2[H][H] | b2[H] | w3[H] | b3
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_H * MLP_IN;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H;
constexpr int MLP_NPARAM = MLP_OFF_B3 + 1;

typedef _Float16 half2v __attribute__((ext_vector_type(2)));

// Inference form (what the builder reads).
struct MlpPacked {
  _Float16 w1[MLP_H * MLP_IN];
  _Float16 w2[MLP_H * MLP_H];
  _Float16 w3[MLP_H];
  float b1[MLP_H];
  float b2[MLP_H];
  float b3;
};

// Normalisation of boxes to the scene box.
struct SceneNorm {
  float lo[3];
  float inv_ext[3];  // 1/extent (0 for a flat axis)
  float inv_area;    // 1/half area of the scene box (0 if flat)
};

// What the builder needs to use a trained model (null: plain SAH).
struct LearnedCost {
  const MlpPacked* P;
  SceneNorm sn;
  uint32_t min_count;  // segments with at most this many primitives use plain SAH
  float sah_slack;     // candidates whose plain SAH cost exceeds (1 + slack) x the best are not considered
};

// Features of a box holding `count` primitives.
__host__ __device__ inline void box_features(const float lo[3], const float hi[3], uint32_t count, const SceneNorm& sn,
                                             float f[MLP_IN]) {
  for (int k = 0; k < 3; k++) {
    f[k] = ((lo[k] + hi[k]) * 0.5f - sn.lo[k]) * sn.inv_ext[k];
    f[3 + k] = (hi[k] - lo[k]) * sn.inv_ext[k];
  }
  const float a = half_area(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]) * sn.inv_area;
  f[6] = log2f(a + 9.3132257e-10f) * 0.0625f;  // 2^-30 floor
  f[7] = log2f((float)(count > 0 ? count : 1)) * 0.0625f;
}

// Forward pass: z1 = W1 x + b1, h1 = half(relu(z1)), z2 = W2 h1 + b2,
// h2 = half(relu(z2)), y = w3 . h2 + b3; every dot product accumulates FP16
// pairs in FP32 from zero, in index order, and adds the bias last.
__device__ inline float mlp_eval(const MlpPacked* __restrict__ P, const float x[MLP_IN]) {
  half2v xh[MLP_IN / 2];
#pragma unroll
  for (int k = 0; k < MLP_IN / 2; k++) xh[k] = half2v{(_Float16)x[2 * k], (_Float16)x[2 * k + 1]};
  const half2v* w1 = reinterpret_cast<const half2v*>(P->w1);
  const half2v* w2 = reinterpret_cast<const half2v*>(P->w2);
  const half2v* w3 = reinterpret_cast<const half2v*>(P->w3);
  half2v h1[MLP_H / 2];
#pragma unroll
  for (int j = 0; j < MLP_H; j++) {
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < MLP_IN / 2; k++) acc = __builtin_amdgcn_fdot2(w1[j * (MLP_IN / 2) + k], xh[k], acc, false);
    const float z = acc + P->b1[j];
    h1[j >> 1][j & 1] = (_Float16)(z > 0.0f ? z : 0.0f);
  }
  half2v h2[MLP_H / 2];
#pragma unroll
  for (int j = 0; j < MLP_H; j++) {
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < MLP_H / 2; k++) acc = __builtin_amdgcn_fdot2(w2[j * (MLP_H / 2) + k], h1[k], acc, false);
    const float z = acc + P->b2[j];
    h2[j >> 1][j & 1] = (_Float16)(z > 0.0f ? z : 0.0f);
  }
  float acc = 0.0f;
#pragma unroll
  for (int k = 0; k < MLP_H / 2; k++) acc = __builtin_amdgcn_fdot2(w3[k], h2[k], acc, false);
  return acc + P->b3;
}

// Host side of a trainable model (one per RTCGSplitModel handle).
struct SplitModel {
  DevBuf params, m, v, grad, packed, loss;  // FP32 master, Adam moments, gradient, inference form
  DevBuf X, T, W;                           // training set: features, targets, sample weights
  DevBuf idx, visits, counts;               // minibatch indices, visit logging scratch
  uint32_t nsamples = 0;
  uint32_t adam_t = 0;
  SceneNorm norm{};                         // of the scene the data came from
  bool has_norm = false;
};

// Device operations (all on stream s).
void mlp_init(SplitModel& m, uint32_t seed, hipStream_t s);
void mlp_set_params(SplitModel& m, const float* host_params, hipStream_t s);
void mlp_get_params(SplitModel& m, float* host_params, hipStream_t s);
void mlp_predict(SplitModel& m, const float* d_x, uint32_t n, float* d_y, hipStream_t s);
// Loss sum_i w_i (y_i - t_i)^2 / n and its gradient (into m.grad, FP32).
float mlp_gradient(SplitModel& m, const float* d_x, const float* d_t, const float* d_w, const uint32_t* d_idx,
                   uint32_t n, hipStream_t s);
// Appends one training sample per node: features f of info[i] and target
// log2((visits[i] + 0.5) / rays) - 16 f[6] (f[6] = log2 of the relative half
// area / 16).
void mlp_append_samples(SplitModel& m, const float* d_info, const uint32_t* d_visits, uint32_t nnodes, uint32_t rays,
                        const SceneNorm& sn, hipStream_t s);
// Adam on random minibatches of the training set; returns the last loss.
float mlp_train(SplitModel& m, uint32_t steps, uint32_t batch, float lr, uint32_t seed, hipStream_t s);

}  // namespace rtg
