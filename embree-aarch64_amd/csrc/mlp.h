// Learned split cost: a small MLP that corrects the SAH's estimate of how
// often a child box is entered when its parent box is. For a candidate child
// C of a segment S -- the per-bin statistics a binned split has for each
// candidate: C's box and primitive count from its bins, S's box and count --
// it predicts y = log2(P(C | S)) - log2(A_C / A_S), where P(C | S) is the
// fraction of the rays of the target distribution that enter S and then C,
// and A is the half area (the SAH assumes y = 0). The binned-SAH builder
// prices the split A_L 2^y_L n_L + A_R 2^y_R n_R in place of
// A_L n_L + A_R n_R.
//
// Network: 12 inputs -> 32 ReLU -> 32 ReLU -> 1, FP16 weights and
// activations with FP32 accumulation (v_dot2_f32_f16), FP32 biases. Training
// keeps FP32 master weights and runs the forward pass in the same FP16 form
// as inference.
#pragma once
#include "bvh_types.h"
#include "gpu_util.h"
#include "sort.h"

namespace rtg {

constexpr int MLP_IN = 12;
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

constexpr float MLP_AREA_FLOOR = 9.3132257e-10f;  // 2^-30
constexpr float MLP_Y_MIN = -3.0f, MLP_Y_MAX = 3.0f;  // bounds of a prediction in use

// Features of child box C (clo..chi, cn primitives) of parent box S
// (plo..phi, pn primitives):
//   f0     log2(A_C / A_S) / 16             the SAH's estimate of P(C | S)
//   f1     n_C / n_S
//   f2-f4  (centre_C - centre_S) / extent_S  per axis (0 on a flat axis)
//   f5-f7  extent_C / extent_S              per axis (1 on a flat axis)
//   f8-f10 (centre_S - scene lo) / scene extent
//   f11    log2(A_S / A_scene) / 16
__host__ __device__ inline void split_features(const float clo[3], const float chi[3], uint32_t cn, const float plo[3],
                                               const float phi[3], uint32_t pn, const SceneNorm& sn,
                                               float f[MLP_IN]) {
  const float ac = half_area(chi[0] - clo[0], chi[1] - clo[1], chi[2] - clo[2]);
  const float ap = half_area(phi[0] - plo[0], phi[1] - plo[1], phi[2] - plo[2]);
  f[0] = ap > 0.0f ? log2f(ac / ap + MLP_AREA_FLOOR) * 0.0625f : 0.0f;
  f[1] = pn > 0 ? (float)cn / (float)pn : 1.0f;
  for (int k = 0; k < 3; k++) {
    const float ep = phi[k] - plo[k];
    const float inv = ep > 0.0f ? 1.0f / ep : 0.0f;
    f[2 + k] = ((chi[k] + clo[k]) - (phi[k] + plo[k])) * 0.5f * inv;
    f[5 + k] = ep > 0.0f ? (chi[k] - clo[k]) * inv : 1.0f;
    f[8 + k] = ((phi[k] + plo[k]) * 0.5f - sn.lo[k]) * sn.inv_ext[k];
  }
  f[11] = log2f(ap * sn.inv_area + MLP_AREA_FLOOR) * 0.0625f;
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
  DevBuf idx, visits, counts, scount;       // minibatch indices, visit logging scratch, appended-sample count
  DevBuf part, lpart;                       // per-block gradient and loss partial sums
  DevBuf skey, skey2, sval, xtmp, ttmp, wtmp;  // canonical ordering of appended samples
  RadixSorter sorter;
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
// Appends one training sample per (parent, internal child) pair of the
// nodes (a binary tree: width 2) whose parent was entered by at least
// MLP_MIN_VISITS rays: split_features of the child's box and count against
// the parent's (info: NodeInfo per node) and the target
// log2((v_C + 0.5) / (v_S + 0.5)) - log2(A_C / A_S + 2^-30), where v counts
// the rays that entered a node. The appended samples are put in an order
// fixed by their contents (a 64-bit hash of features and target), so that
// the node numbering of the tree they came from, which the build does not
// fix, does not change the minibatches training draws.
constexpr uint32_t MLP_MIN_VISITS = 8;
void mlp_append_samples(SplitModel& m, const QNode8* d_nodes, const NodeInfo* d_info, const uint32_t* d_visits,
                        uint32_t nnodes, const SceneNorm& sn, hipStream_t s);
// Adam on random minibatches of the training set; returns the last loss.
// The gradient sums each block's samples and then the blocks in index
// order, so training is deterministic for a given training set and seed.
float mlp_train(SplitModel& m, uint32_t steps, uint32_t batch, float lr, uint32_t seed, hipStream_t s);

}  // namespace rtg
