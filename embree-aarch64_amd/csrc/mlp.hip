// Training and inference kernels of the learned split-cost MLP.
//
// Training step (one minibatch): k_mlp_grad runs one sample per lane,
// recomputes the FP16 forward pass, back-propagates in FP32 through the FP16
// weights, stages per-sample activations and deltas in LDS and sums each
// parameter's gradient over the wave's 64 samples before one atomic add;
// k_mlp_adam then updates the FP32 master weights and rewrites the FP16
// inference form. Loss: sum_i w_i (y_i - t_i)^2 / n.
#include <cmath>
#include <vector>

#include "bvh_types.h"
#include "mlp.h"

namespace rtg {
namespace {

constexpr int WB = 64;        // samples per gradient block (one wave)
constexpr int ROW = MLP_H + 1;  // padded LDS row

__device__ inline void pack_one(MlpPacked* P, int i, float v) {
  if (i < MLP_OFF_B1) P->w1[i - MLP_OFF_W1] = (_Float16)v;
  else if (i < MLP_OFF_W2) P->b1[i - MLP_OFF_B1] = v;
  else if (i < MLP_OFF_B2) P->w2[i - MLP_OFF_W2] = (_Float16)v;
  else if (i < MLP_OFF_W3) P->b2[i - MLP_OFF_B2] = v;
  else if (i < MLP_OFF_B3) P->w3[i - MLP_OFF_W3] = (_Float16)v;
  else P->b3 = v;
}

__global__ void k_mlp_pack(const float* __restrict__ p, MlpPacked* P) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < MLP_NPARAM) pack_one(P, i, p[i]);
}

// Deterministic initialisation: uniform(-a, a) with a = sqrt(6 / fan_in)
// (He-uniform for ReLU layers) from a counter hash; biases zero, so an
// untrained model predicts a visit probability near the relative half area
// (the plain SAH).
__device__ inline float hash_uniform(uint32_t seed, uint32_t i) {
  uint64_t z = ((uint64_t)seed << 32) + i + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (float)(z >> 40) * (1.0f / 16777216.0f) * 2.0f - 1.0f;
}

__global__ void k_mlp_init(float* p, float* m, float* v, float* g, uint32_t seed) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= MLP_NPARAM) return;
  float x = 0.0f;
  if (i < MLP_OFF_B1) x = hash_uniform(seed, i) * sqrtf(6.0f / MLP_IN);
  else if (i >= MLP_OFF_W2 && i < MLP_OFF_B2) x = hash_uniform(seed, i) * sqrtf(6.0f / MLP_H);
  else if (i >= MLP_OFF_W3 && i < MLP_OFF_B3) x = hash_uniform(seed, i) * sqrtf(3.0f / MLP_H);
  p[i] = x;
  m[i] = v[i] = g[i] = 0.0f;
}

__global__ void k_mlp_predict(const MlpPacked* __restrict__ P, const float* __restrict__ X, uint32_t n,
                              float* __restrict__ Y) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float x[MLP_IN];
#pragma unroll
  for (int k = 0; k < MLP_IN; k++) x[k] = X[(size_t)i * MLP_IN + k];
  Y[i] = mlp_eval(P, x);
}

__global__ __launch_bounds__(WB) void k_mlp_grad(const MlpPacked* __restrict__ P, const float* __restrict__ X,
                                                 const float* __restrict__ T, const float* __restrict__ W,
                                                 const uint32_t* __restrict__ idx, uint32_t n, float scale,
                                                 float* __restrict__ part, float* __restrict__ lpart) {
  __shared__ float s_x[WB][MLP_IN];
  __shared__ float s_h1[WB][ROW], s_d1[WB][ROW], s_h2[WB][ROW], s_d2[WB][ROW];
  __shared__ float s_dy[WB];
  const int lane = threadIdx.x;
  const uint32_t i = blockIdx.x * WB + lane;
  const bool valid = i < n;
  const uint32_t src = valid ? (idx ? idx[i] : i) : 0u;
  float x[MLP_IN];
#pragma unroll
  for (int k = 0; k < MLP_IN; k++) x[k] = valid ? X[(size_t)src * MLP_IN + k] : 0.0f;
  // forward, FP16 exactly as mlp_eval; z kept in the delta rows for the masks
  half2v xh[MLP_IN / 2];
#pragma unroll
  for (int k = 0; k < MLP_IN / 2; k++) xh[k] = half2v{(_Float16)x[2 * k], (_Float16)x[2 * k + 1]};
#pragma unroll
  for (int k = 0; k < MLP_IN; k++) s_x[lane][k] = (float)(_Float16)x[k];
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
    const _Float16 h = (_Float16)(z > 0.0f ? z : 0.0f);
    h1[j >> 1][j & 1] = h;
    s_h1[lane][j] = (float)h;
    s_d1[lane][j] = z;
  }
  half2v h2[MLP_H / 2];
#pragma unroll
  for (int j = 0; j < MLP_H; j++) {
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < MLP_H / 2; k++) acc = __builtin_amdgcn_fdot2(w2[j * (MLP_H / 2) + k], h1[k], acc, false);
    const float z = acc + P->b2[j];
    const _Float16 h = (_Float16)(z > 0.0f ? z : 0.0f);
    h2[j >> 1][j & 1] = h;
    s_h2[lane][j] = (float)h;
    s_d2[lane][j] = z;
  }
  float acc = 0.0f;
#pragma unroll
  for (int k = 0; k < MLP_H / 2; k++) acc = __builtin_amdgcn_fdot2(w3[k], h2[k], acc, false);
  const float y = acc + P->b3;
  const float wt = valid ? W[src] : 0.0f;
  const float e = valid ? y - T[src] : 0.0f;
  const float dy = scale * wt * e;
  s_dy[lane] = dy;
  // backward through the FP16 weights, in FP32
#pragma unroll
  for (int j = 0; j < MLP_H; j++) {
    const float z = s_d2[lane][j];
    s_d2[lane][j] = z > 0.0f ? dy * (float)P->w3[j] : 0.0f;
  }
  for (int k = 0; k < MLP_H; k++) {
    float d = 0.0f;
    for (int j = 0; j < MLP_H; j++) d += (float)P->w2[j * MLP_H + k] * s_d2[lane][j];
    const float z = s_d1[lane][k];
    s_d1[lane][k] = z > 0.0f ? d : 0.0f;
  }
  // loss
  float l = wt * e * e;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) l += __shfl_xor(l, d, 64);
  if (lane == 0) lpart[blockIdx.x] = l;
  __syncthreads();
  // parameter gradients summed over the block's samples
  for (int p = lane; p < MLP_NPARAM; p += WB) {
    float s = 0.0f;
    if (p < MLP_OFF_B1) {
      const int j = p / MLP_IN, k = p % MLP_IN;
      for (int q = 0; q < WB; q++) s += s_d1[q][j] * s_x[q][k];
    } else if (p < MLP_OFF_W2) {
      const int j = p - MLP_OFF_B1;
      for (int q = 0; q < WB; q++) s += s_d1[q][j];
    } else if (p < MLP_OFF_B2) {
      const int j = (p - MLP_OFF_W2) / MLP_H, k = (p - MLP_OFF_W2) % MLP_H;
      for (int q = 0; q < WB; q++) s += s_d2[q][j] * s_h1[q][k];
    } else if (p < MLP_OFF_W3) {
      const int j = p - MLP_OFF_B2;
      for (int q = 0; q < WB; q++) s += s_d2[q][j];
    } else if (p < MLP_OFF_B3) {
      const int j = p - MLP_OFF_W3;
      for (int q = 0; q < WB; q++) s += s_dy[q] * s_h2[q][j];
    } else {
      for (int q = 0; q < WB; q++) s += s_dy[q];
    }
    part[(size_t)blockIdx.x * MLP_NPARAM + p] = s;
  }
}

// Sums the blocks' partial gradients (and losses) in block order.
__global__ void k_mlp_reduce(const float* __restrict__ part, const float* __restrict__ lpart, uint32_t nblocks,
                             float* __restrict__ grad, float* __restrict__ loss) {
  const uint32_t p = blockIdx.x * 256 + threadIdx.x;
  if (p < MLP_NPARAM) {
    float s = 0.0f;
    for (uint32_t b = 0; b < nblocks; b++) s += part[(size_t)b * MLP_NPARAM + p];
    grad[p] = s;
  } else if (p == MLP_NPARAM) {
    float s = 0.0f;
    for (uint32_t b = 0; b < nblocks; b++) s += lpart[b];
    *loss = s;
  }
}

// 64-bit content key of sample i (features then target, as bits): two
// independent 32-bit mixes.
__device__ inline uint32_t mix32(uint32_t h, uint32_t v) {
  h ^= v * 0x9E3779B1u;
  h = (h ^ (h >> 15)) * 0x85EBCA77u;
  return h ^ (h >> 13);
}
__global__ void k_mlp_sample_keys(const float* __restrict__ X, const float* __restrict__ T, uint32_t n,
                                  uint32_t* __restrict__ klo, uint32_t* __restrict__ khi, uint32_t* __restrict__ val) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  uint32_t a = 0x243F6A88u, b = 0xB7E15162u;
  for (int k = 0; k < MLP_IN; k++) {
    const uint32_t v = __float_as_uint(X[(size_t)i * MLP_IN + k]);
    a = mix32(a, v);
    b = mix32(b, v ^ 0x5bd1e995u);
  }
  a = mix32(a, __float_as_uint(T[i]));
  b = mix32(b, __float_as_uint(T[i]) ^ 0x5bd1e995u);
  klo[i] = a;
  khi[i] = b;
  val[i] = i;
}
__global__ void k_gather_u32(const uint32_t* __restrict__ src, const uint32_t* __restrict__ idx, uint32_t n,
                             uint32_t* __restrict__ dst) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}
__global__ void k_mlp_permute(const float* __restrict__ X, const float* __restrict__ T, const float* __restrict__ W,
                              const uint32_t* __restrict__ idx, uint32_t n, float* __restrict__ Xo,
                              float* __restrict__ To, float* __restrict__ Wo) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = idx[i];
  for (int k = 0; k < MLP_IN; k++) Xo[(size_t)i * MLP_IN + k] = X[(size_t)j * MLP_IN + k];
  To[i] = T[j];
  Wo[i] = W[j];
}

__global__ void k_mlp_adam(float* p, float* m, float* v, float* g, float lr, float bc1, float bc2, MlpPacked* P) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= MLP_NPARAM) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  p[i] = pi;
  g[i] = 0.0f;
  pack_one(P, i, pi);
}

// One thread per parent node: a sample per internal child, appended at a
// slot reserved with one atomic per node.
__global__ void k_mlp_samples(const QNode8* __restrict__ nodes, const NodeInfo* __restrict__ info,
                              const uint32_t* __restrict__ visits, uint32_t nnodes, SceneNorm sn,
                              float* __restrict__ X, float* __restrict__ T, float* __restrict__ W,
                              uint32_t* __restrict__ count, uint32_t cap) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nnodes) return;
  const uint32_t vs = visits[i];
  if (vs < MLP_MIN_VISITS) return;
  const QNode8 q = nodes[i];
  uint32_t kids[8];
  int nk = 0;
  for (int c = 0; c < 8; c++)
    if (q.meta[c] != 0 && ((q.imask >> c) & 1u)) kids[nk++] = q.child_base + (q.meta[c] & 0x7fu);
  if (nk == 0) return;
  const uint32_t base = atomicAdd(count, (uint32_t)nk);
  const NodeInfo ps = info[i];
  for (int j = 0; j < nk; j++) {
    const uint32_t slot = base + (uint32_t)j;
    if (slot >= cap || kids[j] >= nnodes) break;
    const NodeInfo pc = info[kids[j]];
    float f[MLP_IN];
    split_features(pc.lo, pc.hi, pc.count, ps.lo, ps.hi, ps.count, sn, f);
#pragma unroll
    for (int k = 0; k < MLP_IN; k++) X[(size_t)slot * MLP_IN + k] = f[k];
    // departure of the measured P(C | S) from the SAH's area ratio
    T[slot] = log2f(((float)visits[kids[j]] + 0.5f) / ((float)vs + 0.5f)) - 16.0f * f[0];
    W[slot] = 1.0f;
  }
}

// minibatch indices: a multiplicative permutation of a random window
__global__ void k_mlp_batch(uint32_t* idx, uint32_t batch, uint32_t n, uint32_t seed) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= batch) return;
  uint64_t z = ((uint64_t)seed << 32) + i;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  idx[i] = (uint32_t)(z % n);
}

void ensure(SplitModel& m) {
  m.params.reserve(MLP_NPARAM * 4);
  m.m.reserve(MLP_NPARAM * 4);
  m.v.reserve(MLP_NPARAM * 4);
  m.grad.reserve(MLP_NPARAM * 4);
  m.packed.reserve(sizeof(MlpPacked));
  m.loss.reserve(256);
}

}  // namespace

void mlp_init(SplitModel& m, uint32_t seed, hipStream_t s) {
  ensure(m);
  hipLaunchKernelGGL(k_mlp_init, dim3(div_up(MLP_NPARAM, 256)), dim3(256), 0, s, m.params.as<float>(), m.m.as<float>(),
                     m.v.as<float>(), m.grad.as<float>(), seed);
  hipLaunchKernelGGL(k_mlp_pack, dim3(div_up(MLP_NPARAM, 256)), dim3(256), 0, s, m.params.as<float>(),
                     m.packed.as<MlpPacked>());
  RTG_HIP(hipGetLastError());
  m.adam_t = 0;
}

void mlp_set_params(SplitModel& m, const float* host_params, hipStream_t s) {
  ensure(m);
  RTG_HIP(hipMemcpyAsync(m.params.ptr, host_params, MLP_NPARAM * 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_mlp_pack, dim3(div_up(MLP_NPARAM, 256)), dim3(256), 0, s, m.params.as<float>(),
                     m.packed.as<MlpPacked>());
  RTG_HIP(hipGetLastError());
}

void mlp_get_params(SplitModel& m, float* host_params, hipStream_t s) {
  ensure(m);
  RTG_HIP(hipMemcpyAsync(host_params, m.params.ptr, MLP_NPARAM * 4, hipMemcpyDeviceToHost, s));
  RTG_HIP(hipStreamSynchronize(s));
}

void mlp_predict(SplitModel& m, const float* d_x, uint32_t n, float* d_y, hipStream_t s) {
  ensure(m);
  if (n == 0) return;
  hipLaunchKernelGGL(k_mlp_predict, dim3(div_up(n, 256)), dim3(256), 0, s, m.packed.as<MlpPacked>(), d_x, n, d_y);
  RTG_HIP(hipGetLastError());
}

// Gradient and loss of n samples into m.grad and m.loss (block partial sums,
// then their sum in block order).
void grad_step(SplitModel& m, const float* d_x, const float* d_t, const float* d_w, const uint32_t* d_idx, uint32_t n,
               hipStream_t s) {
  const uint32_t nb = div_up(n, WB);
  m.part.reserve((size_t)nb * MLP_NPARAM * 4);
  m.lpart.reserve((size_t)nb * 4);
  hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(WB), 0, s, m.packed.as<MlpPacked>(), d_x, d_t, d_w, d_idx, n,
                     2.0f / (float)n, m.part.as<float>(), m.lpart.as<float>());
  hipLaunchKernelGGL(k_mlp_reduce, dim3(div_up(MLP_NPARAM + 1, 256)), dim3(256), 0, s, m.part.as<float>(),
                     m.lpart.as<float>(), nb, m.grad.as<float>(), m.loss.as<float>());
  RTG_HIP(hipGetLastError());
}

float mlp_gradient(SplitModel& m, const float* d_x, const float* d_t, const float* d_w, const uint32_t* d_idx,
                   uint32_t n, hipStream_t s) {
  ensure(m);
  RTG_HIP(hipMemsetAsync(m.grad.ptr, 0, MLP_NPARAM * 4, s));
  RTG_HIP(hipMemsetAsync(m.loss.ptr, 0, 4, s));
  if (n == 0) return 0.0f;
  grad_step(m, d_x, d_t, d_w, d_idx, n, s);
  float l = 0.0f;
  RTG_HIP(hipMemcpyAsync(&l, m.loss.ptr, 4, hipMemcpyDeviceToHost, s));
  RTG_HIP(hipStreamSynchronize(s));
  return l / (float)n;
}

void mlp_append_samples(SplitModel& m, const QNode8* d_nodes, const NodeInfo* d_info, const uint32_t* d_visits,
                        uint32_t nnodes, const SceneNorm& sn, hipStream_t s) {
  if (nnodes == 0) return;
  const size_t room = (size_t)nnodes * 8;  // at most one sample per child slot
  const size_t total = (size_t)m.nsamples + room;
  if (total > 0xffffffffull) throw hip_error(hipErrorInvalidValue, "split-model training set too large");
  // grow (keeping the samples already there)
  auto grow = [&](DevBuf& b, size_t bytes_per) {
    if (b.bytes >= total * bytes_per) return;
    DevBuf nb;
    nb.reserve(total * bytes_per * 2);
    if (m.nsamples) RTG_HIP(hipMemcpyAsync(nb.ptr, b.ptr, (size_t)m.nsamples * bytes_per, hipMemcpyDeviceToDevice, s));
    RTG_HIP(hipStreamSynchronize(s));
    std::swap(b.ptr, nb.ptr);
    std::swap(b.bytes, nb.bytes);
  };
  grow(m.X, MLP_IN * 4);
  grow(m.T, 4);
  grow(m.W, 4);
  m.scount.reserve(256);
  uint32_t* cnt = m.scount.as<uint32_t>();
  RTG_HIP(hipMemsetAsync(cnt, 0, 4, s));
  hipLaunchKernelGGL(k_mlp_samples, dim3(div_up(nnodes, 256)), dim3(256), 0, s, d_nodes, d_info, d_visits, nnodes, sn,
                     m.X.as<float>() + (size_t)m.nsamples * MLP_IN, m.T.as<float>() + m.nsamples,
                     m.W.as<float>() + m.nsamples, cnt, (uint32_t)room);
  RTG_HIP(hipGetLastError());
  uint32_t added = 0;
  RTG_HIP(hipMemcpyAsync(&added, cnt, 4, hipMemcpyDeviceToHost, s));
  RTG_HIP(hipStreamSynchronize(s));
  if (added > room) added = (uint32_t)room;
  if (added > 1) {
    // the new samples in the order of their 64-bit content keys (hi, lo)
    float* X0 = m.X.as<float>() + (size_t)m.nsamples * MLP_IN;
    float* T0 = m.T.as<float>() + m.nsamples;
    float* W0 = m.W.as<float>() + m.nsamples;
    m.skey.reserve((size_t)added * 4);
    m.skey2.reserve((size_t)added * 4);
    m.sval.reserve((size_t)added * 4);
    m.xtmp.reserve((size_t)added * MLP_IN * 4);
    m.ttmp.reserve((size_t)added * 4);
    m.wtmp.reserve((size_t)added * 4);
    uint32_t* klo = m.skey.as<uint32_t>();
    uint32_t* khi = m.skey2.as<uint32_t>();
    uint32_t* val = m.sval.as<uint32_t>();
    hipLaunchKernelGGL(k_mlp_sample_keys, dim3(div_up(added, 256)), dim3(256), 0, s, X0, T0, added, klo, khi, val);
    m.sorter.sort_pairs(klo, val, added, 0, 32, s);
    // the high keys in the low keys' order, then a stable sort on them
    hipLaunchKernelGGL(k_gather_u32, dim3(div_up(added, 256)), dim3(256), 0, s, khi, val, added, klo);
    m.sorter.sort_pairs(klo, val, added, 0, 32, s);
    hipLaunchKernelGGL(k_mlp_permute, dim3(div_up(added, 256)), dim3(256), 0, s, X0, T0, W0, val, added,
                       m.xtmp.as<float>(), m.ttmp.as<float>(), m.wtmp.as<float>());
    RTG_HIP(hipMemcpyAsync(X0, m.xtmp.ptr, (size_t)added * MLP_IN * 4, hipMemcpyDeviceToDevice, s));
    RTG_HIP(hipMemcpyAsync(T0, m.ttmp.ptr, (size_t)added * 4, hipMemcpyDeviceToDevice, s));
    RTG_HIP(hipMemcpyAsync(W0, m.wtmp.ptr, (size_t)added * 4, hipMemcpyDeviceToDevice, s));
    RTG_HIP(hipGetLastError());
  }
  m.nsamples += added;
}

float mlp_train(SplitModel& m, uint32_t steps, uint32_t batch, float lr, uint32_t seed, hipStream_t s) {
  ensure(m);
  if (m.nsamples == 0 || steps == 0) return 0.0f;
  if (batch == 0 || batch > m.nsamples) batch = m.nsamples;
  m.idx.reserve((size_t)batch * 4);
  float last = 0.0f;
  for (uint32_t it = 0; it < steps; it++) {
    hipLaunchKernelGGL(k_mlp_batch, dim3(div_up(batch, 256)), dim3(256), 0, s, m.idx.as<uint32_t>(), batch, m.nsamples,
                       seed * 7919u + it);
    grad_step(m, m.X.as<float>(), m.T.as<float>(), m.W.as<float>(), m.idx.as<uint32_t>(), batch, s);
    m.adam_t++;
    const float bc1 = 1.0f - powf(0.9f, (float)m.adam_t), bc2 = 1.0f - powf(0.999f, (float)m.adam_t);
    hipLaunchKernelGGL(k_mlp_adam, dim3(div_up(MLP_NPARAM, 256)), dim3(256), 0, s, m.params.as<float>(),
                       m.m.as<float>(), m.v.as<float>(), m.grad.as<float>(), lr, bc1, bc2, m.packed.as<MlpPacked>());
    RTG_HIP(hipGetLastError());
    if (it + 1 == steps) {
      RTG_HIP(hipMemcpyAsync(&last, m.loss.ptr, 4, hipMemcpyDeviceToHost, s));
      RTG_HIP(hipStreamSynchronize(s));
    }
  }
  return last / (float)batch;
}

}  // namespace rtg
