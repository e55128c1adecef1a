// GPU BVH8 builder (see builder.h for the pipeline).
//
// Reference anchors for what each phase computes:
//   primitive references and validity: kernels/common/scene_triangle_mesh.h:127-147,220-231,
//     kernels/builders/primrefgen.cpp:36-56 (invalid triangles are skipped);
//   Morton codes on 2x centroids, 10 bits per axis: kernels/builders/bvh_builder_morton.h:70-101;
//   radix sort: common/algorithms/parallel_sort.h:437-456 (semantics only);
//   leaf/inner decision by SAH cost and the 8-wide collapse that repeatedly
//     opens the child with the largest half area: kernels/builders/bvh_builder_sah.h:222-330;
//   leaf fill e1 = v0-v1, e2 = v2-v0: kernels/geometry/triangle.h:40-41,98-120.
#include <chrono>
#include <cstring>
#include <string>

#include "builder.h"

namespace rtg {

struct BuildCounters {
  uint32_t glo[3], ghi[3];
  uint32_t clo[3], chi[3];
  uint32_t maxabs;
  uint32_t valid;
  uint32_t node_count;
  uint32_t tri_count;
  uint32_t qcount[3];  // collapse level k reads slot (k + 2) % 3, fills k % 3, clears (k + 1) % 3
  uint32_t max_depth;
  uint32_t error;  // bit 0: refit walk too long, bit 1: collapse index out of range, bit 2: bad tree
  uint32_t max_stack;  // most traversal stack entries any root-to-node path can need
  uint32_t roots_seen; // tree roots met by the tree check
  uint32_t deep;       // a radix-tree node deeper than 254 levels (a broken parent link)
  double sah_nodes, sah_leaves, sah_leaves_tri;
};

constexpr uint32_t NONE = 0xffffffffu;
constexpr int BT = 256;

struct QItem {
  uint32_t bref;  // binary ref: REF_LEAF | sorted prim index, or inner node index
  uint32_t node;  // QNode8 slot
  uint32_t depth;
  uint32_t need;  // stack entries the traversal may hold on reaching this node
};

// ------------------------------------------------------------ reductions --

__device__ inline float wave_min(float v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = fminf(v, __shfl_xor(v, d, 64));
  return v;
}
__device__ inline float wave_max(float v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
  return v;
}
__device__ inline uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}
__device__ inline uint32_t wave_umax(uint32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = max(v, (uint32_t)__shfl_xor(v, d, 64));
  return v;
}
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// Per-scene reductions of a batched build.
struct SceneAcc {
  uint32_t glo[3], ghi[3], clo[3], chi[3];
  uint32_t valid;
  uint32_t pad[3];
};

__global__ void k_init_sacc(SceneAcc* a, uint32_t n) {
  const uint32_t i = blockIdx.x * BT + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < 3; k++) {
    a[i].glo[k] = a[i].clo[k] = f2o(INFINITY);
    a[i].ghi[k] = a[i].chi[k] = f2o(-INFINITY);
  }
  a[i].valid = 0;
}

__global__ void k_fill_none(uint32_t* v, uint32_t n) {
  const uint32_t i = blockIdx.x * BT + threadIdx.x;
  if (i < n) v[i] = 0xffffffffu;
}

__global__ void k_iota(uint32_t* v, uint32_t n) {
  const uint32_t i = blockIdx.x * BT + threadIdx.x;
  if (i < n) v[i] = i;
}

__global__ void k_init_counters(BuildCounters* c, uint32_t nscenes) {
  for (int k = 0; k < 3; k++) {
    c->glo[k] = f2o(INFINITY);
    c->ghi[k] = f2o(-INFINITY);
    c->clo[k] = f2o(INFINITY);
    c->chi[k] = f2o(-INFINITY);
  }
  c->maxabs = f2o(0.0f);
  c->valid = 0;
  c->node_count = nscenes;
  c->roots_seen = 0;
  c->deep = 0;
  c->tri_count = 0;
  c->qcount[0] = 0;
  c->qcount[1] = 0;
  c->qcount[2] = nscenes;  // the collapse's first level reads slot 2: one root item per scene
  c->max_depth = 0;
  c->max_stack = 0;
  c->error = 0;
  c->sah_nodes = c->sah_leaves = c->sah_leaves_tri = 0.0;
}

// ------------------------------------------------------------ primrefs --

__device__ inline bool finite_v(float x) { return (x > -FLT_LARGE_) & (x < FLT_LARGE_); }

// Primitive references of PR_PER * BT consecutive primitives per block
// (PR_PER per thread, BT apart), with the scene bounds and valid count
// reduced over the block before one set of global atomics: one set per
// BT primitives would serialise on those few addresses.
constexpr int PR_PER = 4;

__global__ __launch_bounds__(BT) void k_primrefs(const GeomDesc* __restrict__ geoms, uint32_t ngeoms,
                                                 uint32_t total, float4* __restrict__ lo4,
                                                 float4* __restrict__ hi4, uint32_t* __restrict__ keys,
                                                 BuildCounters* __restrict__ cnt, SceneAcc* __restrict__ sacc) {
  float b[13];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    b[k] = INFINITY;        // geom lo
    b[3 + k] = -INFINITY;   // geom hi
    b[6 + k] = INFINITY;    // cent lo
    b[9 + k] = -INFINITY;   // cent hi
  }
  b[12] = 0.0f;
  uint32_t valid = 0, smin = 0xffffffffu, smax = 0u;
  const uint32_t base = blockIdx.x * (BT * PR_PER) + threadIdx.x;
#pragma unroll
  for (int j = 0; j < PR_PER; j++) {
    const uint32_t i = base + (uint32_t)j * BT;
    if (i >= total) break;
    uint32_t lo = 0, hi = ngeoms - 1;
    while (lo < hi) {  // last geometry with prim_offset <= i
      uint32_t mid = (lo + hi + 1) >> 1;
      if (geoms[mid].prim_offset <= i) lo = mid; else hi = mid - 1;
    }
    const GeomDesc g = geoms[lo];
    const uint32_t prim = i - g.prim_offset;
    bool ok;
    float l[3], h[3];
    if (g.kind == GEOM_INSTANCES) {  // a precomputed world box per instance record
      const float* bx = (const float*)(g.vertices + (uint64_t)prim * g.vstride);
      ok = true;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        l[k] = bx[k];
        h[k] = bx[3 + k];
        ok = ok && finite_v(l[k]) && finite_v(h[k]) && (l[k] <= h[k]);
      }
    } else {
      const uint32_t* ix = (const uint32_t*)(g.indices + (uint64_t)prim * g.istride);
      const uint32_t i0 = ix[0], i1 = ix[1], i2 = ix[2];
      ok = (i0 < g.num_vertices) & (i1 < g.num_vertices) & (i2 < g.num_vertices);
      float v[9];
      if (ok) {
        const float* p0 = (const float*)(g.vertices + (uint64_t)i0 * g.vstride);
        const float* p1 = (const float*)(g.vertices + (uint64_t)i1 * g.vstride);
        const float* p2 = (const float*)(g.vertices + (uint64_t)i2 * g.vstride);
        v[0] = p0[0]; v[1] = p0[1]; v[2] = p0[2];
        v[3] = p1[0]; v[4] = p1[1]; v[5] = p1[2];
        v[6] = p2[0]; v[7] = p2[1]; v[8] = p2[2];
#pragma unroll
        for (int k = 0; k < 9; k++) ok &= finite_v(v[k]);
      }
#pragma unroll
      for (int k = 0; k < 3; k++) {
        l[k] = ok ? fminf(v[k], fminf(v[3 + k], v[6 + k])) : 0.0f;
        h[k] = ok ? fmaxf(v[k], fmaxf(v[3 + k], v[6 + k])) : 0.0f;
      }
    }
    if (ok) {
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const float c2 = l[k] + h[k];
        b[k] = fminf(b[k], l[k]);
        b[3 + k] = fmaxf(b[3 + k], h[k]);
        b[6 + k] = fminf(b[6 + k], c2);
        b[9 + k] = fmaxf(b[9 + k], c2);
        b[12] = fmaxf(b[12], fmaxf(fabsf(l[k]), fabsf(h[k])));
      }
      lo4[i] = make_float4(l[0], l[1], l[2], __uint_as_float(lo));
      hi4[i] = make_float4(h[0], h[1], h[2], __uint_as_float(prim));
      keys[i] = g.scene;
      smin = min(smin, g.scene);
      smax = max(smax, g.scene);
      valid++;
    } else {
      keys[i] = 0xffffffffu;
    }
  }
  // block reduction
  __shared__ float red[BT / 64][13];
  __shared__ uint32_t redc[BT / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 13; k++) {
    const bool mx = (k >= 3 && k < 6) || (k >= 9);
    float x = mx ? wave_max(b[k]) : wave_min(b[k]);
    if (lane == 0) red[wave][k] = x;
  }
  uint32_t vc = wave_sum(valid);
  if (lane == 0) redc[wave] = vc;
  __shared__ uint32_t s_min, s_max;
  if (threadIdx.x == 0) {
    s_min = 0xffffffffu;
    s_max = 0u;
  }
  __syncthreads();
  if (sacc && valid) {
    atomicMin(&s_min, smin);
    atomicMax(&s_max, smax);
  }
  if (threadIdx.x < 13) {
    const int k = threadIdx.x;
    const bool mx = (k >= 3 && k < 6) || (k >= 9);
    float x = red[0][k];
    for (int w = 1; w < BT / 64; w++) x = mx ? fmaxf(x, red[w][k]) : fminf(x, red[w][k]);
    uint32_t* dst = (k < 3) ? &cnt->glo[k] : (k < 6) ? &cnt->ghi[k - 3] : (k < 9) ? &cnt->clo[k - 6]
                    : (k < 12) ? &cnt->chi[k - 9] : &cnt->maxabs;
    if (!__builtin_isinf(x) && !(k == 12 && x == 0.0f)) {
      if (mx) atomicMax(dst, f2o(x)); else atomicMin(dst, f2o(x));
    }
  }
  if (threadIdx.x == 0) {
    uint32_t sum = 0;
    for (int w = 0; w < BT / 64; w++) sum += redc[w];
    if (sum) atomicAdd(&cnt->valid, sum);
  }
  if (!sacc) return;
  // per-scene bounds: the block's reduction when all its valid primitives
  // belong to one scene, else one set of atomics per primitive
  __syncthreads();
  if (s_min == 0xffffffffu) return;
  if (s_min == s_max) {
    SceneAcc* a = sacc + s_min;
    if (threadIdx.x < 12) {
      const int k = threadIdx.x;
      const bool mx = (k >= 3 && k < 6) || (k >= 9);
      float x = red[0][k];
      for (int w = 1; w < BT / 64; w++) x = mx ? fmaxf(x, red[w][k]) : fminf(x, red[w][k]);
      uint32_t* dst = (k < 3) ? &a->glo[k] : (k < 6) ? &a->ghi[k - 3] : (k < 9) ? &a->clo[k - 6] : &a->chi[k - 9];
      if (mx) atomicMax(dst, f2o(x)); else atomicMin(dst, f2o(x));
    }
    if (threadIdx.x == 0) {
      uint32_t sum = 0;
      for (int w = 0; w < BT / 64; w++) sum += redc[w];
      atomicAdd(&a->valid, sum);
    }
  } else {
    for (int j = 0; j < PR_PER; j++) {
      const uint32_t i = base + (uint32_t)j * BT;
      if (i >= total) break;
      const uint32_t sc = keys[i];
      if (sc == 0xffffffffu) continue;
      SceneAcc* a = sacc + sc;
      const float4 l = lo4[i], h = hi4[i];
      const float lv[3] = {l.x, l.y, l.z}, hv[3] = {h.x, h.y, h.z};
      for (int k = 0; k < 3; k++) {
        atomicMin(&a->glo[k], f2o(lv[k]));
        atomicMax(&a->ghi[k], f2o(hv[k]));
        atomicMin(&a->clo[k], f2o(lv[k] + hv[k]));
        atomicMax(&a->chi[k], f2o(lv[k] + hv[k]));
      }
      atomicAdd(&a->valid, 1u);
    }
  }
}

// ---------------------------------------------------------------- morton --

// bitInterleave (reference common/math/math.h:390-409)
__device__ inline uint32_t interleave3(uint32_t x) {
  x = (x | (x << 16)) & 0x030000FFu;
  x = (x | (x << 8)) & 0x0300F00Fu;
  x = (x | (x << 4)) & 0x030C30C3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

__global__ __launch_bounds__(BT) void k_morton(const float4* __restrict__ lo4, const float4* __restrict__ hi4,
                                               uint32_t total, uint32_t* __restrict__ keys,
                                               uint32_t* __restrict__ vals, const BuildCounters* __restrict__ cnt) {
  const uint32_t i = blockIdx.x * BT + threadIdx.x;
  if (i >= total) return;
  vals[i] = i;
  if (keys[i] == 0xffffffffu) return;
  const float4 l = lo4[i], h = hi4[i];
  const float c[3] = {l.x + h.x, l.y + h.y, l.z + h.z};
  uint32_t q[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float base = o2f(cnt->clo[k]);
    const float diag = o2f(cnt->chi[k]) - base;
    const float scale = diag > 1e-19f ? (1024.0f * 0.99f) / diag : 0.0f;
    float x = (c[k] - base) * scale;
    x = fminf(fmaxf(x, 0.0f), 1023.0f);
    q[k] = (uint32_t)x;
  }
  keys[i] = interleave3(q[0]) | (interleave3(q[1]) << 1) | (interleave3(q[2]) << 2);
}

__global__ __launch_bounds__(BT) void k_gather_sorted(const uint32_t* __restrict__ vals,
                                                      const float4* __restrict__ lo4,
                                                      const float4* __restrict__ hi4, uint32_t n,
                                                      float4* __restrict__ slo, float4* __restrict__ shi) {
  const uint32_t j = blockIdx.x * BT + threadIdx.x;
  if (j >= n) return;
  const uint32_t o = vals[j];
  slo[j] = lo4[o];
  shi[j] = hi4[o];
}

// ------------------------------------------------------- radix tree (LBVH) --

__device__ inline int delta(const uint32_t* __restrict__ keys, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t a = keys[i], b = keys[j];
  if (a != b) return __clz(a ^ b);
  return 32 + __clz((uint32_t)i ^ (uint32_t)j);
}

// Binary radix tree over sorted keys (Karras, HPG 2012): inner node i covers
// a contiguous range of sorted primitives and splits it where the highest
// differing key bit changes.
__global__ __launch_bounds__(BT) void k_radix_tree(const uint32_t* __restrict__ keys, int n,
                                                   uint32_t* __restrict__ child, uint32_t* __restrict__ parent_int,
                                                   uint32_t* __restrict__ parent_leaf, uint2* __restrict__ range) {
  const int i = blockIdx.x * BT + threadIdx.x;
  if (i >= n - 1) return;
  const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta(keys, n, i, i - d);
  int lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = delta(keys, n, i, j);
  int s = 0, t = l;
  while (true) {
    t = (t + 1) >> 1;
    if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    if (t <= 1) break;
  }
  const int gamma = i + s * d + (d < 0 ? -1 : 0);
  if (gamma < 0 || gamma + 1 >= n) return;  // impossible for sorted keys; keeps writes in bounds
  const int first = i < j ? i : j, last = i < j ? j : i;
  const uint32_t left = (first == gamma) ? ((uint32_t)gamma | REF_LEAF) : (uint32_t)gamma;
  const uint32_t right = (last == gamma + 1) ? ((uint32_t)(gamma + 1) | REF_LEAF) : (uint32_t)(gamma + 1);
  child[2 * i] = left;
  child[2 * i + 1] = right;
  if (left & REF_LEAF) parent_leaf[left & ~REF_LEAF] = i; else parent_int[left] = i;
  if (right & REF_LEAF) parent_leaf[right & ~REF_LEAF] = i; else parent_int[right] = i;
  range[i] = make_uint2((uint32_t)first, (uint32_t)last);
  if (i == 0) parent_int[0] = NONE;
}

__device__ inline float box_half_area(float4 l, float4 h) { return half_area(h.x - l.x, h.y - l.y, h.z - l.z); }


// Bottom-up bounds and the SAH-optimal 8-wide collapse table. The second
// child to arrive at a node (agent-scope acquire/release counter) finishes it
// and moves up. For every binary node n and slot budget s = 1..8, cost8 holds
// the least SAH cost of covering n's subtree with at most s subtrees:
//   c(n,1) = min(leaf: cost_tri*count*A(n), node: cost_node*A(n) + d(n,8))
//   d(n,s) = min_j c(left,j) + c(right,s-j)      (s >= 2)
//   c(n,s) = min(c(n,s-1), d(n,s))
// split8 records, per s, the best j (low 3 bits) and whether fewer slots win
// (bit 7). A binary leaf costs cost_tri*A(prim) whatever the budget.
// (Collapse formulation after Ylitie, Karras, Laine: "Efficient Incoherent
// Ray Traversal on GPUs Through Compressed Wide BVHs", HPG 2017.)
// A node's width is capped at `width` children (d(n, width) in the inner cost).
// Bottom-up step of one inner node: box of its two children, and the
// SAH-optimal way to spend 1..8 child slots below it (the 8-wide collapse
// DP): cost8[s-1] = least cost of covering the subtree with s slots,
// split8[s-1] = how many of them go left (0x80: fewer slots do as well),
// isleaf = a leaf beats an inner node. Returns false on a broken link.
struct RefitArgs {
  uint32_t n;
  const uint32_t* child;
  const uint2* range;
  const float4* slo;
  const float4* shi;
  float4* nlo;
  float4* nhi;
  float* cost8;
  uint8_t* split8;
  uint8_t* isleaf;
  float cost_node, cost_tri;
  uint32_t max_leaf, width, min_leaf, block;
  uint32_t* err;
};

__device__ inline bool refit_node(const RefitArgs& a, uint32_t node) {
  const uint32_t n = a.n;
  const uint32_t* child = a.child;
  const uint2* range = a.range;
  const float4* slo = a.slo;
  const float4* shi = a.shi;
  float4* nlo = a.nlo;
  float4* nhi = a.nhi;
  float* cost8 = a.cost8;
  uint8_t* split8 = a.split8;
  uint8_t* isleaf = a.isleaf;
  const float cost_node = a.cost_node, cost_tri = a.cost_tri;
  const uint32_t max_leaf = a.max_leaf, width = a.width, min_leaf = a.min_leaf, block = a.block;
  uint32_t* err = a.err;
  float4 bl[2], bh[2];
  float cc[2][8];
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const uint32_t c = child[2 * node + k];
    if ((c & REF_LEAF) ? (c & ~REF_LEAF) >= n : c + 1 >= n) {
      atomicOr(err, 1u);
      return false;
    }
    if (c & REF_LEAF) {
      bl[k] = slo[c & ~REF_LEAF];
      bh[k] = shi[c & ~REF_LEAF];
      const float ct = cost_tri * box_half_area(bl[k], bh[k]);
#pragma unroll
      for (int s = 0; s < 8; s++) cc[k][s] = ct;
    } else {
      bl[k] = nlo[c];
      bh[k] = nhi[c];
      const float4* t = reinterpret_cast<const float4*>(cost8 + 8 * (size_t)c);
      const float4 t0 = t[0], t1 = t[1];
      cc[k][0] = t0.x; cc[k][1] = t0.y; cc[k][2] = t0.z; cc[k][3] = t0.w;
      cc[k][4] = t1.x; cc[k][5] = t1.y; cc[k][6] = t1.z; cc[k][7] = t1.w;
    }
  }
  const float4 l = make_float4(fminf(bl[0].x, bl[1].x), fminf(bl[0].y, bl[1].y), fminf(bl[0].z, bl[1].z), 0.0f);
  const float4 h = make_float4(fmaxf(bh[0].x, bh[1].x), fmaxf(bh[0].y, bh[1].y), fmaxf(bh[0].z, bh[1].z), 0.0f);
  const uint2 r = range[node];
  const uint32_t count = r.y - r.x + 1;
  const float A = box_half_area(l, h);
  float d[9];
  uint32_t dj[9];
#pragma unroll
  for (int s = 2; s <= 8; s++) {
    float best = INFINITY;
    uint32_t bj = 1;
#pragma unroll
    for (int jj = 1; jj < s; jj++) {
      const float v = cc[0][jj - 1] + cc[1][s - jj - 1];
      if (v < best) { best = v; bj = (uint32_t)jj; }
    }
    d[s] = best;
    dj[s] = bj;
  }
  // a leaf pays per block of `block` primitives (the reference's
  // sahBlockSize); subtrees of at most min_leaf primitives are leaves
  const float leafc = count <= max_leaf ? cost_tri * (float)((count + block - 1) / block) * A : INFINITY;
  const float inner = count <= min_leaf ? INFINITY : cost_node * A + d[width];  // at most `width` children
  float c[8];
  uint32_t sp[8];
  c[0] = fminf(leafc, inner);
  sp[0] = 0;
#pragma unroll
  for (int s = 2; s <= 8; s++) {
    if (d[s] < c[s - 2]) { c[s - 1] = d[s]; sp[s - 1] = dj[s]; }
    else { c[s - 1] = c[s - 2]; sp[s - 1] = 0x80u | dj[s]; }
  }
  nlo[node] = l;
  nhi[node] = h;
  float4* t = reinterpret_cast<float4*>(cost8 + 8 * (size_t)node);
  t[0] = make_float4(c[0], c[1], c[2], c[3]);
  t[1] = make_float4(c[4], c[5], c[6], c[7]);
  uint2* sp8 = reinterpret_cast<uint2*>(split8 + 8 * (size_t)node);
  *sp8 = make_uint2(sp[0] | (sp[1] << 8) | (sp[2] << 16) | (sp[3] << 24),
                    sp[4] | (sp[5] << 8) | (sp[6] << 16) | (sp[7] << 24));
  isleaf[node] = leafc <= inner ? 1 : 0;
  return true;
}

// Refit of one depth level of a tree whose inner nodes know their depth:
// the deeper levels are complete (earlier launches), so no thread waits on
// another and no cross-XCD hand-off is needed.
__global__ __launch_bounds__(BT) void k_refit_level(RefitArgs a, const uint32_t* __restrict__ order, uint32_t count,
                                                    uint32_t* visit) {
  const uint32_t i = blockIdx.x * BT + threadIdx.x;
  if (i >= count) return;
  const uint32_t node = order[i];
  if (node + 1 >= a.n) {
    atomicOr(a.err, 1u);
    return;
  }
  if (refit_node(a, node)) visit[node] = 2u;  // what the tree check expects of a refitted node
}

// Depth of every inner node of a radix tree (its ancestors counted by
// walking the parent links); a count past 254 (only a broken link makes one)
// raises *deep and the build fails.
__global__ __launch_bounds__(BT) void k_node_depth(const uint32_t* __restrict__ parent_int, uint32_t m,
                                                   uint8_t* __restrict__ depth, uint32_t* deep) {
  const uint32_t i = blockIdx.x * BT + threadIdx.x;
  if (i >= m) return;
  uint32_t d = 0, p = parent_int[i];
  while (p != NONE && d < 255u) {
    if (p >= m) {  // a broken link: the tree check reports it
      d = 255u;
      break;
    }
    d++;
    p = parent_int[p];
  }
  if (d > 254u) atomicOr(deep, 1u);
  depth[i] = (uint8_t)(d > 254u ? 254u : d);
}

// Inner nodes grouped by depth: a histogram, then a scatter into the
// host-computed level offsets (order within a level does not matter). A
// block covers DEPTH_PER * BT nodes and adds to each global counter once:
// a level's nodes are many and their depths few, and one global atomic per
// (small block, depth) would serialise on those few addresses. Within a
// wave, one LDS atomic serves every lane of one depth.
constexpr int DEPTH_PER = 16;

__device__ inline uint32_t depth_rank(uint32_t* h, uint32_t d, uint32_t none) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t below = lane == 0u ? 0ull : (~0ull >> (64u - lane));
  uint32_t local = 0u;
  uint64_t todo = __ballot(d != none);
  while (todo) {
    const int leader = __ffsll((unsigned long long)todo) - 1;
    const uint32_t dl = (uint32_t)__shfl((int)d, leader, 64);
    const uint64_t same = __ballot(d == dl);
    uint32_t b = 0u;
    if ((int)lane == leader) b = atomicAdd(&h[dl], (uint32_t)__popcll(same));
    b = (uint32_t)__shfl((int)b, leader, 64);
    if (d == dl) local = b + (uint32_t)__popcll(same & below);
    todo &= ~same;
  }
  return local;
}

__global__ __launch_bounds__(BT) void k_depth_hist(const uint8_t* __restrict__ depth, uint32_t m,
                                                   uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0u;
  __syncthreads();
  const uint32_t base = blockIdx.x * (BT * DEPTH_PER) + threadIdx.x;
  for (int j = 0; j < DEPTH_PER; j++) {
    const uint32_t i = base + (uint32_t)j * BT;
    (void)depth_rank(h, i < m ? depth[i] : 256u, 256u);
  }
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

__global__ __launch_bounds__(BT) void k_depth_scatter(const uint8_t* __restrict__ depth, uint32_t m,
                                                      uint32_t* __restrict__ cursor, uint32_t* __restrict__ order) {
  __shared__ uint32_t h[256], base[256];
  h[threadIdx.x] = 0u;
  __syncthreads();
  const uint32_t first = blockIdx.x * (BT * DEPTH_PER) + threadIdx.x;
  uint32_t d[DEPTH_PER], local[DEPTH_PER];
#pragma unroll
  for (int j = 0; j < DEPTH_PER; j++) {
    const uint32_t i = first + (uint32_t)j * BT;
    d[j] = i < m ? depth[i] : 255u;  // 255: no node, or an id no tree uses
    local[j] = depth_rank(h, d[j], 255u);
  }
  __syncthreads();
  if (h[threadIdx.x]) base[threadIdx.x] = atomicAdd(&cursor[threadIdx.x], h[threadIdx.x]);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < DEPTH_PER; j++)
    if (d[j] != 255u) order[base[d[j]] + local[j]] = first + (uint32_t)j * BT;
}

// Integrity check of the sorted keys and the radix tree (cheap, one pass):
// bit 3 keys out of order, bit 4 inner node not finished by the refit,
// bit 5 parent/child links disagree.
__global__ __launch_bounds__(BT) void k_check_tree(const uint32_t* __restrict__ keys, uint32_t n, uint32_t* roots_seen,
                                                   const uint32_t* __restrict__ child,
                                                   const uint32_t* __restrict__ parent_int,
                                                   const uint32_t* __restrict__ parent_leaf,
                                                   const uint32_t* __restrict__ visit, uint32_t* err) {
  const uint32_t i = blockIdx.x * BT + threadIdx.x;
  if (keys && i + 1 < n && keys[i] > keys[i + 1]) atomicOr(err, 8u);
  if (i < n) {
    const uint32_t p = parent_leaf[i];
    if (p == NONE) atomicAdd(roots_seen, 1u);  // a scene of one primitive
    else if (p >= n - 1 || (child[2 * p] != (i | REF_LEAF) && child[2 * p + 1] != (i | REF_LEAF))) atomicOr(err, 32u);
  }
  if (i + 1 < n && parent_int[i] != REF_EMPTY) {
    if (visit[i] != 2u) atomicOr(err, 16u);
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const uint32_t c = child[2 * i + k];
      const bool ok = (c & REF_LEAF) ? ((c & ~REF_LEAF) < n && parent_leaf[c & ~REF_LEAF] == i)
                                     : (c + 1 < n && parent_int[c] == i);
      if (!ok) atomicOr(err, 32u);
    }
    if (parent_int[i] == NONE) {
      atomicAdd(roots_seen, 1u);
    } else {
      const uint32_t p = parent_int[i];
      if (p >= n - 1 || (child[2 * p] != i && child[2 * p + 1] != i)) atomicOr(err, 32u);
    }
  }
}

// ---------------------------------------------------------------- collapse --

struct CollapseArgs {
  const QItem* qin;
  QItem* qout;
  const uint32_t* nin;  // items in qin (a device counter)
  uint32_t qout_slot;
  uint32_t* qclear;     // the count the next level fills, cleared here (no kernel of this level uses it)
  const uint32_t* child;
  const uint2* range;
  const float4* nlo;
  const float4* nhi;
  const uint8_t* isleaf;
  const uint8_t* split8;
  const float4* slo;
  const float4* shi;
  QNode8* nodes;
  uint32_t* leaf_src; // position of each leaf record's primitive, bit 31: last of its leaf
  NodeInfo* info;
  BuildCounters* cnt;
  uint32_t n;         // capacity of nodes and of each queue
  uint32_t tri_cap;   // capacity of tris
  uint32_t width;     // children per node (2..8)
};

constexpr uint32_t LEAF_SRC_LAST = 0x80000000u;  // positions are below 2^31 (REF_LEAF marks them likewise)

__device__ inline bool is_inner(uint32_t ref, const uint8_t* isleaf) {
  return !(ref & REF_LEAF) && !isleaf[ref];
}

// Quantises one axis: returns the biased exponent byte of the grid step.
__device__ inline uint32_t grid_exponent(float p, float hi) {
  const float ext = hi - p;
  int eb = (int)((__float_as_uint(ext) >> 23) & 0xff) - 7;
  if (eb < 1) eb = 1;
  for (int it = 0; it < 64 && eb < 254; it++) {
    const float s = __uint_as_float((uint32_t)eb << 23);
    if (__builtin_fmaf(s, 254.0f, p) >= hi) break;
    eb++;
  }
  return (uint32_t)eb;
}

__device__ inline uint32_t quant_lo(float v, float p, float s) {
  float q = floorf((v - p) / s);
  q = fminf(fmaxf(q, 0.0f), 255.0f);
  uint32_t qi = (uint32_t)q;
  while (qi > 0 && __builtin_fmaf(s, (float)qi, p) > v) qi--;
  return qi;
}
__device__ inline uint32_t quant_hi(float v, float p, float s) {
  float q = ceilf((v - p) / s);
  q = fminf(fmaxf(q, 0.0f), 255.0f);
  uint32_t qi = (uint32_t)q;
  while (qi < 255 && __builtin_fmaf(s, (float)qi, p) < v) qi++;
  return qi;
}

// Byte k of a node's eight split bytes loaded as one uint2.
__device__ inline uint32_t split_byte(uint2 s8, uint32_t k) { return ((k < 4 ? s8.x : s8.y) >> (8 * (k & 3))) & 0xffu; }

// Statistics of the nodes a thread wrote, reduced over the wave before
// the global atomics (one per item would serialise on their addresses).
struct CollapseStats {
  uint32_t max_depth = 0, max_stack = 0;
  double sah_nodes = 0.0, sah_leaves = 0.0, sah_leaves_tri = 0.0;
};

// Node, record and queue slots for the workgroup's items in one atomic per
// counter (atomics on one address from every wave would serialise): the
// bit-sliced ballot sums give each lane its offset within the wave, LDS
// the waves' offsets within the workgroup. Called by every thread of the
// workgroup; ni (children that are inner nodes) is below 16 and nt (leaf
// records) below 512.
__device__ inline void block_alloc(BuildCounters* cnt, uint32_t qslot, uint32_t ni, uint32_t nt, uint32_t& child_base,
                                   uint32_t& tri_base, uint32_t& qbase) {
  __shared__ uint32_t wtot[BT / 64][2], bbase[3];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint64_t below = lane == 0u ? 0ull : (~0ull >> (64u - lane));
  uint32_t pi = 0u, ti = 0u, pt = 0u, tt = 0u;
#pragma unroll
  for (int b = 0; b < 4; b++) {
    const uint64_t m = __ballot((ni >> b) & 1u);
    pi += (uint32_t)__popcll(m & below) << b;
    ti += (uint32_t)__popcll(m) << b;
  }
#pragma unroll
  for (int b = 0; b < 9; b++) {
    const uint64_t m = __ballot((nt >> b) & 1u);
    pt += (uint32_t)__popcll(m & below) << b;
    tt += (uint32_t)__popcll(m) << b;
  }
  if (lane == 0u) {
    wtot[wave][0] = ti;
    wtot[wave][1] = tt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t si = 0u, st = 0u;
    for (int w = 0; w < BT / 64; w++) {
      si += wtot[w][0];
      st += wtot[w][1];
    }
    bbase[0] = si ? atomicAdd(&cnt->node_count, si) : 0u;
    bbase[1] = st ? atomicAdd(&cnt->tri_count, st) : 0u;
    bbase[2] = si ? atomicAdd(&cnt->qcount[qslot], si) : 0u;
  }
  __syncthreads();
  uint32_t oi = 0u, ot = 0u;
  for (uint32_t w = 0; w < wave; w++) {
    oi += wtot[w][0];
    ot += wtot[w][1];
  }
  child_base = bbase[0] + oi + pi;
  tri_base = bbase[1] + ot + pt;
  qbase = bbase[2] + oi + pi;
  __syncthreads();  // wtot and bbase are reused by the next item
}

// One item (an 8-wide node to write) per thread, for the levels with many
// items (throughput); every thread of the workgroup calls it (active: the
// thread has an item), for block_alloc.
__device__ inline void collapse_item1(const CollapseArgs& a, uint32_t t, bool active, CollapseStats& st) {
  bool go = active;
  QItem item{};
  if (go) item = a.qin[t];
  const float eps = o2f(a.cnt->maxabs) * BOX_EPS_REL;
  if (go && item.bref == REF_EMPTY) {  // root of a scene without primitives: a node with no children
    go = false;
    if (item.node >= a.n) {
      atomicOr(&a.cnt->error, 2u);
    } else {
    QNode8 q;
    q.px = q.py = q.pz = 0.0f;
    q.ex = q.ey = q.ez = 127;
    q.imask = 0;
    q.child_base = q.tri_base = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      q.meta[k] = 0;
      q.qlox[k] = q.qloy[k] = q.qloz[k] = 255;
      q.qhix[k] = q.qhiy[k] = q.qhiz[k] = 0;
    }
    a.nodes[item.node] = q;
    NodeInfo inf;
    inf.lo[0] = inf.lo[1] = inf.lo[2] = 0.0f;
    inf.hi[0] = inf.hi[1] = inf.hi[2] = 0.0f;
    inf.count = 0;
    inf.depth = item.depth;
    a.info[item.node] = inf;
    }
  }

  // children: the optimal distribution of 8 slots below item.bref
  uint32_t ch[8];
  int n = 0;
  if (!go) {
  } else if (!is_inner(item.bref, a.isleaf)) {
    ch[0] = item.bref;  // whole (small) scene as one leaf
    n = 1;
  } else {
    uint32_t st_ref[8], st_s[8];
    int sp = 0;
    {
      const uint32_t j8 = (a.split8[8 * (size_t)item.bref + a.width - 1]) & 7u;
      st_ref[0] = a.child[2 * item.bref + 1]; st_s[0] = a.width - j8;
      st_ref[1] = a.child[2 * item.bref];     st_s[1] = j8;
      sp = 2;
    }
    while (sp > 0 && n < 8) {
      sp--;
      uint32_t c = st_ref[sp], slots = st_s[sp];
      uint2 s8 = make_uint2(0u, 0u), cp = make_uint2(0u, 0u);
      if (!(c & REF_LEAF)) {
        // the node's eight split bytes and its child pair: two 8 B loads
        // issued together, not a byte load per slot count and then the pair
        s8 = *reinterpret_cast<const uint2*>(a.split8 + 8 * (size_t)c);
        cp = *reinterpret_cast<const uint2*>(a.child + 2 * (size_t)c);
        while (slots > 1 && (split_byte(s8, slots - 1) & 0x80u)) slots--;
      } else {
        slots = 1;
      }
      if (slots <= 1) {
        ch[n++] = c;
      } else {
        const uint32_t jj = split_byte(s8, slots - 1) & 7u;
        if (sp + 2 > 8) {
          atomicOr(&a.cnt->error, 128u);
          go = false;
          n = 0;
          break;
        }
        st_ref[sp] = cp.y; st_s[sp] = slots - jj; sp++;
        st_ref[sp] = cp.x; st_s[sp] = jj;         sp++;
      }
    }
  }

  // per-child primitive ranges live in LDS (one row per thread)
  __shared__ uint32_t s_first[BT][8];
  __shared__ uint32_t s_count[BT][8];
  const uint32_t* rng32 = reinterpret_cast<const uint32_t*>(a.range);
  float bl[8][3], bh[8][3];
  uint32_t ni = 0, nt = 0, ninner_prims = 0;
  uint32_t innerm = 0;  // bit k: child k is an inner node (kept for the writes below)
  float ul[3] = {INFINITY, INFINITY, INFINITY}, uh[3] = {-INFINITY, -INFINITY, -INFINITY};
  double sah_leaf = 0.0, sah_leaf_tri = 0.0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    if (k < n) {
      const uint32_t c = ch[k];
      float4 l, h;
      uint32_t f, cnt;
      if (c & REF_LEAF) {
        l = a.slo[c & ~REF_LEAF];
        h = a.shi[c & ~REF_LEAF];
        f = c & ~REF_LEAF;
        cnt = 1u;
      } else {
        l = a.nlo[c];
        h = a.nhi[c];
        f = rng32[2 * c];
        cnt = rng32[2 * c + 1] - f + 1u;
      }
      s_first[threadIdx.x][k] = f;
      s_count[threadIdx.x][k] = cnt;
      bl[k][0] = l.x; bl[k][1] = l.y; bl[k][2] = l.z;
      bh[k][0] = h.x; bh[k][1] = h.y; bh[k][2] = h.z;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        ul[d] = fminf(ul[d], bl[k][d]);
        uh[d] = fmaxf(uh[d], bh[k][d]);
      }
      if (is_inner(c, a.isleaf)) {
        innerm |= 1u << k;
        ni++;
        ninner_prims += cnt;
      } else {
        nt += cnt;
        const float A = box_half_area(l, h);
        sah_leaf += (double)A * (double)((cnt + 3) / 4);
        sah_leaf_tri += (double)A * (double)cnt;
      }
    }
  }
  // the children partition the parent's range: a second count of the leaf primitives
  if (go) {
    const uint32_t b = item.bref;
    const uint32_t total_prims = (b & REF_LEAF) ? 1u : (rng32[2 * b + 1] - rng32[2 * b] + 1u);
    if (total_prims - ninner_prims != nt) atomicOr(&a.cnt->error, 64u);
  }
  // node, record and queue slots: one atomic per workgroup and counter
  uint32_t child_base, tri_base, qbase;
  block_alloc(a.cnt, a.qout_slot, go ? ni : 0u, go ? nt : 0u, child_base, tri_base, qbase);
  if (!go) return;
  if ((uint64_t)child_base + ni > a.n || (uint64_t)tri_base + nt > a.tri_cap || (uint64_t)qbase + ni > a.n ||
      item.node >= a.n) {
    atomicOr(&a.cnt->error, 2u);
    return;
  }
  st.max_depth = max(st.max_depth, item.depth);
  // a node pushes at most n-1 of its n children
  const uint32_t need = item.need + (uint32_t)n - 1u;
  st.max_stack = max(st.max_stack, need);
  st.sah_nodes += (double)half_area(uh[0] - ul[0], uh[1] - ul[1], uh[2] - ul[2]);
  if (nt) {
    st.sah_leaves += sah_leaf;
    st.sah_leaves_tri += sah_leaf_tri;
  }

  // quantisation grid over the widened union box
  QNode8 q;
  float p[3], s[3];
  uint32_t e[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    p[d] = ul[d] - eps;
    e[d] = grid_exponent(p[d], uh[d] + eps);
    s[d] = __uint_as_float(e[d] << 23);
  }
  q.px = p[0]; q.py = p[1]; q.pz = p[2];
  q.ex = (uint8_t)e[0]; q.ey = (uint8_t)e[1]; q.ez = (uint8_t)e[2];
  q.child_base = child_base;
  q.tri_base = tri_base;
  uint32_t imask = 0, rank_i = 0, off_t = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    if (k < n) {
      const uint32_t c = ch[k];
      const uint32_t qlx = quant_lo(bl[k][0] - eps, p[0], s[0]), qhx = quant_hi(bh[k][0] + eps, p[0], s[0]);
      const uint32_t qly = quant_lo(bl[k][1] - eps, p[1], s[1]), qhy = quant_hi(bh[k][1] + eps, p[1], s[1]);
      const uint32_t qlz = quant_lo(bl[k][2] - eps, p[2], s[2]), qhz = quant_hi(bh[k][2] + eps, p[2], s[2]);
      q.qlox[k] = (uint8_t)qlx; q.qhix[k] = (uint8_t)qhx;
      q.qloy[k] = (uint8_t)qly; q.qhiy[k] = (uint8_t)qhy;
      q.qloz[k] = (uint8_t)qlz; q.qhiz[k] = (uint8_t)qhz;
      if ((innerm >> k) & 1u) {
        imask |= 1u << k;
        q.meta[k] = (uint8_t)(0x80u | rank_i);
        a.qout[qbase + rank_i] = QItem{c, child_base + rank_i, item.depth + 1, need};
        rank_i++;
      } else {
        q.meta[k] = (uint8_t)(1u + off_t);
        // leaf records are written afterwards, one thread each (k_write_leaves)
        const uint32_t cntk = s_count[threadIdx.x][k];
        const uint32_t firstk = s_first[threadIdx.x][k];
        for (uint32_t j = 0; j < cntk; j++)
          a.leaf_src[tri_base + off_t + j] = (firstk + j) | ((j + 1 == cntk) ? LEAF_SRC_LAST : 0u);
        off_t += cntk;
      }
    } else {
      q.meta[k] = 0;
      q.qlox[k] = q.qloy[k] = q.qloz[k] = 255;
      q.qhix[k] = q.qhiy[k] = q.qhiz[k] = 0;
    }
  }
  q.imask = (uint8_t)imask;
  a.nodes[item.node] = q;
  NodeInfo inf;
  inf.lo[0] = ul[0]; inf.lo[1] = ul[1]; inf.lo[2] = ul[2];
  inf.hi[0] = uh[0]; inf.hi[1] = uh[1]; inf.hi[2] = uh[2];
  inf.count = ninner_prims + nt;
  inf.depth = item.depth;
  a.info[item.node] = inf;
}

// One item (an 8-wide node to write) per group of CG = 8 lanes, for the
// levels with few items (latency), lane k of the group taking slot k of the
// node: the group descends the binary tree
// together (every lane the same way) to the node's children, lane k keeps
// child k, and the lanes then load, count, quantise and write their own
// child's slot, with the node's union box, counts and slot ranks formed by
// reductions and scans over the group. Every thread of the workgroup calls
// it (active: the group has an item), for block_alloc.
constexpr uint32_t CG = 8;
constexpr uint32_t COLLAPSE_GROUPED_MAX = 8192;  // items of a level up to which groups take them

template <class T>
__device__ inline T group_sum(T v) {
#pragma unroll
  for (int d = 1; d < (int)CG; d <<= 1) v += __shfl_xor(v, d, CG);
  return v;
}
__device__ inline float group_min(float v) {
#pragma unroll
  for (int d = 1; d < (int)CG; d <<= 1) v = fminf(v, __shfl_xor(v, d, CG));
  return v;
}
__device__ inline float group_max(float v) {
#pragma unroll
  for (int d = 1; d < (int)CG; d <<= 1) v = fmaxf(v, __shfl_xor(v, d, CG));
  return v;
}
// exclusive prefix over the group's lanes
__device__ inline uint32_t group_prefix(uint32_t v, uint32_t k) {
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < (int)CG; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, CG);
    if (k >= (uint32_t)d) x += y;
  }
  return x - v;
}

__device__ inline void collapse_item(const CollapseArgs& a, uint32_t t, bool active, uint32_t k, CollapseStats& st) {
  bool go = active;
  QItem item{};
  if (go) item = a.qin[t];
  const float eps = o2f(a.cnt->maxabs) * BOX_EPS_REL;
  if (go && item.bref == REF_EMPTY) {  // root of a scene without primitives: a node with no children
    go = false;
    if (k == 0) {
      if (item.node >= a.n) {
        atomicOr(&a.cnt->error, 2u);
      } else {
        QNode8 q;
        q.px = q.py = q.pz = 0.0f;
        q.ex = q.ey = q.ez = 127;
        q.imask = 0;
        q.child_base = q.tri_base = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
          q.meta[j] = 0;
          q.qlox[j] = q.qloy[j] = q.qloz[j] = 255;
          q.qhix[j] = q.qhiy[j] = q.qhiz[j] = 0;
        }
        a.nodes[item.node] = q;
        NodeInfo inf;
        inf.lo[0] = inf.lo[1] = inf.lo[2] = 0.0f;
        inf.hi[0] = inf.hi[1] = inf.hi[2] = 0.0f;
        inf.count = 0;
        inf.depth = item.depth;
        a.info[item.node] = inf;
      }
    }
  }

  // children: the optimal distribution of 8 slots below item.bref (the
  // group's lanes take the same path); lane k keeps child k
  uint32_t mine = 0u;
  int n = 0;
  if (!go) {
  } else if (!is_inner(item.bref, a.isleaf)) {
    mine = item.bref;  // whole (small) scene as one leaf
    n = 1;
  } else {
    uint32_t st_ref[8], st_s[8];
    int sp = 0;
    {
      const uint32_t j8 = (a.split8[8 * (size_t)item.bref + a.width - 1]) & 7u;
      st_ref[0] = a.child[2 * item.bref + 1]; st_s[0] = a.width - j8;
      st_ref[1] = a.child[2 * item.bref];     st_s[1] = j8;
      sp = 2;
    }
    while (sp > 0 && n < 8) {
      sp--;
      uint32_t c = st_ref[sp], slots = st_s[sp];
      uint2 s8 = make_uint2(0u, 0u), cp = make_uint2(0u, 0u);
      if (!(c & REF_LEAF)) {
        // the node's eight split bytes and its child pair: two 8 B loads
        // issued together, not a byte load per slot count and then the pair
        s8 = *reinterpret_cast<const uint2*>(a.split8 + 8 * (size_t)c);
        cp = *reinterpret_cast<const uint2*>(a.child + 2 * (size_t)c);
        while (slots > 1 && (split_byte(s8, slots - 1) & 0x80u)) slots--;
      } else {
        slots = 1;
      }
      if (slots <= 1) {
        if ((uint32_t)n == k) mine = c;
        n++;
      } else {
        const uint32_t jj = split_byte(s8, slots - 1) & 7u;
        if (sp + 2 > 8) {
          if (k == 0) atomicOr(&a.cnt->error, 128u);
          go = false;
          n = 0;
          break;
        }
        st_ref[sp] = cp.y; st_s[sp] = slots - jj; sp++;
        st_ref[sp] = cp.x; st_s[sp] = jj;         sp++;
      }
    }
  }

  // this lane's child: box, primitive range and kind
  const bool has = go && k < (uint32_t)n;
  const uint32_t* rng32 = reinterpret_cast<const uint32_t*>(a.range);
  float4 l = make_float4(INFINITY, INFINITY, INFINITY, 0.0f), h = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
  uint32_t f = 0u, cnt = 0u;
  bool inner = false;
  if (has) {
    if (mine & REF_LEAF) {
      l = a.slo[mine & ~REF_LEAF];
      h = a.shi[mine & ~REF_LEAF];
      f = mine & ~REF_LEAF;
      cnt = 1u;
    } else {
      l = a.nlo[mine];
      h = a.nhi[mine];
      f = rng32[2 * mine];
      cnt = rng32[2 * mine + 1] - f + 1u;
    }
    inner = is_inner(mine, a.isleaf);
  }
  const float ul[3] = {group_min(l.x), group_min(l.y), group_min(l.z)};
  const float uh[3] = {group_max(h.x), group_max(h.y), group_max(h.z)};
  const uint32_t ci = inner ? 1u : 0u, ct = (has && !inner) ? cnt : 0u;
  const uint32_t ni = group_sum(ci), nt = group_sum(ct);
  const uint32_t ninner_prims = group_sum(inner ? cnt : 0u);
  const uint32_t rank_i = group_prefix(ci, k), off_t = group_prefix(ct, k);
  double sah_leaf = 0.0, sah_leaf_tri = 0.0;
  if (has && !inner) {
    const float A = box_half_area(l, h);
    sah_leaf = (double)A * (double)((cnt + 3) / 4);
    sah_leaf_tri = (double)A * (double)cnt;
  }
  sah_leaf = group_sum(sah_leaf);
  sah_leaf_tri = group_sum(sah_leaf_tri);
  // the children partition the parent's range: a second count of the leaf primitives
  if (go && k == 0) {
    const uint32_t b = item.bref;
    const uint32_t total_prims = (b & REF_LEAF) ? 1u : (rng32[2 * b + 1] - rng32[2 * b] + 1u);
    if (total_prims - ninner_prims != nt) atomicOr(&a.cnt->error, 64u);
  }
  // node, record and queue slots: one atomic per workgroup and counter (the
  // group's first lane asks for the group)
  uint32_t child_base, tri_base, qbase;
  block_alloc(a.cnt, a.qout_slot, (go && k == 0) ? ni : 0u, (go && k == 0) ? nt : 0u, child_base, tri_base, qbase);
  const int first_lane = (int)((threadIdx.x & 63u) & ~(CG - 1u));
  child_base = (uint32_t)__shfl((int)child_base, first_lane, 64);
  tri_base = (uint32_t)__shfl((int)tri_base, first_lane, 64);
  qbase = (uint32_t)__shfl((int)qbase, first_lane, 64);
  if (!go) return;
  if ((uint64_t)child_base + ni > a.n || (uint64_t)tri_base + nt > a.tri_cap || (uint64_t)qbase + ni > a.n ||
      item.node >= a.n) {
    if (k == 0) atomicOr(&a.cnt->error, 2u);
    return;
  }
  // a node pushes at most n-1 of its n children
  const uint32_t need = item.need + (uint32_t)n - 1u;
  if (k == 0) {
    st.max_depth = max(st.max_depth, item.depth);
    st.max_stack = max(st.max_stack, need);
    st.sah_nodes += (double)half_area(uh[0] - ul[0], uh[1] - ul[1], uh[2] - ul[2]);
    if (nt) {
      st.sah_leaves += sah_leaf;
      st.sah_leaves_tri += sah_leaf_tri;
    }
  }

  // quantisation grid over the widened union box (every lane the same)
  float p[3], sc[3];
  uint32_t e[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    p[d] = ul[d] - eps;
    e[d] = grid_exponent(p[d], uh[d] + eps);
    sc[d] = __uint_as_float(e[d] << 23);
  }
  uint8_t* node = reinterpret_cast<uint8_t*>(a.nodes + item.node);
  uint32_t meta = 0u, qlx = 255u, qhx = 0u, qly = 255u, qhy = 0u, qlz = 255u, qhz = 0u;
  if (has) {
    qlx = quant_lo(l.x - eps, p[0], sc[0]);
    qhx = quant_hi(h.x + eps, p[0], sc[0]);
    qly = quant_lo(l.y - eps, p[1], sc[1]);
    qhy = quant_hi(h.y + eps, p[1], sc[1]);
    qlz = quant_lo(l.z - eps, p[2], sc[2]);
    qhz = quant_hi(h.z + eps, p[2], sc[2]);
    if (inner) {
      meta = 0x80u | rank_i;
      a.qout[qbase + rank_i] = QItem{mine, child_base + rank_i, item.depth + 1, need};
    } else {
      meta = 1u + off_t;
      // leaf records are written afterwards, one thread each (k_write_leaves)
      for (uint32_t j = 0; j < cnt; j++)
        a.leaf_src[tri_base + off_t + j] = (f + j) | ((j + 1 == cnt) ? LEAF_SRC_LAST : 0u);
    }
  }
  node[offsetof(QNode8, meta) + k] = (uint8_t)meta;
  node[offsetof(QNode8, qlox) + k] = (uint8_t)qlx;
  node[offsetof(QNode8, qhix) + k] = (uint8_t)qhx;
  node[offsetof(QNode8, qloy) + k] = (uint8_t)qly;
  node[offsetof(QNode8, qhiy) + k] = (uint8_t)qhy;
  node[offsetof(QNode8, qloz) + k] = (uint8_t)qlz;
  node[offsetof(QNode8, qhiz) + k] = (uint8_t)qhz;
  const uint32_t imask = (uint32_t)group_sum(inner ? (1u << k) : 0u);
  if (k == 0) {
    uint32_t* hdr = reinterpret_cast<uint32_t*>(node);
    hdr[0] = __float_as_uint(p[0]);
    hdr[1] = __float_as_uint(p[1]);
    hdr[2] = __float_as_uint(p[2]);
    hdr[3] = e[0] | (e[1] << 8) | (e[2] << 16) | (imask << 24);
    hdr[4] = child_base;
    hdr[5] = tri_base;
    NodeInfo inf;
    inf.lo[0] = ul[0]; inf.lo[1] = ul[1]; inf.lo[2] = ul[2];
    inf.hi[0] = uh[0]; inf.hi[1] = uh[1]; inf.hi[2] = uh[2];
    inf.count = ninner_prims + nt;
    inf.depth = item.depth;
    a.info[item.node] = inf;
  }
}

// Leaf record r: the primitive at position leaf_src[r] (primref vals[pos],
// whose .w fields hold its geometry and primitive index) as a TriRec
// (v0, v0 - v1, v2 - v0; the three vertices for the robust test) or as the
// instance's InstRec, with TRI_LAST on the last record of each leaf.
__global__ __launch_bounds__(BT) void k_write_leaves(const uint32_t* __restrict__ leaf_src,
                                                     const uint32_t* __restrict__ vals, const float4* __restrict__ lo4,
                                                     const float4* __restrict__ hi4, const GeomDesc* __restrict__ geoms,
                                                     const BuildCounters* cnt, TriRec* __restrict__ tris,
                                                     InstRec* __restrict__ insts, uint32_t robust) {
  const uint32_t r = blockIdx.x * BT + threadIdx.x;
  if (cnt->error || r >= cnt->tri_count) return;
  const uint32_t e = leaf_src[r];
  const uint32_t o = vals[e & ~LEAF_SRC_LAST];
  const uint32_t last = (e & LEAF_SRC_LAST) ? TRI_LAST : 0u;
  const uint32_t gi = __float_as_uint(lo4[o].w);
  const uint32_t prim = __float_as_uint(hi4[o].w);
  const GeomDesc g = geoms[gi];
  if (g.kind == GEOM_INSTANCES) {
    InstRec ir = *(const InstRec*)(g.indices + (uint64_t)prim * g.istride);
    ir.flags = (ir.flags & ~TRI_LAST) | last;
    insts[r] = ir;
    return;
  }
  const uint32_t* ix = (const uint32_t*)(g.indices + (uint64_t)prim * g.istride);
  const float* p0 = (const float*)(g.vertices + (uint64_t)ix[0] * g.vstride);
  const float* p1 = (const float*)(g.vertices + (uint64_t)ix[1] * g.vstride);
  const float* p2 = (const float*)(g.vertices + (uint64_t)ix[2] * g.vstride);
  TriRec tr;
  tr.v0x = p0[0]; tr.v0y = p0[1]; tr.v0z = p0[2];
  if (robust) {
    tr.e1x = p1[0]; tr.e1y = p1[1]; tr.e1z = p1[2];
    tr.e2x = p2[0]; tr.e2y = p2[1]; tr.e2z = p2[2];
  } else {
    tr.e1x = p0[0] - p1[0]; tr.e1y = p0[1] - p1[1]; tr.e1z = p0[2] - p1[2];
    tr.e2x = p2[0] - p0[0]; tr.e2y = p2[1] - p0[1]; tr.e2z = p2[2] - p0[2];
  }
  tr.primID = prim;
  tr.geomID = g.geomID;
  tr.flags = last;
  tris[r] = tr;
}

// One level of the collapse: the grid strides over the level's queue, whose
// length it reads on the device. A tree that failed its check (error bits
// set) is not followed.
__global__ __launch_bounds__(BT, 3) void k_collapse(CollapseArgs a) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *a.qclear = 0u;
  // the workgroup decides together (block_alloc synchronises it)
  __shared__ uint32_t s_run;
  if (threadIdx.x == 0) s_run = a.cnt->error ? 0u : 1u;
  __syncthreads();
  const uint32_t nin = *a.nin;
  // few items: a group of CG lanes per item, which shortens each item's
  // chain of work; many: a thread per item, fewer workgroup steps (each
  // step's block_alloc atomics serialise on their addresses)
  const bool grouped = nin <= COLLAPSE_GROUPED_MAX;
  if (s_run && blockIdx.x * (grouped ? BT / CG : BT) < nin) {
    CollapseStats st;
    // block-uniform loops: every thread takes part in each item's block_alloc
    if (grouped) {
      constexpr uint32_t PER_BLOCK = BT / CG;
      for (uint32_t base = blockIdx.x * PER_BLOCK; base < nin; base += gridDim.x * PER_BLOCK) {
        const uint32_t t = base + threadIdx.x / CG;
        collapse_item(a, t, t < nin, threadIdx.x % CG, st);
      }
    } else {
      for (uint32_t base = blockIdx.x * BT; base < nin; base += gridDim.x * BT)
        collapse_item1(a, base + threadIdx.x, base + threadIdx.x < nin, st);
    }
    const uint32_t md = wave_umax(st.max_depth), ms = wave_umax(st.max_stack);
    const double sn = wave_sum(st.sah_nodes), sl = wave_sum(st.sah_leaves), slt = wave_sum(st.sah_leaves_tri);
    if ((threadIdx.x & 63u) == 0u) {
      if (md) atomicMax(&a.cnt->max_depth, md);
      if (ms) atomicMax(&a.cnt->max_stack, ms);
      if (sn != 0.0) atomicAdd(&a.cnt->sah_nodes, sn);
      if (sl != 0.0) atomicAdd(&a.cnt->sah_leaves, sl);
      if (slt != 0.0) atomicAdd(&a.cnt->sah_leaves_tri, slt);
    }
  }
}


// ------------------------------------------------------------------ host --

void build_bvh(const std::vector<GeomDesc>& geoms, const BuildParams& prm, BuildScratch& S, DeviceBVH& out,
               hipStream_t s, uint32_t nscenes, UserTree* user) {
  auto t0 = std::chrono::steady_clock::now();
  out.generation++;
  out.stats = BuildStats();
  out.robust = prm.robust ? 1u : 0u;
  uint64_t total64 = 0;
  for (const auto& g : geoms) total64 += g.num_tris;
  // positions and primitive references carry a flag in bit 31
  if (total64 >= REF_LEAF) throw hip_error(hipErrorInvalidValue, "a build holds fewer than 2^31 primitives");
  const uint32_t total = (uint32_t)total64;
  out.slots.assign(nscenes, SceneSlot());
  if (user) *user = UserTree();
  if (geoms.empty() || total == 0) {
    out.stats.num_prims = 0;
    out.stats.num_nodes = 0;
    return;
  }
  if (user && nscenes != 1) throw hip_error(hipErrorInvalidValue, "a user tree is one scene");
  // a QNode8 leaf slot addresses at most 15 records; a user tree's leaves
  // are handed to the caller and may hold up to 32
  const uint32_t leaf_cap = user ? 32u : 15u;
  const uint32_t max_leaf = prm.max_leaf < 1 ? 1 : (prm.max_leaf > leaf_cap ? leaf_cap : prm.max_leaf);
  const uint32_t width = prm.width < 2 ? 2 : (prm.width > 8 ? 8 : prm.width);
  const uint32_t min_leaf = prm.min_leaf < 1 ? 1 : (prm.min_leaf > max_leaf ? max_leaf : prm.min_leaf);
  const uint32_t block = prm.block < 1 ? 1 : prm.block;

  S.geoms.reserve(geoms.size() * sizeof(GeomDesc));
  RTG_HIP(hipMemcpyAsync(S.geoms.ptr, geoms.data(), geoms.size() * sizeof(GeomDesc), hipMemcpyHostToDevice, s));
  S.counters.reserve(sizeof(BuildCounters));
  S.lo4.reserve((size_t)total * 16);
  S.hi4.reserve((size_t)total * 16);
  S.keys.reserve((size_t)total * 4);
  S.vals.reserve((size_t)total * 4);
  BuildCounters* cnt = S.counters.as<BuildCounters>();

  if (nscenes == 0 || nscenes > 65535) throw hip_error(hipErrorInvalidValue, "scene count out of range");
  const bool batched = nscenes > 1;
  SceneAcc* sacc = nullptr;
  if (batched) {
    S.sacc.reserve((size_t)nscenes * sizeof(SceneAcc));
    sacc = S.sacc.as<SceneAcc>();
    hipLaunchKernelGGL(k_init_sacc, dim3(div_up(nscenes, BT)), dim3(BT), 0, s, sacc, nscenes);
  }
  hipLaunchKernelGGL(k_init_counters, dim3(1), dim3(1), 0, s, cnt, nscenes);
  hipLaunchKernelGGL(k_primrefs, dim3(div_up(total, (size_t)BT * PR_PER)), dim3(BT), 0, s, S.geoms.as<GeomDesc>(),
                     (uint32_t)geoms.size(), total, S.lo4.as<float4>(), S.hi4.as<float4>(), S.keys.as<uint32_t>(), cnt,
                     sacc);
  if (batched) {
    // group the references by scene (stable, so each scene keeps primref order); invalid ones sort last
    hipLaunchKernelGGL(k_iota, dim3(div_up(total, BT)), dim3(BT), 0, s, S.vals.as<uint32_t>(), total);
    RTG_HIP(hipGetLastError());
    S.sorter.sort_pairs(S.keys.as<uint32_t>(), S.vals.as<uint32_t>(), total, 0, 16, s);
  } else {
    hipLaunchKernelGGL(k_morton, dim3(div_up(total, BT)), dim3(BT), 0, s, S.lo4.as<float4>(), S.hi4.as<float4>(),
                       total, S.keys.as<uint32_t>(), S.vals.as<uint32_t>(), cnt);
    RTG_HIP(hipGetLastError());
    S.sorter.sort_pairs(S.keys.as<uint32_t>(), S.vals.as<uint32_t>(), total, 0, 32, s);
  }

  BuildCounters hc;
  RTG_HIP(hipMemcpyAsync(&hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, s));
  std::vector<SceneAcc> hs(batched ? nscenes : 0);
  if (batched) RTG_HIP(hipMemcpyAsync(hs.data(), sacc, nscenes * sizeof(SceneAcc), hipMemcpyDeviceToHost, s));
  RTG_HIP(hipStreamSynchronize(s));
  uint32_t n = hc.valid;
  out.stats.num_prims = n;
  for (int k = 0; k < 3; k++) {
    out.stats.bounds_lo[k] = o2f(hc.glo[k]);
    out.stats.bounds_hi[k] = o2f(hc.ghi[k]);
  }
  // scenes: position ranges, bounds, and the binary-tree reference of each root
  out.slots.assign(nscenes, SceneSlot());
  std::vector<uint32_t> first(nscenes, 0), count(nscenes, 0);
  std::vector<SahRoot> trees;
  std::vector<uint32_t> tree_scene;
  {
    uint32_t pos = 0;
    for (uint32_t sI = 0; sI < nscenes; sI++) {
      SceneSlot& sl = out.slots[sI];
      sl.root = sI;
      count[sI] = batched ? hs[sI].valid : n;
      first[sI] = pos;
      pos += count[sI];
      sl.num_prims = count[sI];
      for (int k = 0; k < 3; k++) {
        sl.bounds_lo[k] = o2f(batched ? hs[sI].glo[k] : hc.glo[k]);
        sl.bounds_hi[k] = o2f(batched ? hs[sI].ghi[k] : hc.ghi[k]);
      }
      if (count[sI] >= 2) {
        SahRoot r;
        r.first = first[sI];
        r.count = count[sI];
        for (int k = 0; k < 3; k++) {
          r.clo[k] = o2f(batched ? hs[sI].clo[k] : hc.clo[k]);
          r.chi[k] = o2f(batched ? hs[sI].chi[k] : hc.chi[k]);
        }
        trees.push_back(r);
        tree_scene.push_back(sI);
      }
    }
    if (pos != n) throw hip_error(hipErrorUnknown, "per-scene primitive counts disagree with the total");
  }
  if (n == 0) {
    out.stats.num_nodes = 0;
    return;
  }
  // spatial splits (one scene of triangles, binned-SAH tree): the tree is
  // built over the pieces, whose centroids may lie anywhere in the scene box
  if (prm.split_budget > 0.0f && !batched && !user && geoms[0].kind == GEOM_TRIANGLES && trees.size() == 1 &&
      (prm.builder == BUILDER_SAH || prm.learned)) {
    const uint32_t nr = spatial_presplit(n, prm.split_budget, S.geoms.as<GeomDesc>(), out.stats.bounds_lo,
                                         out.stats.bounds_hi, S.vals, S.lo4, S.hi4, S.spatial, s);
    if (nr != n) {
      out.stats.num_split = nr - n;
      n = nr;
      out.stats.num_prims = n;
      out.slots[0].num_prims = n;
      count[0] = n;
      trees[0].count = n;
      for (int k = 0; k < 3; k++) {
        trees[0].clo[k] = 2.0f * out.stats.bounds_lo[k];
        trees[0].chi[k] = 2.0f * out.stats.bounds_hi[k];
      }
    }
  }
  if (user) user->n = n;

  S.slo.reserve((size_t)n * 16);
  S.shi.reserve((size_t)n * 16);
  hipLaunchKernelGGL(k_gather_sorted, dim3(div_up(n, BT)), dim3(BT), 0, s, S.vals.as<uint32_t>(), S.lo4.as<float4>(),
                     S.hi4.as<float4>(), n, S.slo.as<float4>(), S.shi.as<float4>());

  const uint32_t ninner = n > 1 ? n - 1 : 1;
  S.child.reserve((size_t)ninner * 8);
  S.parent_int.reserve((size_t)ninner * 4);
  S.parent_leaf.reserve((size_t)n * 4);
  S.range.reserve((size_t)ninner * 8);
  S.nlo.reserve((size_t)ninner * 16);
  S.nhi.reserve((size_t)ninner * 16);
  S.cost.reserve((size_t)ninner * 32);
  S.split8.reserve((size_t)ninner * 8);
  S.isleaf.reserve((size_t)ninner);
  S.visit.reserve((size_t)ninner * 4);
  S.depth.reserve((size_t)ninner);
  std::vector<uint32_t> root_ref(nscenes, REF_EMPTY);
  for (uint32_t sI = 0; sI < nscenes; sI++)
    if (count[sI] == 1) root_ref[sI] = first[sI] | REF_LEAF;
  const bool use_sah = batched || prm.builder == BUILDER_SAH || prm.learned;
  if (!trees.empty()) {
    RTG_HIP(hipMemsetAsync(S.visit.ptr, 0, (size_t)ninner * 4, s));
    if (batched)
      hipLaunchKernelGGL(k_fill_none, dim3(div_up(n, BT)), dim3(BT), 0, s, S.parent_leaf.as<uint32_t>(), n);
    if (use_sah) {
      SahTreeOut so;
      build_sah_tree(n, S.vals.as<uint32_t>(), S.slo.as<float4>(), S.shi.as<float4>(), trees, S.child.as<uint32_t>(),
                     S.parent_int.as<uint32_t>(), S.parent_leaf.as<uint32_t>(), S.range.as<uint2>(),
                     S.depth.as<uint8_t>(), prm.learned, prm.median_below, S.sah, so, s);
      for (size_t k = 0; k < trees.size(); k++) root_ref[tree_scene[k]] = so.roots[k];
      // boxes in the final primitive order
      hipLaunchKernelGGL(k_gather_sorted, dim3(div_up(n, BT)), dim3(BT), 0, s, S.vals.as<uint32_t>(),
                         S.lo4.as<float4>(), S.hi4.as<float4>(), n, S.slo.as<float4>(), S.shi.as<float4>());
    } else {
      hipLaunchKernelGGL(k_radix_tree, dim3(div_up(n - 1, BT)), dim3(BT), 0, s, S.keys.as<uint32_t>(), (int)n,
                         S.child.as<uint32_t>(), S.parent_int.as<uint32_t>(), S.parent_leaf.as<uint32_t>(),
                         S.range.as<uint2>());
      root_ref[0] = 0;
    }
    RefitArgs ra{n, S.child.as<uint32_t>(), S.range.as<uint2>(), S.slo.as<float4>(), S.shi.as<float4>(),
                 S.nlo.as<float4>(), S.nhi.as<float4>(), S.cost.as<float>(), S.split8.as<uint8_t>(),
                 S.isleaf.as<uint8_t>(), prm.cost_node, prm.cost_tri, max_leaf, width, min_leaf, block, &cnt->error};
    // every inner node's depth (recorded by the binned-SAH builder, counted
    // for the radix tree): refit one level per launch, deepest first. A
    // radix tree over 32-bit codes with index tie-breaks splits each range
    // at a longer common prefix than its parent's, so it is at most 65 deep;
    // a count beyond 254 means a broken parent link.
    const uint32_t m = n - 1;
    uint32_t* deep = &cnt->deep;
    if (!use_sah)
      hipLaunchKernelGGL(k_node_depth, dim3(div_up(m, BT)), dim3(BT), 0, s, S.parent_int.as<uint32_t>(), m,
                         S.depth.as<uint8_t>(), deep);
    S.lvl_hist.reserve(2 * 256 * 4);
    S.lvl_order.reserve((size_t)m * 4);
    S.lvl_host.reserve(2 * 257 * 4);
    uint32_t* hist = S.lvl_hist.as<uint32_t>();
    uint32_t* hh = S.lvl_host.as<uint32_t>();  // [0, 256): histogram, 256: deep flag, [257, 513): offsets
    RTG_HIP(hipMemsetAsync(hist, 0, 2 * 256 * 4, s));
    hipLaunchKernelGGL(k_depth_hist, dim3(div_up(m, (size_t)BT * DEPTH_PER)), dim3(BT), 0, s, S.depth.as<uint8_t>(), m,
                       hist);
    RTG_HIP(hipMemcpyAsync(hh, hist, 256 * 4, hipMemcpyDeviceToHost, s));
    RTG_HIP(hipMemcpyAsync(hh + 256, deep, 4, hipMemcpyDeviceToHost, s));
    RTG_HIP(hipStreamSynchronize(s));
    if (hh[256]) throw hip_error(hipErrorUnknown, "radix tree deeper than 254 levels: broken parent links");
    uint32_t* off = hh + 257;
    uint32_t total = 0;
    for (int d = 0; d < 256; d++) {
      off[d] = total;
      total += d < 255 ? hh[d] : 0u;
    }
    RTG_HIP(hipMemcpyAsync(hist + 256, off, 256 * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_depth_scatter, dim3(div_up(m, (size_t)BT * DEPTH_PER)), dim3(BT), 0, s,
                       S.depth.as<uint8_t>(), m, hist + 256, S.lvl_order.as<uint32_t>());
    for (int d = 254; d >= 0; d--)
      if (hh[d])
        hipLaunchKernelGGL(k_refit_level, dim3(div_up(hh[d], BT)), dim3(BT), 0, s, ra,
                           S.lvl_order.as<uint32_t>() + off[d], hh[d], S.visit.as<uint32_t>());
    hipLaunchKernelGGL(k_check_tree, dim3(div_up(n, BT)), dim3(BT), 0, s,
                       use_sah ? nullptr : S.keys.as<uint32_t>(), n, &cnt->roots_seen,
                       S.child.as<uint32_t>(), S.parent_int.as<uint32_t>(), S.parent_leaf.as<uint32_t>(),
                       S.visit.as<uint32_t>(), &cnt->error);
    RTG_HIP(hipGetLastError());
  }
  // tree roots the check must meet (its error bits stop the collapse on the
  // device; both are read back with the collapse's results)
  uint32_t roots_expected = (uint32_t)trees.size();
  for (uint32_t sI = 0; sI < nscenes; sI++) roots_expected += count[sI] == 1 ? 1u : 0u;
  if (user) {
    RTG_HIP(hipMemcpyAsync(&hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, s));
    RTG_HIP(hipStreamSynchronize(s));
    if (!trees.empty() && (hc.error || hc.roots_seen != roots_expected))
      throw hip_error(hipErrorUnknown, "binary tree check failed: error bits " + std::to_string(hc.error) +
                                           ", roots " + std::to_string(hc.roots_seen) + " of " +
                                           std::to_string(roots_expected));
  }

  if (user) {  // the binary tree with its collapse decisions, for the caller's own nodes
    const uint32_t ni = n > 1 ? n - 1 : 0;
    user->root = root_ref[0];
    user->child.resize(2 * (size_t)ni);
    user->range.resize(2 * (size_t)ni);
    user->split8.resize(8 * (size_t)ni);
    user->isleaf.resize(ni);
    user->nbox.resize(8 * (size_t)ni);
    user->pbox.resize(8 * (size_t)n);
    if (ni) {
      RTG_HIP(hipMemcpyAsync(user->child.data(), S.child.ptr, 8 * (size_t)ni, hipMemcpyDeviceToHost, s));
      RTG_HIP(hipMemcpyAsync(user->range.data(), S.range.ptr, 8 * (size_t)ni, hipMemcpyDeviceToHost, s));
      RTG_HIP(hipMemcpyAsync(user->split8.data(), S.split8.ptr, 8 * (size_t)ni, hipMemcpyDeviceToHost, s));
      RTG_HIP(hipMemcpyAsync(user->isleaf.data(), S.isleaf.ptr, ni, hipMemcpyDeviceToHost, s));
      RTG_HIP(hipMemcpy2DAsync(user->nbox.data(), 32, S.nlo.ptr, 16, 16, ni, hipMemcpyDeviceToHost, s));
      RTG_HIP(hipMemcpy2DAsync(user->nbox.data() + 4, 32, S.nhi.ptr, 16, 16, ni, hipMemcpyDeviceToHost, s));
    }
    RTG_HIP(hipMemcpy2DAsync(user->pbox.data(), 32, S.slo.ptr, 16, 16, n, hipMemcpyDeviceToHost, s));
    RTG_HIP(hipMemcpy2DAsync(user->pbox.data() + 4, 32, S.shi.ptr, 16, 16, n, hipMemcpyDeviceToHost, s));
    RTG_HIP(hipStreamSynchronize(s));
    user->width = width;
    out.stats.num_prims = n;
    out.stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return;
  }

  // outputs: at most n + nscenes nodes, exactly n triangle records
  const uint32_t cap = n + nscenes;
  out.nodes.reserve((size_t)cap * sizeof(QNode8));
  const bool inst_build = geoms[0].kind == GEOM_INSTANCES;
  for (const auto& g : geoms)
    if ((g.kind == GEOM_INSTANCES) != inst_build)
      throw hip_error(hipErrorInvalidValue, "a build holds either triangles or instance records");
  out.instances = inst_build;
  if (inst_build) out.insts.reserve((size_t)n * sizeof(InstRec));
  else out.tris.reserve((size_t)n * sizeof(TriRec));
  out.info.reserve((size_t)cap * sizeof(NodeInfo));
  S.queue.reserve((size_t)2 * cap * sizeof(QItem));
  QItem* q0 = S.queue.as<QItem>();
  QItem* q1 = q0 + cap;
  std::vector<QItem> roots(nscenes);
  for (uint32_t sI = 0; sI < nscenes; sI++) roots[sI] = QItem{root_ref[sI], sI, 1u, 0u};
  RTG_HIP(hipMemcpyAsync(q0, roots.data(), nscenes * sizeof(QItem), hipMemcpyHostToDevice, s));

  CollapseArgs ca;
  ca.child = S.child.as<uint32_t>();
  ca.range = S.range.as<uint2>();
  ca.nlo = S.nlo.as<float4>();
  ca.nhi = S.nhi.as<float4>();
  ca.isleaf = S.isleaf.as<uint8_t>();
  ca.split8 = S.split8.as<uint8_t>();
  ca.slo = S.slo.as<float4>();
  ca.shi = S.shi.as<float4>();
  ca.nodes = out.nodes.as<QNode8>();
  S.leaf_src.reserve((size_t)n * 4);
  ca.leaf_src = S.leaf_src.as<uint32_t>();
  ca.info = out.info.as<NodeInfo>();
  ca.cnt = cnt;
  ca.n = cap;
  ca.tri_cap = n;
  ca.width = width;
  // one launch per level of the 8-wide tree; the kernels read their queue
  // lengths on the device, so the host queues level k + 1 before it learns
  // (through a pinned slot) whether level k produced any items
  // 2048 workgroups: a thread per item covers the large levels in one step
  const uint32_t cgrid = std::min<uint32_t>(div_up(cap, BT), 2048u);
  auto enqueue_level = [&](uint32_t k) {
    const uint32_t o = k & 1u;
    ca.qin = o ? q1 : q0;
    ca.qout = o ? q0 : q1;
    ca.nin = &cnt->qcount[(k + 2u) % 3u];
    ca.qout_slot = k % 3u;
    ca.qclear = &cnt->qcount[(k + 1u) % 3u];
    hipLaunchKernelGGL(k_collapse, dim3(cgrid), dim3(BT), 0, s, ca);
    RTG_HIP(hipGetLastError());
    RTG_HIP(hipMemcpyAsync(S.rb.slot((int)k), cnt, sizeof(BuildCounters), hipMemcpyDeviceToHost, s));
    RTG_HIP(hipEventRecord(S.rb.ev[k % LevelReadback::K], s));
  };
  S.rb.ensure(sizeof(BuildCounters));
  std::string level_sizes;
  enqueue_level(0);
  for (uint32_t k = 1;; k++) {
    if (k > 4096) {
      (void)hipStreamSynchronize(s);
      throw hip_error(hipErrorUnknown, "BVH collapse did not terminate");
    }
    enqueue_level(k);  // speculative: without input it does nothing
    RTG_HIP(hipEventSynchronize(S.rb.ev[(k - 1) % LevelReadback::K]));
    memcpy(&hc, S.rb.slot((int)(k - 1)), sizeof(hc));
    const uint32_t next = hc.qcount[(k - 1) % 3u];
    level_sizes += " " + std::to_string(next);
    if (next == 0 || hc.error) break;
    if (next > cap) {
      (void)hipStreamSynchronize(s);
      throw hip_error(hipErrorUnknown, "BVH collapse queue overflow");
    }
  }
  hipLaunchKernelGGL(k_write_leaves, dim3(div_up(n, BT)), dim3(BT), 0, s, S.leaf_src.as<uint32_t>(),
                     S.vals.as<uint32_t>(), S.lo4.as<float4>(), S.hi4.as<float4>(), S.geoms.as<GeomDesc>(), cnt,
                     inst_build ? nullptr : out.tris.as<TriRec>(), inst_build ? out.insts.as<InstRec>() : nullptr,
                     prm.robust ? 1u : 0u);
  RTG_HIP(hipGetLastError());
  RTG_HIP(hipMemcpyAsync(&hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, s));
  RTG_HIP(hipStreamSynchronize(s));
  if (!trees.empty() && (hc.roots_seen != roots_expected || (hc.error & 5u)))
    throw hip_error(hipErrorUnknown, "binary tree check failed: error bits " + std::to_string(hc.error) + ", roots " +
                                         std::to_string(hc.roots_seen) + " of " + std::to_string(roots_expected));
  if (hc.error || hc.tri_count != n)
    throw hip_error(hipErrorUnknown, "BVH build consistency check failed: error bits " + std::to_string(hc.error) +
                                         ", primitives " + std::to_string(n) + ", triangle records " +
                                         std::to_string(hc.tri_count) + ", nodes " + std::to_string(hc.node_count) +
                                         ", queue sizes per level" + level_sizes);
  if (hc.max_stack + 1 > TRACE_STACK)
    throw hip_error(hipErrorUnknown, "BVH too deep for the traversal stack: " + std::to_string(hc.max_stack + 1) +
                                         " entries needed, " + std::to_string(TRACE_STACK) + " available");
  out.stats.num_nodes = hc.node_count;
  out.stats.max_depth = hc.max_depth;
  out.stats.max_stack = hc.max_stack + 1;
  const double root_area = half_area(out.stats.bounds_hi[0] - out.stats.bounds_lo[0],
                                     out.stats.bounds_hi[1] - out.stats.bounds_lo[1],
                                     out.stats.bounds_hi[2] - out.stats.bounds_lo[2]);
  if (root_area > 0) {
    out.stats.sah = (hc.sah_nodes + hc.sah_leaves) / root_area;
    out.stats.sah_tri = (hc.sah_nodes + hc.sah_leaves_tri) / root_area;
  }
  if (hc.error) throw hip_error(hipErrorUnknown, "BVH build consistency check failed (" + std::to_string(hc.error) + ")");
  if (hc.tri_count != n) throw hip_error(hipErrorUnknown, "BVH build lost triangles");
  out.stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace rtg
