// Top-down binned-SAH binary tree, one level of splits per step.
//
// Reference anchors for the heuristic:
//   bin mapping on doubled centroids lo+hi with scale 0.99*num/diag and the
//     bin index clamped to [0, num-1]: kernels/builders/heuristic_binning.h:16-112;
//   per-bin counts and boxes, right-to-left suffix areas, left-to-right
//     sweep keeping the first minimum, partition by bin < pos:
//     kernels/builders/heuristic_binning.h:211-258,340-393,
//     kernels/builders/heuristic_binning_array_aligned.h:37-143;
//   object-median fallback when no split separates the centroids:
//     heuristic_binning_array_aligned.h (splitFallback).
// Differences, none of which can change a hit: 32 bins per axis (min(32, n)
// for n < 32); the split cost is A_L*n_L + A_R*n_R per primitive (leaf sizes
// are decided afterwards by the 8-wide collapse); splitting continues down to
// single primitives; below depth 48 every split is an object median, which
// bounds the depth.
//
// Segments (contiguous position ranges still to split) are processed in three
// size classes:
//   big   (> 4096): positions in chunks of 4096 bin into LDS, the chunk bins
//                   merge into per-segment bins in HBM with atomics, one
//                   thread per segment picks the split, and a global scan of
//                   left flags gives every position its stable destination;
//   mid   (65..4096): one 256-lane workgroup per segment bins, picks and
//                   partitions in LDS;
//   small (9..64): one 64-lane workgroup per segment, same code;
//   tiny  (2..8): one thread finishes the whole subtree with exact SAH
//                 splits (k_tiny), launched right after the level that made
//                 the segment; a root segment of 2..64 starts as small.
// Every inner node gets the id of its split position (first + n_left - 1),
// which is unique, so the output does not depend on the order in which
// segments are processed.
#include <string>

#include "builder_sah.h"
#include "mlp.h"

namespace rtg {
namespace {

constexpr uint32_t NIL = 0xffffffffu;
constexpr int NB = 32;                  // bins per axis
constexpr int BIN_WORDS = 3 * NB * 7;   // counts, then lo xyz, then hi xyz, per (axis, bin)
constexpr uint32_t SMALL_MAX = 64;
constexpr uint32_t TINY_MAX = 16;        // segments up to this size are finished by one thread (exact SAH)
constexpr uint32_t MID_MAX = 4096;
constexpr uint32_t CHUNK = 4096;        // positions per block of the big-segment binning
constexpr uint32_t MEDIAN_DEPTH = 48;   // from this binary depth on, splits are object medians
constexpr int TBS = 256;

__device__ inline int wc(int a, int b) { return a * NB + b; }
__device__ inline int wl(int a, int b, int k) { return 3 * NB + (a * NB + b) * 3 + k; }
__device__ inline int wh(int a, int b, int k) { return 3 * NB + 9 * NB + (a * NB + b) * 3 + k; }

// parent of a tree's root segment: ROOT_TAG | tree index
constexpr uint32_t ROOT_TAG = 0x80000000u;

struct Seg {
  uint32_t first, count, parent, meta;  // meta = depth << 1 | side of the parent
  float clo[3], chi[3];                 // bounds of lo+hi over the segment
};

struct Split {
  float ofs[3], scale[3];
  int axis;  // 0..2: binned split "bin < k" on this axis; 3: object median by position
  int k, nb;
  uint32_t nleft;
  uint32_t lsid, rsid;  // ids of the children in the next big list (NIL if not big)
};

struct Counters {
  uint32_t nseg[4];  // next-level segment counts: big, mid, small, tiny
  uint32_t error;    // bit 0: a segment list overflowed, bit 1: bin counts disagree, bit 2: destination out of range,
                     // bit 3: median subtree stack overflow, bit 4: a tiny segment out of range,
                     // bit 5: a depth beyond 254
};

struct Lists {
  Seg* seg[4];
  uint32_t cap[4];
  uint32_t fin_max;  // segments up to this size are finished at once by median splits
};

struct Tree {
  uint32_t* child;
  uint32_t* parent_int;
  uint32_t* parent_leaf;
  uint2* range;
  uint32_t* vals;
  Counters* cnt;
  uint32_t* roots;  // per tree
  uint8_t* depth;   // binary depth of each inner node (0xff: unused id)
};

__device__ inline void set_depth(const Tree& t, uint32_t m, uint32_t d) {
  if (d > 254u) atomicOr(&t.cnt->error, 32u);
  t.depth[m] = (uint8_t)(d > 254u ? 254u : d);
}

struct Mapping {
  float ofs[3], scale[3];
  int nb;
};

__device__ inline Mapping make_mapping(const Seg& s) {
  Mapping m;
  m.nb = s.count < (uint32_t)NB ? (int)s.count : NB;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float diag = s.chi[k] - s.clo[k];
    m.ofs[k] = s.clo[k];
    m.scale[k] = diag > 1e-34f ? 0.99f * (float)m.nb / diag : 0.0f;
  }
  return m;
}

__device__ inline int bin_of(float c2, float ofs, float scale, int nb) {
  const int b = (int)floorf((c2 - ofs) * scale);
  return b < 0 ? 0 : (b > nb - 1 ? nb - 1 : b);
}

__device__ inline float comp(float4 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

__device__ inline bool goes_left(const Split& sp, float4 l, float4 h, uint32_t rel) {
  if (sp.axis == 3) return rel < sp.nleft;
  const float c2 = comp(l, sp.axis) + comp(h, sp.axis);
  return bin_of(c2, sp.ofs[sp.axis], sp.scale[sp.axis], sp.nb) < sp.k;
}

__device__ inline void init_bins(uint32_t* w, int tid, int nthreads) {
  for (int i = tid; i < BIN_WORDS; i += nthreads) w[i] = i < 3 * NB ? 0u : (i < 3 * NB + 9 * NB ? NIL : 0u);
}

__device__ inline void bin_prim(uint32_t* w, const Mapping& m, float4 l, float4 h) {
  const float c[3] = {l.x + h.x, l.y + h.y, l.z + h.z};
  const uint32_t ol[3] = {f2o(l.x), f2o(l.y), f2o(l.z)};
  const uint32_t oh[3] = {f2o(h.x), f2o(h.y), f2o(h.z)};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const int b = bin_of(c[a], m.ofs[a], m.scale[a], m.nb);
    atomicAdd(&w[wc(a, b)], 1u);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      atomicMin(&w[wl(a, b, k)], ol[k]);
      atomicMax(&w[wh(a, b, k)], oh[k]);
    }
  }
}

// Best "bin < k" split along axis a from bin words w (LDS or HBM): least
// A_L*n_L + A_R*n_R, first minimum in k. Returns false when no k separates
// the primitives.
__device__ inline bool best_on_axis(const uint32_t* w, const Mapping& m, int a, float& cost, int& kbest,
                                    uint32_t& nleft) {
  if (m.scale[a] == 0.0f) return false;
  float ra[NB];
  uint32_t rn[NB];
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  uint32_t nn = 0;
  for (int b = m.nb - 1; b >= 1; b--) {
    const uint32_t cb = w[wc(a, b)];
    if (cb) {
      nn += cb;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        lo[k] = fminf(lo[k], o2f(w[wl(a, b, k)]));
        hi[k] = fmaxf(hi[k], o2f(w[wh(a, b, k)]));
      }
    }
    rn[b] = nn;
    ra[b] = nn ? half_area(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]) : 0.0f;
  }
  bool found = false;
  cost = INFINITY;
  nn = 0;
#pragma unroll
  for (int k = 0; k < 3; k++) { lo[k] = INFINITY; hi[k] = -INFINITY; }
  for (int b = 0; b + 1 < m.nb; b++) {
    const uint32_t cb = w[wc(a, b)];
    if (cb) {
      nn += cb;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        lo[k] = fminf(lo[k], o2f(w[wl(a, b, k)]));
        hi[k] = fmaxf(hi[k], o2f(w[wh(a, b, k)]));
      }
    }
    const uint32_t nr = rn[b + 1];
    if (nn == 0 || nr == 0) continue;
    const float c = half_area(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]) * (float)nn + ra[b + 1] * (float)nr;
    if (c < cost) {
      cost = c;
      kbest = b + 1;
      nleft = nn;
      found = true;
    }
  }
  return found;
}

// Box of bins [b0, b1) on axis a.
__device__ inline void bins_box(const uint32_t* w, int a, int b0, int b1, float lo[3], float hi[3]) {
#pragma unroll
  for (int k = 0; k < 3; k++) { lo[k] = INFINITY; hi[k] = -INFINITY; }
  for (int b = b0; b < b1; b++) {
    if (!w[wc(a, b)]) continue;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      lo[k] = fminf(lo[k], o2f(w[wl(a, b, k)]));
      hi[k] = fmaxf(hi[k], o2f(w[wh(a, b, k)]));
    }
  }
}

// Records inner node m (the split of segment sg) in the tree.
__device__ inline void link_inner(const Tree& t, const Seg& sg, uint32_t m) {
  if (sg.parent & ROOT_TAG) {
    t.roots[sg.parent & ~ROOT_TAG] = m;
    t.parent_int[m] = NIL;
  } else {
    t.child[2 * sg.parent + (sg.meta & 1u)] = m;
    t.parent_int[m] = sg.parent;
  }
  t.range[m] = make_uint2(sg.first, sg.first + sg.count - 1);
  set_depth(t, m, sg.meta >> 1);
}

// The whole subtree of a small segment [first, first+count) (child `side`
// of inner node m at depth+1), split at the median position all the way
// down. Positions keep the order the binned splits left them in (stable
// partitions of the Morton-sorted references), so these are spatial-median
// splits in Morton order; nothing moves, only links are written.
__device__ inline void finish_median(const Tree& t, uint32_t m, uint32_t side, uint32_t depth, uint32_t first,
                                     uint32_t count) {
  uint32_t sf[16], sc[16], sp_[16], sm[16];
  int sp = 0;
  sf[0] = first; sc[0] = count; sp_[0] = m; sm[0] = ((depth + 1) << 1) | side;
  sp = 1;
  while (sp > 0) {
    sp--;
    const uint32_t f = sf[sp], c = sc[sp], p = sp_[sp], meta = sm[sp];
    const uint32_t nl = c / 2, mid = f + nl - 1, d = meta >> 1;
    t.child[2 * p + (meta & 1u)] = mid;
    t.parent_int[mid] = p;
    t.range[mid] = make_uint2(f, f + c - 1);
    set_depth(t, mid, d);
    const uint32_t cf[2] = {f, f + nl}, cc[2] = {nl, c - nl};
#pragma unroll
    for (int s = 0; s < 2; s++) {
      if (cc[s] == 1) {
        t.child[2 * mid + s] = cf[s] | REF_LEAF;
        t.parent_leaf[cf[s]] = mid;
      } else if (sp < 16) {
        sf[sp] = cf[s]; sc[sp] = cc[s]; sp_[sp] = mid; sm[sp] = ((d + 1) << 1) | (uint32_t)s;
        sp++;
      } else {
        atomicOr(&t.cnt->error, 8u);
      }
    }
  }
}

// Child segment [first, first+count) of inner node m: a leaf link for a
// single primitive, a finished median subtree up to nx.fin_max primitives,
// else a new segment in the list of its size class. Returns the id in the
// big list, or NIL.
__device__ inline uint32_t emit_child(const Tree& t, const Lists& nx, uint32_t m, uint32_t side, uint32_t depth,
                                      uint32_t first, uint32_t count, const float clo[3], const float chi[3]) {
  if (count == 1) {
    t.child[2 * m + side] = first | REF_LEAF;
    t.parent_leaf[first] = m;
    return NIL;
  }
  if (count <= nx.fin_max) {
    finish_median(t, m, side, depth, first, count);
    return NIL;
  }
  const int cls = count > MID_MAX ? 0 : (count > SMALL_MAX ? 1 : (count > TINY_MAX ? 2 : 3));
  const uint32_t id = atomicAdd(&t.cnt->nseg[cls], 1u);
  if (id >= nx.cap[cls]) {
    atomicOr(&t.cnt->error, 1u);
    return NIL;
  }
  Seg c;
  c.first = first;
  c.count = count;
  c.parent = m;
  c.meta = ((depth + 1) << 1) | side;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    c.clo[k] = clo[k];
    c.chi[k] = chi[k];
  }
  nx.seg[cls][id] = c;
  return cls == 0 ? id : NIL;
}

// Workgroup-level choice of the split of segment sg from its bins w (LDS or
// HBM); every thread returns the same Split. Plain SAH: one thread per
// axis sweeps its bins. Learned cost (lc.P set and the segment larger than
// lc.min_count): the workgroup evaluates the model on the left and right
// boxes of all 3*(nb-1) candidates (their boxes and counts from the bins),
// then one thread per axis keeps the least P(L)*n_L + P(R)*n_R with
// P = halfArea * 2^model (bounds below) among the candidates whose plain
// cost A(L)*n_L + A(R)*n_R is within (1 + lc.sah_slack) of the least over
// all axes.
// Ties keep the lowest axis and the lowest k; without a separating
// candidate the split is the object median.
struct ChooseLds {
  float yp, pp;  // learned log rate and visit weight of the segment's own box
  float cost[3];
  float amin[3];  // learned: least plain cost per axis
  int k[3];
  uint32_t nl[3];
  float p[2 * 3 * (NB - 1)];  // learned: visit weight of each candidate's left and right box
  float a[2 * 3 * (NB - 1)];  // learned: their half areas
  Split sp;
};

__device__ inline void range_box(const uint32_t* w, int a, int b0, int b1, float lo[3], float hi[3], uint32_t& cnt) {
  bins_box(w, a, b0, b1, lo, hi);
  cnt = 0;
  for (int b = b0; b < b1; b++) cnt += w[wc(a, b)];
}

// Plain SAH cost of a candidate from its two half areas and counts.
__device__ inline float split_area_cost(float al, float ar, uint32_t nl, uint32_t nr) {
  return al * (float)nl + ar * (float)nr;
}

template <int BLOCK>
__device__ inline Split choose_split(const uint32_t* w, const Mapping& m, const Seg& sg, const LearnedCost& lc,
                                     ChooseLds& L, uint32_t* err) {
  const int tid = threadIdx.x;
  const uint32_t depth = sg.meta >> 1;
  const bool learned = lc.P != nullptr && sg.count > lc.min_count && depth < MEDIAN_DEPTH;
  const int ncand = m.nb - 1;
  if (learned) {
    // the segment's own prediction bounds its children's: a child is never
    // more likely to be visited than its parent, and its learned rate per
    // unit area stays within 2x of the parent's, which stays within 4x of
    // the uniform-ray rate. A model queried far from its training boxes
    // could otherwise price a split at almost nothing and the tree would
    // degenerate (measured: 6 Mrays/s after 3000 training steps with an
    // 8x band).
    if (tid == 0) {
      float lo[3], hi[3];
      uint32_t cnt;
      range_box(w, 0, 0, m.nb, lo, hi, cnt);
      float f[MLP_IN];
      box_features(lo, hi, sg.count, lc.sn, f);
      L.yp = fminf(fmaxf(mlp_eval(lc.P, f), -2.0f), 2.0f);
      L.pp = half_area(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]) * exp2f(L.yp);
    }
    __syncthreads();
    const float yp = L.yp, pp = L.pp;
    for (int e = tid; e < 2 * 3 * ncand; e += BLOCK) {
      const int c = e >> 1, side = e & 1;
      const int a = c / ncand, k = c % ncand + 1;
      float lo[3], hi[3];
      uint32_t cnt;
      range_box(w, a, side ? k : 0, side ? m.nb : k, lo, hi, cnt);
      float pv = -1.0f, av = -1.0f;
      if (cnt > 0 && m.scale[a] != 0.0f) {
        float f[MLP_IN];
        box_features(lo, hi, cnt, lc.sn, f);
        const float y = fminf(fmaxf(mlp_eval(lc.P, f), yp - 1.0f), yp + 1.0f);
        av = half_area(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]);
        pv = fminf(av * exp2f(y), pp);
      }
      L.p[e] = pv;
      L.a[e] = av;
    }
    __syncthreads();
    if (tid < 3) {
      float c = INFINITY;
      uint32_t run = 0, total = 0;
      for (int b = 0; b < m.nb; b++) total += w[wc(tid, b)];
      for (int kk = 1; kk <= ncand; kk++) {
        run += w[wc(tid, kk - 1)];
        const float al = L.a[2 * (tid * ncand + kk - 1)], ar = L.a[2 * (tid * ncand + kk - 1) + 1];
        if (run == 0 || run == total || al < 0.0f || ar < 0.0f) continue;
        c = fminf(c, split_area_cost(al, ar, run, total - run));
      }
      L.amin[tid] = c;
    }
    __syncthreads();
  }
  if (tid < 3) {
    float c = INFINITY;
    int k = 0;
    uint32_t nl = 0;
    uint32_t total = 0;
    for (int b = 0; b < m.nb; b++) total += w[wc(tid, b)];
    if (total != sg.count) atomicOr(err, 2u);
    if (depth < MEDIAN_DEPTH && total == sg.count) {
      if (!learned) {
        if (!best_on_axis(w, m, tid, c, k, nl)) c = INFINITY;
      } else if (m.scale[tid] != 0.0f) {
        // an unbounded slack admits every candidate (also where the least
        // plain cost is 0, which would make 0 * inf a NaN)
        const float guard = __builtin_isinf(lc.sah_slack)
                                ? INFINITY
                                : fminf(fminf(L.amin[0], L.amin[1]), L.amin[2]) * (1.0f + lc.sah_slack);
        uint32_t run = 0;
        for (int kk = 1; kk <= ncand; kk++) {
          run += w[wc(tid, kk - 1)];
          const int e = 2 * (tid * ncand + kk - 1);
          const float pl = L.p[e], pr = L.p[e + 1];
          if (run == 0 || run == total || pl < 0.0f || pr < 0.0f) continue;
          if (!(split_area_cost(L.a[e], L.a[e + 1], run, total - run) <= guard)) continue;
          const float v = pl * (float)run + pr * (float)(total - run);
          if (v < c) {
            c = v;
            k = kk;
            nl = run;
          }
        }
      }
    }
    L.cost[tid] = c;
    L.k[tid] = k;
    L.nl[tid] = nl;
  }
  __syncthreads();
  if (tid == 0) {
    Split sp;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      sp.ofs[k] = m.ofs[k];
      sp.scale[k] = m.scale[k];
    }
    sp.nb = m.nb;
    sp.axis = 3;
    sp.k = 0;
    sp.nleft = sg.count / 2;
    float best = INFINITY;
    for (int a = 0; a < 3; a++) {
      if (L.cost[a] < best) {
        best = L.cost[a];
        sp.axis = a;
        sp.k = L.k[a];
        sp.nleft = L.nl[a];
      }
    }
    sp.lsid = sp.rsid = NIL;
    L.sp = sp;
  }
  __syncthreads();
  return L.sp;
}

// ------------------------------------------------------------ big segments --

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  const size_t i = (size_t)blockIdx.x * TBS + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_init_gbins(uint32_t* g, uint32_t nseg) {
  const size_t i = (size_t)blockIdx.x * TBS + threadIdx.x;
  if (i >= (size_t)nseg * BIN_WORDS) return;
  const int w = (int)(i % BIN_WORDS);
  g[i] = w < 3 * NB ? 0u : (w < 3 * NB + 9 * NB ? NIL : 0u);
}

__global__ void k_zero_counts(Counters* c) {
  if (threadIdx.x < 4) c->nseg[threadIdx.x] = 0u;
}

// A chunk of CHUNK positions meets at most two big segments (each is longer
// than the chunk): they are binned in LDS and merged into HBM; any other
// segment id (not expected) bins straight into HBM.
__global__ __launch_bounds__(TBS) void k_big_bins(const uint32_t* __restrict__ sid, const float4* __restrict__ blo,
                                                  const float4* __restrict__ bhi, uint32_t n,
                                                  const Seg* __restrict__ segs, uint32_t* gbins) {
  __shared__ uint32_t sb[2][BIN_WORDS];
  __shared__ uint32_t s_lo, s_hi;
  __shared__ Mapping sm[2];
  const int tid = threadIdx.x;
  const uint32_t c0 = blockIdx.x * CHUNK;
  const uint32_t c1 = c0 + CHUNK < n ? c0 + CHUNK : n;
  if (tid == 0) {
    s_lo = NIL;
    s_hi = 0u;
  }
  init_bins(sb[0], tid, TBS);
  init_bins(sb[1], tid, TBS);
  __syncthreads();
  uint32_t mn = NIL, mx = 0u;
  for (uint32_t i = c0 + tid; i < c1; i += TBS) {
    const uint32_t s = sid[i];
    if (s != NIL) {
      mn = s < mn ? s : mn;
      mx = s > mx ? s : mx;
    }
  }
  if (mn != NIL) {
    atomicMin(&s_lo, mn);
    atomicMax(&s_hi, mx);
  }
  __syncthreads();
  const uint32_t slo_id = s_lo, shi_id = s_hi;
  if (slo_id == NIL) return;  // no big segment here (uniform)
  if (tid == 0) sm[0] = make_mapping(segs[slo_id]);
  if (tid == 1) sm[1] = make_mapping(segs[shi_id]);
  __syncthreads();
  for (uint32_t i = c0 + tid; i < c1; i += TBS) {
    const uint32_t s = sid[i];
    if (s == NIL) continue;
    const float4 l = blo[i], h = bhi[i];
    if (s == slo_id) bin_prim(sb[0], sm[0], l, h);
    else if (s == shi_id) bin_prim(sb[1], sm[1], l, h);
    else bin_prim(gbins + (size_t)s * BIN_WORDS, make_mapping(segs[s]), l, h);
  }
  __syncthreads();
  for (int slot = 0; slot < 2; slot++) {
    if (slot == 1 && shi_id == slo_id) break;
    uint32_t* g = gbins + (size_t)(slot ? shi_id : slo_id) * BIN_WORDS;
    const uint32_t* w = sb[slot];
    for (int i = tid; i < BIN_WORDS; i += TBS) {
      if (i < 3 * NB) {
        if (w[i]) atomicAdd(&g[i], w[i]);
      } else if (i < 3 * NB + 9 * NB) {
        if (w[(i - 3 * NB) / 3]) atomicMin(&g[i], w[i]);
      } else {
        if (w[(i - 3 * NB - 9 * NB) / 3]) atomicMax(&g[i], w[i]);
      }
    }
  }
}

__global__ __launch_bounds__(TBS) void k_big_split(const Seg* __restrict__ segs, const uint32_t* __restrict__ gbins,
                                                   Split* __restrict__ splits, Lists nx, Tree t, LearnedCost lc) {
  __shared__ ChooseLds L;
  const uint32_t i = blockIdx.x;
  const Seg sg = segs[i];
  const uint32_t depth = sg.meta >> 1;
  const Mapping m = make_mapping(sg);
  const uint32_t* w = gbins + (size_t)i * BIN_WORDS;
  Split sp = choose_split<TBS>(w, m, sg, lc, L, &t.cnt->error);
  if (threadIdx.x != 0) return;
  // child centroid bounds: lo+hi of a primitive lies in [2 lo, 2 hi] of its
  // side's box, and in the parent's bounds
  float lcl[3], lch[3], rcl[3], rch[3];
  if (sp.axis < 3) {
    float bl[3], bh[3];
    bins_box(w, sp.axis, 0, sp.k, bl, bh);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      lcl[k] = fmaxf(sg.clo[k], 2.0f * bl[k]);
      lch[k] = fminf(sg.chi[k], 2.0f * bh[k]);
    }
    bins_box(w, sp.axis, sp.k, sp.nb, bl, bh);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      rcl[k] = fmaxf(sg.clo[k], 2.0f * bl[k]);
      rch[k] = fminf(sg.chi[k], 2.0f * bh[k]);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      lcl[k] = rcl[k] = sg.clo[k];
      lch[k] = rch[k] = sg.chi[k];
    }
  }
  const uint32_t mid = sg.first + sp.nleft - 1;
  link_inner(t, sg, mid);
  sp.lsid = emit_child(t, nx, mid, 0, depth, sg.first, sp.nleft, lcl, lch);
  sp.rsid = emit_child(t, nx, mid, 1, depth, sg.first + sp.nleft, sg.count - sp.nleft, rcl, rch);
  splits[i] = sp;
}

__global__ __launch_bounds__(TBS) void k_big_flags(const uint32_t* __restrict__ sid, const float4* __restrict__ blo,
                                                   const float4* __restrict__ bhi, uint32_t n,
                                                   const Seg* __restrict__ segs, const Split* __restrict__ splits,
                                                   uint32_t* __restrict__ flags) {
  const uint32_t i = blockIdx.x * TBS + threadIdx.x;
  if (i >= n) return;
  const uint32_t s = sid[i];
  uint32_t f = 0u;
  if (s != NIL) f = goes_left(splits[s], blo[i], bhi[i], i - segs[s].first) ? 1u : 0u;
  flags[i] = f;
}

__global__ __launch_bounds__(TBS) void k_big_scatter(const uint32_t* __restrict__ sid, const uint32_t* __restrict__ idx,
                                                     const float4* __restrict__ blo, const float4* __restrict__ bhi,
                                                     uint32_t n, const Seg* __restrict__ segs,
                                                     const Split* __restrict__ splits,
                                                     const uint32_t* __restrict__ flags,
                                                     const uint32_t* __restrict__ scan, uint32_t* __restrict__ sid_o,
                                                     uint32_t* __restrict__ idx_o, float4* __restrict__ blo_o,
                                                     float4* __restrict__ bhi_o, uint32_t* __restrict__ vals,
                                                     Counters* cnt) {
  const uint32_t i = blockIdx.x * TBS + threadIdx.x;
  if (i >= n) return;
  const uint32_t s = sid[i];
  if (s == NIL) return;
  const uint32_t first = segs[s].first;
  const Split& sp = splits[s];
  const uint32_t lr = scan[i] - scan[first];
  const bool left = flags[i] != 0u;
  const uint32_t dst = left ? first + lr : first + sp.nleft + (i - first - lr);
  if (dst >= first + segs[s].count) {
    atomicOr(&cnt->error, 4u);
    return;
  }
  const uint32_t v = idx[i];
  idx_o[dst] = v;
  blo_o[dst] = blo[i];
  bhi_o[dst] = bhi[i];
  vals[dst] = v;
  sid_o[dst] = left ? sp.lsid : sp.rsid;
}

// ------------------------------------------------- mid and small segments --

template <int BLOCK, int PER>
__global__ __launch_bounds__(BLOCK) void k_seg(const Seg* __restrict__ segs, const uint32_t* __restrict__ idx,
                                               const float4* __restrict__ blo, const float4* __restrict__ bhi,
                                               uint32_t* __restrict__ idx_o, float4* __restrict__ blo_o,
                                               float4* __restrict__ bhi_o, Lists nx, Tree t, LearnedCost lc) {
  static_assert(PER <= 32, "flags are kept in one 32-bit mask");
  constexpr int NW = BLOCK / 64;
  __shared__ uint32_t sb[BIN_WORDS];
  __shared__ ChooseLds L;
  __shared__ uint32_t s_wsum[NW];
  __shared__ uint32_t s_cb[12];  // ordered bounds of lo+hi: left lo3, left hi3, right lo3, right hi3
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const Seg sg = segs[blockIdx.x];
  const uint32_t depth = sg.meta >> 1;
  const Mapping m = make_mapping(sg);
  init_bins(sb, tid, BLOCK);
  if (tid < 12) s_cb[tid] = (tid % 6) < 3 ? NIL : 0u;
  __syncthreads();
  const uint32_t end = sg.first + sg.count;
  const uint32_t p0 = sg.first + (uint32_t)tid * PER;
#pragma unroll 4
  for (int j = 0; j < PER; j++) {
    const uint32_t p = p0 + j;
    if (p < end) bin_prim(sb, m, blo[p], bhi[p]);
  }
  __syncthreads();
  const Split sp = choose_split<BLOCK>(sb, m, sg, lc, L, &t.cnt->error);
  // flags of this thread's positions, and the stable destinations
  uint32_t mask = 0u, nl = 0u;
#pragma unroll 4
  for (int j = 0; j < PER; j++) {
    const uint32_t p = p0 + j;
    if (p < end && goes_left(sp, blo[p], bhi[p], p - sg.first)) {
      mask |= 1u << j;
      nl++;
    }
  }
  uint32_t x = nl;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) s_wsum[wave] = x;
  __syncthreads();
  uint32_t lbase = x - nl;
  for (int w = 0; w < wave; w++) lbase += s_wsum[w];
  float bnd[12];
#pragma unroll
  for (int k = 0; k < 12; k++) bnd[k] = (k % 6) < 3 ? INFINITY : -INFINITY;
  uint32_t lr = 0u;
  for (int j = 0; j < PER; j++) {
    const uint32_t p = p0 + j;
    if (p >= end) break;
    const bool left = (mask >> j) & 1u;
    const uint32_t rel = p - sg.first;
    const uint32_t dst = left ? sg.first + lbase + lr : sg.first + sp.nleft + (rel - lbase - lr);
    if (left) lr++;
    if (dst >= end) {
      atomicOr(&t.cnt->error, 4u);
      continue;
    }
    const float4 l = blo[p], h = bhi[p];
    const uint32_t v = idx[p];
    idx_o[dst] = v;
    blo_o[dst] = l;
    bhi_o[dst] = h;
    t.vals[dst] = v;
    const float c[3] = {l.x + h.x, l.y + h.y, l.z + h.z};
    const int o = left ? 0 : 6;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      bnd[o + k] = fminf(bnd[o + k], c[k]);
      bnd[o + 3 + k] = fmaxf(bnd[o + 3 + k], c[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < 12; k++) {
    float v = bnd[k];
    const bool mx = (k % 6) >= 3;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
      const float y = __shfl_xor(v, d, 64);
      v = mx ? fmaxf(v, y) : fminf(v, y);
    }
    if (lane == 0 && !__builtin_isinf(v)) {
      if (mx) atomicMax(&s_cb[k], f2o(v)); else atomicMin(&s_cb[k], f2o(v));
    }
  }
  __syncthreads();
  if (tid == 0) {
    const uint32_t mid = sg.first + sp.nleft - 1;
    link_inner(t, sg, mid);
    float lcl[3], lch[3], rcl[3], rch[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      lcl[k] = o2f(s_cb[k]);
      lch[k] = o2f(s_cb[3 + k]);
      rcl[k] = o2f(s_cb[6 + k]);
      rch[k] = o2f(s_cb[9 + k]);
    }
    emit_child(t, nx, mid, 0, depth, sg.first, sp.nleft, lcl, lch);
    emit_child(t, nx, mid, 1, depth, sg.first + sp.nleft, sg.count - sp.nleft, rcl, rch);
  }
}

// ------------------------------------------------------------ tiny segments --

constexpr int TT = 32;  // threads per block of k_tiny, one segment each

// A thread's entries in LDS, entry j at j * TT (the rows of a block's
// threads interleave, so that the lanes of a wave touch consecutive words
// instead of one bank).
struct TinyRows {
  float4* lo;
  float4* hi;
  uint32_t* id;
  __device__ inline float4& L(uint32_t j) const { return lo[j * TT]; }
  __device__ inline float4& H(uint32_t j) const { return hi[j * TT]; }
  __device__ inline uint32_t& I(uint32_t j) const { return id[j * TT]; }
};

// Stable insertion sort of entries [f, f+c) by centroid (lo + hi) on axis a.
__device__ inline void tiny_sort(const TinyRows& e, uint32_t f, uint32_t c, int a) {
  for (uint32_t j = f + 1; j < f + c; j++) {
    const float4 l = e.L(j), h = e.H(j);
    const uint32_t v = e.I(j);
    const float key = comp(l, a) + comp(h, a);
    uint32_t k = j;
    while (k > f && comp(e.L(k - 1), a) + comp(e.H(k - 1), a) > key) {
      e.L(k) = e.L(k - 1);
      e.H(k) = e.H(k - 1);
      e.I(k) = e.I(k - 1);
      k--;
    }
    e.L(k) = l;
    e.H(k) = h;
    e.I(k) = v;
  }
}

// One thread finishes the whole subtree of a segment of 2..TINY_MAX
// references: every split is the exact SAH optimum over the references
// sorted by centroid on each axis (least A_L*n_L + A_R*n_R, first minimum
// over axes then positions), inner node ids are split positions as for the
// binned levels, and vals receives the final order. The references are
// staged in this thread's rows of LDS.
__global__ __launch_bounds__(TT) void k_tiny(const Seg* __restrict__ segs, uint32_t nseg,
                                             const uint32_t* __restrict__ idx, const float4* __restrict__ blo,
                                             const float4* __restrict__ bhi, Tree t) {
  __shared__ float4 s_lo[TINY_MAX][TT], s_hi[TINY_MAX][TT];
  __shared__ uint32_t s_id[TINY_MAX][TT];
  const uint32_t i = blockIdx.x * TT + threadIdx.x;
  if (i >= nseg) return;
  const Seg sg = segs[i];
  if (sg.count < 2 || sg.count > TINY_MAX) {
    atomicOr(&t.cnt->error, 16u);
    return;
  }
  const TinyRows e{&s_lo[0][threadIdx.x], &s_hi[0][threadIdx.x], &s_id[0][threadIdx.x]};
  const uint32_t m = sg.count;
  for (uint32_t j = 0; j < m; j++) {
    e.L(j) = blo[sg.first + j];
    e.H(j) = bhi[sg.first + j];
    e.I(j) = idx[sg.first + j];
  }
  uint32_t sf[TINY_MAX], sc[TINY_MAX], spar[TINY_MAX], smeta[TINY_MAX];
  int sp = 0;
  sf[0] = 0; sc[0] = m; spar[0] = sg.parent; smeta[0] = sg.meta;
  sp = 1;
  while (sp > 0) {
    sp--;
    const uint32_t f = sf[sp], c = sc[sp], p = spar[sp], meta = smeta[sp];
    const uint32_t side = meta & 1u, depth = meta >> 1;
    if (c == 1) {
      t.child[2 * p + side] = (sg.first + f) | REF_LEAF;
      t.parent_leaf[sg.first + f] = p;
      continue;
    }
    float best = INFINITY;
    int ba = 0;
    uint32_t bk = 1;
    for (int a = 0; a < 3; a++) {
      tiny_sort(e, f, c, a);
      float ra[TINY_MAX];
      float bl[3] = {INFINITY, INFINITY, INFINITY}, bh[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
      for (int k = TINY_MAX - 1; k >= 1; k--) {
        if ((uint32_t)k < c) {
          const float4 l = e.L(f + k), h = e.H(f + k);
          bl[0] = fminf(bl[0], l.x); bl[1] = fminf(bl[1], l.y); bl[2] = fminf(bl[2], l.z);
          bh[0] = fmaxf(bh[0], h.x); bh[1] = fmaxf(bh[1], h.y); bh[2] = fmaxf(bh[2], h.z);
          ra[k] = half_area(bh[0] - bl[0], bh[1] - bl[1], bh[2] - bl[2]);
        }
      }
      bl[0] = bl[1] = bl[2] = INFINITY;
      bh[0] = bh[1] = bh[2] = -INFINITY;
#pragma unroll
      for (int k = 1; k < (int)TINY_MAX; k++) {
        if ((uint32_t)k < c) {
          const float4 l = e.L(f + k - 1), h = e.H(f + k - 1);
          bl[0] = fminf(bl[0], l.x); bl[1] = fminf(bl[1], l.y); bl[2] = fminf(bl[2], l.z);
          bh[0] = fmaxf(bh[0], h.x); bh[1] = fmaxf(bh[1], h.y); bh[2] = fmaxf(bh[2], h.z);
          const float cost = half_area(bh[0] - bl[0], bh[1] - bl[1], bh[2] - bl[2]) * (float)k +
                             ra[k] * (float)(c - (uint32_t)k);
          if (cost < best) {
            best = cost;
            ba = a;
            bk = (uint32_t)k;
          }
        }
      }
    }
    if (ba != 2) tiny_sort(e, f, c, ba);  // the last sort left axis 2's order
    const uint32_t mid = sg.first + f + bk - 1;
    t.child[2 * p + side] = mid;
    t.parent_int[mid] = p;
    t.range[mid] = make_uint2(sg.first + f, sg.first + f + c - 1);
    set_depth(t, mid, depth);
    if (sp + 2 > (int)TINY_MAX) {
      atomicOr(&t.cnt->error, 16u);
      return;
    }
    sf[sp] = f + bk; sc[sp] = c - bk; spar[sp] = mid; smeta[sp] = ((depth + 1) << 1) | 1u; sp++;
    sf[sp] = f;      sc[sp] = bk;     spar[sp] = mid; smeta[sp] = ((depth + 1) << 1);      sp++;
  }
  for (uint32_t j = 0; j < m; j++) t.vals[sg.first + j] = e.I(j);
}

// Writes the big-segment id of every position of the initial big segments.
__global__ void k_fill_seg_ids(uint32_t* sid, const Seg* segs) {
  const Seg sg = segs[blockIdx.x];
  for (uint32_t i = threadIdx.x; i < sg.count; i += TBS) sid[sg.first + i] = blockIdx.x;
}

}  // namespace

void build_sah_tree(uint32_t n, uint32_t* vals, const float4* slo, const float4* shi,
                    const std::vector<SahRoot>& trees, uint32_t* child, uint32_t* parent_int,
                    uint32_t* parent_leaf, uint2* range, uint8_t* depth, const LearnedCost* learned,
                    uint32_t fin_max, SahScratch& S, SahTreeOut& out, hipStream_t s) {
  if (fin_max > SMALL_MAX) fin_max = SMALL_MAX;
  LearnedCost lcost{};
  if (learned) lcost = *learned;
  if (trees.empty()) throw hip_error(hipErrorInvalidValue, "SAH build without a tree");
  if (trees.size() >= ROOT_TAG) throw hip_error(hipErrorInvalidValue, "too many trees");
  for (const SahRoot& r : trees)
    if (r.count < 2 || (uint64_t)r.first + r.count > n) throw hip_error(hipErrorInvalidValue, "bad SAH tree range");
  const uint32_t cap[4] = {n / MID_MAX + 2, n / (SMALL_MAX + 1) + 2, n / (TINY_MAX + 1) + 2, n / 2 + 2};
  for (int p = 0; p < 2; p++) {
    S.idx[p].reserve((size_t)n * 4);
    S.blo[p].reserve((size_t)n * 16);
    S.bhi[p].reserve((size_t)n * 16);
    S.sid[p].reserve((size_t)n * 4);
    for (int c = 0; c < 4; c++) S.segs[c][p].reserve((size_t)cap[c] * sizeof(Seg));
  }
  S.bins.reserve((size_t)cap[0] * BIN_WORDS * 4);
  S.splits.reserve((size_t)cap[0] * sizeof(Split));
  S.flags.reserve((size_t)n * 4);
  S.scan.reserve((size_t)n * 4);
  S.counters.reserve(sizeof(Counters));
  S.roots.reserve(trees.size() * 4);
  Counters* cnt = S.counters.as<Counters>();

  RTG_HIP(hipMemcpyAsync(S.idx[0].ptr, vals, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
  RTG_HIP(hipMemcpyAsync(S.blo[0].ptr, slo, (size_t)n * 16, hipMemcpyDeviceToDevice, s));
  RTG_HIP(hipMemcpyAsync(S.bhi[0].ptr, shi, (size_t)n * 16, hipMemcpyDeviceToDevice, s));
  // inner node ids no tree uses stay marked
  hipLaunchKernelGGL(k_fill_u32, dim3(div_up(n - 1, TBS)), dim3(TBS), 0, s, parent_int, (size_t)(n - 1), REF_EMPTY);
  hipLaunchKernelGGL(k_fill_u32, dim3(div_up(trees.size(), TBS)), dim3(TBS), 0, s, S.roots.as<uint32_t>(),
                     trees.size(), NIL);
  RTG_HIP(hipMemsetAsync(depth, 0xff, (size_t)(n - 1), s));

  std::vector<Seg> init[3];
  for (size_t i = 0; i < trees.size(); i++) {
    const SahRoot& r = trees[i];
    Seg sg;
    sg.first = r.first;
    sg.count = r.count;
    sg.parent = ROOT_TAG | (uint32_t)i;
    sg.meta = 0;
    for (int k = 0; k < 3; k++) {
      sg.clo[k] = r.clo[k];
      sg.chi[k] = r.chi[k];
    }
    init[r.count > MID_MAX ? 0 : (r.count > SMALL_MAX ? 1 : 2)].push_back(sg);
  }
  uint32_t cur[3];
  for (int c = 0; c < 3; c++) {
    cur[c] = (uint32_t)init[c].size();
    if (cur[c])
      RTG_HIP(hipMemcpyAsync(S.segs[c][0].ptr, init[c].data(), init[c].size() * sizeof(Seg), hipMemcpyHostToDevice, s));
  }
  if (cur[0]) {
    hipLaunchKernelGGL(k_fill_u32, dim3(div_up(n, TBS)), dim3(TBS), 0, s, S.sid[0].as<uint32_t>(), (size_t)n, NIL);
    hipLaunchKernelGGL(k_fill_seg_ids, dim3(cur[0]), dim3(TBS), 0, s, S.sid[0].as<uint32_t>(), S.segs[0][0].as<Seg>());
  }
  Counters hc{};
  hc.error = 0;
  RTG_HIP(hipMemcpyAsync(cnt, &hc, sizeof(hc), hipMemcpyHostToDevice, s));
  RTG_HIP(hipStreamSynchronize(s));  // the host staging above must outlive its copies

  Tree t{child, parent_int, parent_leaf, range, vals, cnt, S.roots.as<uint32_t>(), depth};
  S.side.ensure();
  bool tiny_pending[2] = {false, false};
  int par = 0;
  uint32_t levels = 0;
  // a failed build leaves no finisher running on the scratch it reuses
  auto fail = [&](const std::string& msg) {
    (void)hipStreamSynchronize(S.side.s);
    throw hip_error(hipErrorUnknown, msg);
  };
  while (cur[0] + cur[1] + cur[2] > 0) {
    if (++levels > 4096) fail("SAH build did not terminate");
    const int in = par, o = par ^ 1;
    // this level rewrites the segment lists of parity o, which the finisher
    // of two levels back may still be reading
    if (tiny_pending[o]) {
      RTG_HIP(hipStreamWaitEvent(s, S.side.done[o], 0));
      tiny_pending[o] = false;
    }
    Lists nx;
    for (int c = 0; c < 4; c++) {
      nx.seg[c] = S.segs[c][o].as<Seg>();
      nx.cap[c] = cap[c];
    }
    nx.fin_max = fin_max;
    hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(64), 0, s, cnt);
    if (cur[0]) {
      const Seg* bs = S.segs[0][in].as<Seg>();
      hipLaunchKernelGGL(k_init_gbins, dim3(div_up((size_t)cur[0] * BIN_WORDS, TBS)), dim3(TBS), 0, s,
                         S.bins.as<uint32_t>(), cur[0]);
      hipLaunchKernelGGL(k_fill_u32, dim3(div_up(n, TBS)), dim3(TBS), 0, s, S.sid[o].as<uint32_t>(), (size_t)n, NIL);
      hipLaunchKernelGGL(k_big_bins, dim3(div_up(n, CHUNK)), dim3(TBS), 0, s, S.sid[in].as<uint32_t>(),
                         S.blo[in].as<float4>(), S.bhi[in].as<float4>(), n, bs, S.bins.as<uint32_t>());
      hipLaunchKernelGGL(k_big_split, dim3(cur[0]), dim3(TBS), 0, s, bs, S.bins.as<uint32_t>(), S.splits.as<Split>(),
                         nx, t, lcost);
      hipLaunchKernelGGL(k_big_flags, dim3(div_up(n, TBS)), dim3(TBS), 0, s, S.sid[in].as<uint32_t>(),
                         S.blo[in].as<float4>(), S.bhi[in].as<float4>(), n, bs, S.splits.as<Split>(),
                         S.flags.as<uint32_t>());
      S.scanner.exclusive(S.flags.as<uint32_t>(), S.scan.as<uint32_t>(), n, s);
      hipLaunchKernelGGL(k_big_scatter, dim3(div_up(n, TBS)), dim3(TBS), 0, s, S.sid[in].as<uint32_t>(),
                         S.idx[in].as<uint32_t>(), S.blo[in].as<float4>(), S.bhi[in].as<float4>(), n, bs,
                         S.splits.as<Split>(), S.flags.as<uint32_t>(), S.scan.as<uint32_t>(), S.sid[o].as<uint32_t>(),
                         S.idx[o].as<uint32_t>(), S.blo[o].as<float4>(), S.bhi[o].as<float4>(), vals, cnt);
    }
    if (cur[1])
      hipLaunchKernelGGL((k_seg<256, 16>), dim3(cur[1]), dim3(256), 0, s, S.segs[1][in].as<Seg>(),
                         S.idx[in].as<uint32_t>(), S.blo[in].as<float4>(), S.bhi[in].as<float4>(),
                         S.idx[o].as<uint32_t>(), S.blo[o].as<float4>(), S.bhi[o].as<float4>(), nx, t, lcost);
    if (cur[2])
      hipLaunchKernelGGL((k_seg<64, 1>), dim3(cur[2]), dim3(64), 0, s, S.segs[2][in].as<Seg>(),
                         S.idx[in].as<uint32_t>(), S.blo[in].as<float4>(), S.bhi[in].as<float4>(),
                         S.idx[o].as<uint32_t>(), S.blo[o].as<float4>(), S.bhi[o].as<float4>(), nx, t, lcost);
    RTG_HIP(hipGetLastError());
    RTG_HIP(hipMemcpyAsync(&hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, s));
    RTG_HIP(hipStreamSynchronize(s));
    if (hc.error) fail("SAH build failed: error bits " + std::to_string(hc.error));
    for (int c = 0; c < 3; c++) cur[c] = hc.nseg[c];
    // this level's tiny children, finished in place on the side stream while
    // the next levels run (their positions are outside every segment the
    // next levels touch; this level's work is complete after the
    // synchronisation above)
    if (hc.nseg[3]) {
      hipLaunchKernelGGL(k_tiny, dim3(div_up(hc.nseg[3], TT)), dim3(TT), 0, S.side.s, S.segs[3][o].as<Seg>(),
                         hc.nseg[3], S.idx[o].as<uint32_t>(), S.blo[o].as<float4>(), S.bhi[o].as<float4>(), t);
      RTG_HIP(hipGetLastError());
      RTG_HIP(hipEventRecord(S.side.done[o], S.side.s));
      tiny_pending[o] = true;
    }
    par = o;
  }
  for (int p = 0; p < 2; p++)
    if (tiny_pending[p]) RTG_HIP(hipStreamWaitEvent(s, S.side.done[p], 0));
  out.roots.assign(trees.size(), NIL);
  RTG_HIP(hipMemcpyAsync(out.roots.data(), S.roots.ptr, trees.size() * 4, hipMemcpyDeviceToHost, s));
  RTG_HIP(hipMemcpyAsync(&hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, s));
  RTG_HIP(hipStreamSynchronize(s));
  if (hc.error) throw hip_error(hipErrorUnknown, "SAH build failed: error bits " + std::to_string(hc.error));
  for (uint32_t r : out.roots)
    if (r == NIL || r + 1 >= n) throw hip_error(hipErrorUnknown, "SAH build produced no root");
  out.levels = levels;
}

}  // namespace rtg
