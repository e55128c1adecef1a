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
// Segments (contiguous position ranges still to split) are split by size
// class; the class fixes how a segment is split, whatever kernel does it:
//   big   (> 4096): positions in chunks of 1024 bin into LDS, the chunk bins
//                   merge into per-segment bins in HBM with atomics, one
//                   workgroup per segment picks the split, and a global scan
//                   of left flags gives every position its stable destination;
//   mid   (513..4096): a 256-lane workgroup bins, picks and partitions in LDS;
//   small (17..512): one wave, same code;
//   tiny  (2..16): exact SAH splits over the references sorted by centroid
//                  on each axis, the whole subtree at once; a tree's own root
//                  of 2..512 references is split as small.
// Big, mid and small segments are split one level per step of the level
// loop, whose kernels read their segment counts on the device, so the host
// queues a level before it learns whether the previous one left work; tiny
// segments collect in a list that k_tiny finishes after the loop, four per
// wave.
// Every inner node gets the id of its split position (first + n_left - 1),
// which is unique, so the output does not depend on the order in which
// segments are processed.
#include <algorithm>
#include <cstring>
#include <string>

#include "builder_sah.h"
#include "mlp.h"

namespace rtg {
namespace {

constexpr uint32_t NIL = 0xffffffffu;
constexpr int NB = 32;                  // bins per axis
constexpr int BIN_WORDS = 3 * NB * 7;   // counts, then lo xyz, then hi xyz, per (axis, bin)
constexpr uint32_t MEDIAN_MAX = 64;      // largest median_below honoured
constexpr uint32_t TINY_MAX = 16;        // segments up to this size are finished by one thread (exact SAH)
constexpr uint32_t MID_MAX = 4096;
constexpr uint32_t SMALL_MAX = 512;     // segments up to this size are split by one wave
constexpr uint32_t CHUNK = 1024;        // positions per block of the big-segment binning
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

// A tiny segment waiting for k_tiny, with the parity of the buffers
// holding its references.
struct TinyItem {
  Seg s;
  uint32_t par;
};

struct Counters {
  uint32_t nseg[3][3];  // segment counts by list slot: big, mid, small (level L reads slot L % 3, fills (L + 1) % 3
                        // and clears (L + 2) % 3 for the level after)
  uint32_t ntiny;       // tiny segments so far (all levels)
  uint32_t error;    // bit 0: a segment list overflowed, bit 1: bin counts disagree, bit 2: destination out of range,
                     // bit 3: median subtree stack overflow, bit 4: a tiny segment out of range,
                     // bit 5: a depth beyond 254
};

struct Lists {
  Seg* seg[3];       // next level's big, mid and small segments
  uint32_t* nseg;    // and their counts
  uint32_t cap[3];
  TinyItem* tiny;    // tiny segments
  uint32_t tiny_cap;
  uint32_t par;      // parity of the buffers this level writes
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

// Bins the references at positions [p, p + cnt), consecutive in the
// Morton order: a run of references in one bin adds its count and box in
// one set of atomics.
__device__ inline void bin_run(uint32_t* w, const Mapping& m, const float4* __restrict__ blo,
                               const float4* __restrict__ bhi, uint32_t p, uint32_t cnt) {
  int cb[3] = {-1, -1, -1};
  uint32_t cn[3] = {0u, 0u, 0u}, ql[3][3], qh[3][3];
  auto flush = [&](int a) {
    if (cn[a] == 0u) return;
    atomicAdd(&w[wc(a, cb[a])], cn[a]);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      atomicMin(&w[wl(a, cb[a], k)], ql[a][k]);
      atomicMax(&w[wh(a, cb[a], k)], qh[a][k]);
    }
    cn[a] = 0u;
  };
  // loads in batches of four, issued before their use
  constexpr uint32_t BATCH = 4;
  for (uint32_t j0 = 0; j0 < cnt; j0 += BATCH) {
    float4 bl[BATCH], bh[BATCH];
#pragma unroll
    for (uint32_t u = 0; u < BATCH; u++)
      if (j0 + u < cnt) {
        bl[u] = blo[p + j0 + u];
        bh[u] = bhi[p + j0 + u];
      }
#pragma unroll
    for (uint32_t u = 0; u < BATCH; u++) {
      if (j0 + u >= cnt) break;
      const float4 l = bl[u], h = bh[u];
      const float c[3] = {l.x + h.x, l.y + h.y, l.z + h.z};
      const uint32_t ol[3] = {f2o(l.x), f2o(l.y), f2o(l.z)};
      const uint32_t oh[3] = {f2o(h.x), f2o(h.y), f2o(h.z)};
#pragma unroll
      for (int a = 0; a < 3; a++) {
        const int b = bin_of(c[a], m.ofs[a], m.scale[a], m.nb);
        if (b != cb[a] || cn[a] == 0u) {
          flush(a);
          cb[a] = b;
#pragma unroll
          for (int k = 0; k < 3; k++) {
            ql[a][k] = ol[k];
            qh[a][k] = oh[k];
          }
        } else {
#pragma unroll
          for (int k = 0; k < 3; k++) {
            ql[a][k] = ol[k] < ql[a][k] ? ol[k] : ql[a][k];
            qh[a][k] = oh[k] > qh[a][k] ? oh[k] : qh[a][k];
          }
        }
        cn[a]++;
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 3; a++) flush(a);
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

// Child segment [first, first+count) of inner node m at depth+1: a leaf
// link for a single primitive and a finished median subtree up to
// nx.fin_max primitives are written here (result -1); any other child
// goes to the list of its size class, which is returned (0 big, 1 mid,
// 2 small, 3 tiny).
constexpr int LIST_TINY = 3;

__device__ inline int child_list(const Tree& t, const Lists& nx, uint32_t m, uint32_t side, uint32_t depth,
                                 uint32_t first, uint32_t count) {
  if (count == 1) {
    t.child[2 * m + side] = first | REF_LEAF;
    t.parent_leaf[first] = m;
    return -1;
  }
  if (count <= nx.fin_max) {
    finish_median(t, m, side, depth, first, count);
    return -1;
  }
  return count <= TINY_MAX ? LIST_TINY : (count > MID_MAX ? 0 : (count > SMALL_MAX ? 1 : 2));
}

__device__ inline uint32_t* list_counter(const Tree& t, const Lists& nx, int cls) {
  return cls == LIST_TINY ? &t.cnt->ntiny : &nx.nseg[cls];
}

// Writes the child into slot id of list cls (error bit 0 past its capacity).
__device__ inline void put_child(const Tree& t, const Lists& nx, int cls, uint32_t id, uint32_t m, uint32_t side,
                                 uint32_t depth, uint32_t first, uint32_t count, const float clo[3],
                                 const float chi[3]) {
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
  if (id >= (cls == LIST_TINY ? nx.tiny_cap : nx.cap[cls])) {
    atomicOr(&t.cnt->error, 1u);
    return;
  }
  if (cls == LIST_TINY) nx.tiny[id] = TinyItem{c, nx.par};
  else nx.seg[cls][id] = c;
}

// child_list and put_child with a slot of its own; returns the id in the
// big list, or NIL.
__device__ inline uint32_t emit_child(const Tree& t, const Lists& nx, uint32_t m, uint32_t side, uint32_t depth,
                                      uint32_t first, uint32_t count, const float clo[3], const float chi[3]) {
  const int cls = child_list(t, nx, m, side, depth, first, count);
  if (cls < 0) return NIL;
  const uint32_t id = atomicAdd(list_counter(t, nx, cls), 1u);
  put_child(t, nx, cls, id, m, side, depth, first, count, clo, chi);
  return cls == 0 && id < nx.cap[0] ? id : NIL;
}

// Synchronisation of the threads that share a segment's LDS: the whole
// workgroup, or one wave. A wave's LDS instructions execute in order, so
// between its lanes it takes only a wait for its LDS operations
// (lgkmcnt(0); vmcnt and expcnt left at their maxima) and a compiler fence
// that keeps loads and stores on their side; it does not wait for the
// wave's global stores.
template <bool WAVE>
__device__ inline void group_sync() {
  if constexpr (WAVE) {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
  } else {
    __syncthreads();
  }
}


// Moves within the 16-lane rows of a wave (DPP): lane i of a row gets
// lane i - D's value (ROW_SHR + D), lane i + D's (ROW_SHL + D) or lane
// (i - D) mod 16's (ROW_ROR + D); a lane without a source in its row keeps
// its own value. Called with the whole wave active.
constexpr int ROW_SHL = 0x100, ROW_SHR = 0x110, ROW_ROR = 0x120;
template <int CTRL>
__device__ inline uint32_t row_move(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ inline float row_move(float v) {
  return __uint_as_float(row_move<CTRL>(__float_as_uint(v)));
}
template <int CTRL>
__device__ inline int row_move(int v) {
  return (int)row_move<CTRL>((uint32_t)v);
}

// Plain binned-SAH sweep of axis a by one half-wave (32 lanes; lane b =
// lane & 31 holds bin b): the least A_L*n_L + A_R*n_R over the "bin < k"
// candidates, first minimum in k, with the float operations of a
// left-to-right scan (prefix and suffix boxes are exact minima and maxima,
// so the costs are the scan's bit for bit). cost is INFINITY when no k
// separates the references; every lane of the half returns the result and
// the axis's reference count. Called by a whole wave (both halves); the
// scans move values within 16-lane rows (DPP) and across the half's two
// rows by lane reads.
__device__ inline float read_lane(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
__device__ inline uint32_t read_lane(uint32_t v, int l) { return (uint32_t)__builtin_amdgcn_readlane((int)v, l); }
__device__ inline int read_lane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

template <int D>
__device__ inline void sweep_scan_step(int r, uint32_t& nl, float llo[3], float lhi[3], uint32_t& nr, float rlo[3],
                                       float rhi[3]) {
  const uint32_t un = row_move<ROW_SHR + D>(nl), dn = row_move<ROW_SHL + D>(nr);
  float ul[3], uh[3], dl[3], dh[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    ul[k] = row_move<ROW_SHR + D>(llo[k]);
    uh[k] = row_move<ROW_SHR + D>(lhi[k]);
    dl[k] = row_move<ROW_SHL + D>(rlo[k]);
    dh[k] = row_move<ROW_SHL + D>(rhi[k]);
  }
  if (r >= D) {
    nl += un;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      llo[k] = fminf(llo[k], ul[k]);
      lhi[k] = fmaxf(lhi[k], uh[k]);
    }
  }
  if (r + D < 16) {
    nr += dn;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      rlo[k] = fminf(rlo[k], dl[k]);
      rhi[k] = fmaxf(rhi[k], dh[k]);
    }
  }
}

template <int D>
__device__ inline void sweep_min_step(int r, float& cst, int& kk, uint32_t& nlv) {
  const float oc = row_move<ROW_SHR + D>(cst);
  const int ok = row_move<ROW_SHR + D>(kk);
  const uint32_t on = row_move<ROW_SHR + D>(nlv);
  if (r >= D && (oc < cst || (oc == cst && ok < kk))) {
    cst = oc;
    kk = ok;
    nlv = on;
  }
}

__device__ inline void sweep_axis_half(const uint32_t* w, const Mapping& m, int a, int lane, float& cost, int& kbest,
                                       uint32_t& nleft, uint32_t& total) {
  const int b = lane & 31, r = lane & 15;
  const bool upper = lane >= 32;
  const uint32_t c = b < m.nb ? w[wc(a, b)] : 0u;
  float llo[3], lhi[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    llo[k] = c ? o2f(w[wl(a, b, k)]) : INFINITY;
    lhi[k] = c ? o2f(w[wh(a, b, k)]) : -INFINITY;
  }
  float rlo[3] = {llo[0], llo[1], llo[2]}, rhi[3] = {lhi[0], lhi[1], lhi[2]};
  uint32_t nl = c, nr = c;
  // prefix over the row's bins up to b, suffix from b
  sweep_scan_step<1>(r, nl, llo, lhi, nr, rlo, rhi);
  sweep_scan_step<2>(r, nl, llo, lhi, nr, rlo, rhi);
  sweep_scan_step<4>(r, nl, llo, lhi, nr, rlo, rhi);
  sweep_scan_step<8>(r, nl, llo, lhi, nr, rlo, rhi);
  // the half's second row adds its first row's total to the prefix, the
  // first row the second row's to the suffix
  const uint32_t pn = upper ? read_lane(nl, 47) : read_lane(nl, 15);
  const uint32_t sn = upper ? read_lane(nr, 48) : read_lane(nr, 16);
  float pbl[3], pbh[3], sbl[3], sbh[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    pbl[k] = upper ? read_lane(llo[k], 47) : read_lane(llo[k], 15);
    pbh[k] = upper ? read_lane(lhi[k], 47) : read_lane(lhi[k], 15);
    sbl[k] = upper ? read_lane(rlo[k], 48) : read_lane(rlo[k], 16);
    sbh[k] = upper ? read_lane(rhi[k], 48) : read_lane(rhi[k], 16);
  }
  if (b >= 16) {
    nl += pn;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      llo[k] = fminf(llo[k], pbl[k]);
      lhi[k] = fmaxf(lhi[k], pbh[k]);
    }
  } else {
    nr += sn;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      rlo[k] = fminf(rlo[k], sbl[k]);
      rhi[k] = fmaxf(rhi[k], sbh[k]);
    }
  }
  // candidate k = b + 1: bins [0, b] left, the suffix from bin b + 1 right
  // (for b = 15 that is the second row's first suffix, read above)
  uint32_t nr1 = row_move<ROW_SHL + 1>(nr);
  float r1l[3], r1h[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    r1l[k] = row_move<ROW_SHL + 1>(rlo[k]);
    r1h[k] = row_move<ROW_SHL + 1>(rhi[k]);
  }
  if (b == 15) {
    nr1 = sn;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      r1l[k] = sbl[k];
      r1h[k] = sbh[k];
    }
  }
  float cst = INFINITY;
  if (m.scale[a] != 0.0f && b + 1 < m.nb && nl != 0u && nr1 != 0u)
    cst = half_area(lhi[0] - llo[0], lhi[1] - llo[1], lhi[2] - llo[2]) * (float)nl +
          half_area(r1h[0] - r1l[0], r1h[1] - r1l[1], r1h[2] - r1l[2]) * (float)nr1;
  int kk = b + 1;
  uint32_t nlv = nl;
  total = upper ? read_lane(nl, 63) : read_lane(nl, 31);
  // least (cost, k) of each row reaches its last lane, then the two rows'
  sweep_min_step<1>(r, cst, kk, nlv);
  sweep_min_step<2>(r, cst, kk, nlv);
  sweep_min_step<4>(r, cst, kk, nlv);
  sweep_min_step<8>(r, cst, kk, nlv);
  const float c0 = upper ? read_lane(cst, 47) : read_lane(cst, 15);
  const int k0 = upper ? read_lane(kk, 47) : read_lane(kk, 15);
  const uint32_t n0 = upper ? read_lane(nlv, 47) : read_lane(nlv, 15);
  const float c1 = upper ? read_lane(cst, 63) : read_lane(cst, 31);
  const int k1 = upper ? read_lane(kk, 63) : read_lane(kk, 31);
  const uint32_t n1 = upper ? read_lane(nlv, 63) : read_lane(nlv, 31);
  const bool second = c1 < c0 || (c1 == c0 && k1 < k0);
  cost = second ? c1 : c0;
  kbest = second ? k1 : k0;
  nleft = second ? n1 : n0;
}

// Workgroup-level choice of the split of segment sg from its bins w (LDS or
// HBM); every thread returns the same Split. Plain SAH: one half-wave per
// axis sweeps its bins. Learned cost (lc.P set and the segment larger than
// lc.min_count): the workgroup evaluates the model on the left and right
// sides of all 3*(nb-1) candidates (each side's box and count from the
// bins, against the segment's), then one thread per axis keeps the least
// P(L)*n_L + P(R)*n_R with P = halfArea * 2^model (bounds below) among the
// candidates whose plain cost A(L)*n_L + A(R)*n_R is within
// (1 + lc.sah_slack) of the least over all axes.
// Ties keep the lowest axis and the lowest k; without a separating
// candidate the split is the object median.
struct ChooseLds {
  float slo[3], shi[3];  // learned: the segment's box
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

template <int BLOCK, bool WAVE, bool LEARNED>
__device__ inline Split choose_split(const uint32_t* w, const Mapping& m, const Seg& sg, const LearnedCost& lc,
                                     ChooseLds& L, uint32_t* err, int tid) {
  const uint32_t depth = sg.meta >> 1;
  const bool learned = LEARNED && lc.P != nullptr && sg.count > lc.min_count && depth < MEDIAN_DEPTH;
  const int ncand = m.nb - 1;
  if (learned) {
    // the segment's box and count (the union of one axis's bins)
    if (tid == 0) {
      uint32_t cnt;
      range_box(w, 0, 0, m.nb, L.slo, L.shi, cnt);
    }
    group_sync<WAVE>();
    const float slo[3] = {L.slo[0], L.slo[1], L.slo[2]}, shi[3] = {L.shi[0], L.shi[1], L.shi[2]};
    // every candidate's two sides: the per-bin statistics (box and count of
    // the side's bins, against the segment's) go through the model, which
    // corrects the SAH's entry-probability estimate A_C / A_S by 2^y; a
    // prediction is held within [MLP_Y_MIN, MLP_Y_MAX] and never makes a
    // child more likely to be entered than its parent
    for (int e = tid; e < 2 * 3 * ncand; e += BLOCK) {
      const int c = e >> 1, side = e & 1;
      const int a = c / ncand, k = c % ncand + 1;
      float lo[3], hi[3];
      uint32_t cnt;
      range_box(w, a, side ? k : 0, side ? m.nb : k, lo, hi, cnt);
      float pv = -1.0f, av = -1.0f;
      if (cnt > 0 && m.scale[a] != 0.0f) {
        float f[MLP_IN];
        split_features(lo, hi, cnt, slo, shi, sg.count, lc.sn, f);
        const float ymax = fminf(MLP_Y_MAX, -16.0f * f[0]);
        const float y = fminf(fmaxf(mlp_eval(lc.P, f), MLP_Y_MIN), ymax);
        av = half_area(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]);
        pv = av * exp2f(y);
      }
      L.p[e] = pv;
      L.a[e] = av;
    }
    group_sync<WAVE>();
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
    group_sync<WAVE>();
  }
  if (!learned && tid < 64) {
    // one half-wave per axis: x and y, then z (the upper half sweeps z as
    // well, for the whole-wave moves, and keeps nothing)
    const int half = tid >> 5, b = tid & 31;
    for (int pass = 0; pass < 2; pass++) {
      const int a = pass == 0 ? half : 2;
      float c;
      int k;
      uint32_t nl, total;
      sweep_axis_half(w, m, a, tid, c, k, nl, total);
      if (b == 0 && !(pass == 1 && half == 1)) {
        if (total != sg.count) atomicOr(err, 2u);
        L.cost[a] = depth < MEDIAN_DEPTH && total == sg.count ? c : INFINITY;
        L.k[a] = k;
        L.nl[a] = nl;
      }
    }
  }
  if (learned && tid < 3) {
    float c = INFINITY;
    int k = 0;
    uint32_t nl = 0;
    uint32_t total = 0;
    for (int b = 0; b < m.nb; b++) total += w[wc(tid, b)];
    if (total != sg.count) atomicOr(err, 2u);
    if (depth < MEDIAN_DEPTH && total == sg.count) {
      if (m.scale[tid] != 0.0f) {
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
  group_sync<WAVE>();
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
  group_sync<WAVE>();
  return L.sp;
}

// ------------------------------------------------------------ big segments --

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  const size_t i = (size_t)blockIdx.x * TBS + threadIdx.x;
  if (i < n) p[i] = v;
}

// Bins of the level's big segments; the grid covers the list's capacity
// and the count is read on the device.
__global__ void k_init_gbins(uint32_t* g, const uint32_t* nseg) {
  const size_t i = (size_t)blockIdx.x * TBS + threadIdx.x;
  if (i >= (size_t)*nseg * BIN_WORDS) return;
  const int w = (int)(i % BIN_WORDS);
  g[i] = w < 3 * NB ? 0u : (w < 3 * NB + 9 * NB ? NIL : 0u);
}


// A chunk of CHUNK positions meets at most two big segments (each is longer
// than the chunk): they are binned in LDS and merged into HBM; any other
// segment id (not expected) bins straight into HBM.
__global__ __launch_bounds__(TBS) void k_big_bins(const uint32_t* __restrict__ sid, const float4* __restrict__ blo,
                                                  const float4* __restrict__ bhi, uint32_t n,
                                                  const Seg* __restrict__ segs, uint32_t* gbins) {
  // one set of bins per wave and segment slot (fewer lanes per LDS atomic),
  // merged before the merge into HBM
  constexpr int NW = TBS / 64;
  __shared__ uint32_t sb[2][NW][BIN_WORDS];
  __shared__ uint32_t s_lo, s_hi;
  __shared__ Mapping sm[2];
  const int tid = threadIdx.x, wave = tid >> 6;
  const uint32_t c0 = blockIdx.x * CHUNK;
  const uint32_t c1 = c0 + CHUNK < n ? c0 + CHUNK : n;
  if (tid == 0) {
    s_lo = NIL;
    s_hi = 0u;
  }
  for (int w = 0; w < NW; w++) {
    init_bins(sb[0][w], tid, TBS);
    init_bins(sb[1][w], tid, TBS);
  }
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
  // each thread bins a run of RUN consecutive positions; neighbours in the
  // Morton order mostly share their bins, so a thread sums each axis's run
  // of equal bins in registers and adds the sum to the bins when the bin
  // (or the segment) changes: few LDS atomics, on few shared addresses
  constexpr int RUN = (int)(CHUNK / TBS);
  const uint32_t r0 = c0 + (uint32_t)tid * RUN;
  uint32_t rs[RUN];
  float4 rl[RUN], rh[RUN];
#pragma unroll
  for (int j = 0; j < RUN; j++) {
    const uint32_t i = r0 + j;
    rs[j] = i < c1 ? sid[i] : NIL;
    if (rs[j] != NIL) {
      rl[j] = blo[i];
      rh[j] = bhi[i];
    }
  }
  int cslot = -1;
  int cb[3] = {-1, -1, -1};
  uint32_t cn[3] = {0u, 0u, 0u}, ql[3][3], qh[3][3];
  auto flush = [&](int a) {
    if (cn[a] == 0u) return;
    uint32_t* w = sb[cslot][wave];
    atomicAdd(&w[wc(a, cb[a])], cn[a]);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      atomicMin(&w[wl(a, cb[a], k)], ql[a][k]);
      atomicMax(&w[wh(a, cb[a], k)], qh[a][k]);
    }
    cn[a] = 0u;
  };
#pragma unroll
  for (int j = 0; j < RUN; j++) {
    const uint32_t s = rs[j];
    if (s == NIL) continue;
    const float4 l = rl[j], h = rh[j];
    const int slot = s == slo_id ? 0 : (s == shi_id ? 1 : 2);
    if (slot == 2) {
      bin_prim(gbins + (size_t)s * BIN_WORDS, make_mapping(segs[s]), l, h);
      continue;
    }
    if (slot != cslot) {
#pragma unroll
      for (int a = 0; a < 3; a++) flush(a);
      cslot = slot;
    }
    const Mapping& m = sm[slot];
    const float c[3] = {l.x + h.x, l.y + h.y, l.z + h.z};
    const uint32_t ol[3] = {f2o(l.x), f2o(l.y), f2o(l.z)};
    const uint32_t oh[3] = {f2o(h.x), f2o(h.y), f2o(h.z)};
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const int b = bin_of(c[a], m.ofs[a], m.scale[a], m.nb);
      if (b != cb[a] || cn[a] == 0u) {
        flush(a);
        cb[a] = b;
#pragma unroll
        for (int k = 0; k < 3; k++) {
          ql[a][k] = ol[k];
          qh[a][k] = oh[k];
        }
      } else {
#pragma unroll
        for (int k = 0; k < 3; k++) {
          ql[a][k] = ol[k] < ql[a][k] ? ol[k] : ql[a][k];
          qh[a][k] = oh[k] > qh[a][k] ? oh[k] : qh[a][k];
        }
      }
      cn[a]++;
    }
  }
  if (cslot >= 0) {
#pragma unroll
    for (int a = 0; a < 3; a++) flush(a);
  }
  __syncthreads();
  for (int slot = 0; slot < 2; slot++) {
    if (slot == 1 && shi_id == slo_id) break;
    uint32_t* g = gbins + (size_t)(slot ? shi_id : slo_id) * BIN_WORDS;
    for (int i = tid; i < BIN_WORDS; i += TBS) {
      // the waves' words: counts add, ordered bounds take min / max
      const int kind = i < 3 * NB ? 0 : (i < 3 * NB + 9 * NB ? 1 : 2);
      const int cw = kind == 0 ? i : (kind == 1 ? (i - 3 * NB) / 3 : (i - 3 * NB - 9 * NB) / 3);
      uint32_t v = sb[slot][0][i], c = sb[slot][0][cw];
      for (int w = 1; w < NW; w++) {
        const uint32_t x = sb[slot][w][i];
        c += sb[slot][w][cw];
        v = kind == 0 ? v + x : (kind == 1 ? (x < v ? x : v) : (x > v ? x : v));
      }
      if (kind == 0) {
        if (v) atomicAdd(&g[i], v);
      } else if (c) {
        if (kind == 1) atomicMin(&g[i], v); else atomicMax(&g[i], v);
      }
    }
  }
}

template <bool LEARNED>
__global__ __launch_bounds__(TBS, 4) void k_big_split(const Seg* __restrict__ segs, const uint32_t* __restrict__ nseg,
                                                   const uint32_t* __restrict__ gbins, Split* __restrict__ splits,
                                                   Lists nx, Tree t, LearnedCost lc) {
  __shared__ ChooseLds L;
  __shared__ uint32_t w[BIN_WORDS];  // the segment's bins, read once from HBM
  const uint32_t i = blockIdx.x;
  if (i >= *nseg) return;
  const Seg sg = segs[i];
  const uint32_t depth = sg.meta >> 1;
  const Mapping m = make_mapping(sg);
  for (int k = threadIdx.x; k < BIN_WORDS; k += TBS) w[k] = gbins[(size_t)i * BIN_WORDS + k];
  __syncthreads();
  Split sp = choose_split<TBS, false, LEARNED>(w, m, sg, lc, L, &t.cnt->error, threadIdx.x);
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
  if (s == NIL) {  // not in a big segment: the position keeps its place
    sid_o[i] = NIL;
    return;
  }
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

// LDS of one segment split by a group (a workgroup, or one wave).
template <int BLOCK>
struct SegLds {
  uint32_t sb[BIN_WORDS];
  ChooseLds L;
  uint32_t wsum[BLOCK / 64];
  uint32_t cb[12];  // ordered bounds of lo+hi: left lo3, left hi3, right lo3, right hi3
};

// Bins, splits and stably partitions segment sg with the BLOCK threads of a
// group (tid 0..BLOCK-1, each a run of per = ceil(count / BLOCK) <= PER
// consecutive positions, so that small segments spread over the group's
// lanes): references read from
// (idx, blo, bhi), written to (idx_o, blo_o, bhi_o) and their ids to
// t.vals at the same positions; links the inner node (the split position)
// into the tree. Every thread returns the split and, in lcl..rch, the
// centroid bounds of the two children.
template <int BLOCK, int PER, bool WAVE, bool LEARNED>
__device__ inline Split split_segment(const Seg& sg, const uint32_t* __restrict__ idx, const float4* __restrict__ blo,
                                      const float4* __restrict__ bhi, uint32_t* __restrict__ idx_o,
                                      float4* __restrict__ blo_o, float4* __restrict__ bhi_o, const Tree& t,
                                      const LearnedCost& lc, SegLds<BLOCK>& S, int tid, float lcl[3], float lch[3],
                                      float rcl[3], float rch[3]) {
  static_assert(PER <= 32, "flags are kept in one 32-bit mask");
  const int lane = tid & 63, wave = tid >> 6;
  const Mapping m = make_mapping(sg);
  init_bins(S.sb, tid, BLOCK);
  if (tid < 12) S.cb[tid] = (tid % 6) < 3 ? NIL : 0u;
  group_sync<WAVE>();
  const uint32_t end = sg.first + sg.count;
  const int per = PER == 1 ? 1 : (int)((sg.count + BLOCK - 1) / BLOCK);
  const uint32_t p0 = sg.first + (uint32_t)tid * (uint32_t)per;
  // One position per thread (a segment of at most BLOCK references): its
  // reference is loaded once and kept in registers for the binning, the
  // flag and the scatter, instead of three loads from global memory.
  const bool single = PER == 1 || per == 1;
  const bool have = p0 < end;
  float4 l1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), h1 = l1;
  uint32_t v1 = 0u;
  if (single) {
    if (have) {
      l1 = blo[p0];
      h1 = bhi[p0];
      v1 = idx[p0];
      bin_prim(S.sb, m, l1, h1);
    }
  } else if (have) {
    bin_run(S.sb, m, blo, bhi, p0, min((uint32_t)per, end - p0));
  }
  group_sync<WAVE>();
  const Split sp = choose_split<BLOCK, WAVE, LEARNED>(S.sb, m, sg, lc, S.L, &t.cnt->error, tid);
  // flags of this thread's positions, and the stable destinations
  uint32_t mask = 0u, nl = 0u;
  constexpr int BATCH = PER < 4 ? PER : 4;
  if (single) {
    if (have && goes_left(sp, l1, h1, p0 - sg.first)) {
      mask = 1u;
      nl = 1u;
    }
  } else for (int j0 = 0; j0 < per; j0 += BATCH) {
    float4 bl[BATCH], bh[BATCH];
#pragma unroll
    for (int u = 0; u < BATCH; u++) {
      const uint32_t p = p0 + j0 + u;
      if (j0 + u < per && p < end) {
        bl[u] = blo[p];
        bh[u] = bhi[p];
      }
    }
#pragma unroll
    for (int u = 0; u < BATCH; u++) {
      const uint32_t p = p0 + j0 + u;
      if (j0 + u < per && p < end && goes_left(sp, bl[u], bh[u], p - sg.first)) {
        mask |= 1u << (j0 + u);
        nl++;
      }
    }
  }
  // the wave's exclusive prefix of nl (at most PER) and its total, bit by
  // bit over ballots
  constexpr int NL_BITS = 32 - __builtin_clz((unsigned)PER);
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  uint32_t x = 0u, wtotal = 0u;
#pragma unroll
  for (int b = 0; b < NL_BITS; b++) {
    const uint64_t m = __ballot((nl >> b) & 1u);
    x += (uint32_t)__popcll(m & below) << b;
    wtotal += (uint32_t)__popcll(m) << b;
  }
  if (lane == 0) S.wsum[wave] = wtotal;
  group_sync<WAVE>();
  uint32_t lbase = x;
  for (int w = 0; w < wave; w++) lbase += S.wsum[w];
  // centroid bounds of the left (lb) and right (rb) references: lo3, hi3
  float lb[6], rb[6];
#pragma unroll
  for (int k = 0; k < 6; k++) lb[k] = rb[k] = k < 3 ? INFINITY : -INFINITY;
  uint32_t lr = 0u;
  // the reference at position p (rel = p - sg.first) to its stable
  // destination; centroid bounds of the side it joins
  auto place = [&](uint32_t rel, bool left, float4 l, float4 h, uint32_t v) {
    const uint32_t dst = left ? sg.first + lbase + lr : sg.first + sp.nleft + (rel - lbase - lr);
    if (left) lr++;
    if (dst >= end) {
      atomicOr(&t.cnt->error, 4u);
      return;
    }
    idx_o[dst] = v;
    blo_o[dst] = l;
    bhi_o[dst] = h;
    t.vals[dst] = v;
    const float c[3] = {l.x + h.x, l.y + h.y, l.z + h.z};
    if (left) {
#pragma unroll
      for (int k = 0; k < 3; k++) {
        lb[k] = fminf(lb[k], c[k]);
        lb[3 + k] = fmaxf(lb[3 + k], c[k]);
      }
    } else {
#pragma unroll
      for (int k = 0; k < 3; k++) {
        rb[k] = fminf(rb[k], c[k]);
        rb[3 + k] = fmaxf(rb[3 + k], c[k]);
      }
    }
  };
  if (single) {
    if (have) place(p0 - sg.first, mask & 1u, l1, h1, v1);
  } else for (int j0 = 0; j0 < per; j0 += BATCH) {
    float4 bl[BATCH], bh[BATCH];
    uint32_t bv[BATCH];
#pragma unroll
    for (int u = 0; u < BATCH; u++) {  // loads in batches, issued before their use
      const uint32_t p = p0 + j0 + u;
      if (j0 + u < per && p < end) {
        bl[u] = blo[p];
        bh[u] = bhi[p];
        bv[u] = idx[p];
      }
    }
#pragma unroll
    for (int u = 0; u < BATCH; u++) {
      const int j = j0 + u;
      const uint32_t p = p0 + j;
      if (j >= per || p >= end) break;
      place(p - sg.first, (mask >> j) & 1u, bl[u], bh[u], bv[u]);
    }
  }
#pragma unroll
  for (int k = 0; k < 12; k++) {
    // the least (greatest) of each 16-lane row reaches the row's last lane
    float v = k < 6 ? lb[k] : rb[k - 6];
    const bool mx = (k % 6) >= 3;
    float y = row_move<ROW_SHR + 1>(v);
    v = mx ? fmaxf(v, y) : fminf(v, y);
    y = row_move<ROW_SHR + 2>(v);
    v = mx ? fmaxf(v, y) : fminf(v, y);
    y = row_move<ROW_SHR + 4>(v);
    v = mx ? fmaxf(v, y) : fminf(v, y);
    y = row_move<ROW_SHR + 8>(v);
    v = mx ? fmaxf(v, y) : fminf(v, y);
    if ((lane & 15) == 15 && !__builtin_isinf(v)) {
      if (mx) atomicMax(&S.cb[k], f2o(v)); else atomicMin(&S.cb[k], f2o(v));
    }
  }
  group_sync<WAVE>();
  if (tid == 0) link_inner(t, sg, sg.first + sp.nleft - 1);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    lcl[k] = o2f(S.cb[k]);
    lch[k] = o2f(S.cb[3 + k]);
    rcl[k] = o2f(S.cb[6 + k]);
    rch[k] = o2f(S.cb[9 + k]);
  }
  return sp;
}

// One workgroup per mid segment of a level (between SMALL_MAX and MID_MAX
// references).
template <bool LEARNED>
__global__ __launch_bounds__(256, 4) void k_seg(const Seg* __restrict__ segs, const uint32_t* __restrict__ nseg,
                                             const uint32_t* __restrict__ idx, const float4* __restrict__ blo,
                                             const float4* __restrict__ bhi, uint32_t* __restrict__ idx_o,
                                             float4* __restrict__ blo_o, float4* __restrict__ bhi_o, Lists nx,
                                             Tree t, LearnedCost lc, uint32_t* __restrict__ clear) {
  __shared__ SegLds<256> S;
  // the counts the next level fills (no kernel of this level uses them)
  if (blockIdx.x == 0 && threadIdx.x < 3) clear[threadIdx.x] = 0u;
  if (blockIdx.x >= *nseg) return;
  const Seg sg = segs[blockIdx.x];
  float lcl[3], lch[3], rcl[3], rch[3];
  const Split sp = split_segment<256, 16, false, LEARNED>(sg, idx, blo, bhi, idx_o, blo_o, bhi_o, t, lc, S, threadIdx.x, lcl,
                                                 lch, rcl, rch);
  if (threadIdx.x == 0) {
    const uint32_t mid = sg.first + sp.nleft - 1, depth = sg.meta >> 1;
    emit_child(t, nx, mid, 0, depth, sg.first, sp.nleft, lcl, lch);
    emit_child(t, nx, mid, 1, depth, sg.first + sp.nleft, sg.count - sp.nleft, rcl, rch);
  }
}

// One wave per small segment of a level (up to SMALL_MAX references, or a
// tree's root of any size up to that), SEGW_WAVES per workgroup; the grid
// covers the list's capacity and the count is read on the device. The
// workgroup reserves its children's list slots with one global atomic per
// list: the level's segments number in the tens of thousands, and an atomic
// per child on one address would serialise them. (A grid-stride loop over
// the count doubles the kernel's registers.)
constexpr int SEGW_TB = 512;
constexpr int SEGW_WAVES = SEGW_TB / 64;
constexpr int SEGW_PER = (int)(SMALL_MAX / 64);

template <bool LEARNED>
__global__ __launch_bounds__(SEGW_TB) void k_seg_wave(const Seg* __restrict__ segs, const uint32_t* __restrict__ nseg,
                                                      const uint32_t* __restrict__ idx, const float4* __restrict__ blo,
                                                      const float4* __restrict__ bhi, uint32_t* __restrict__ idx_o,
                                                      float4* __restrict__ blo_o, float4* __restrict__ bhi_o, Lists nx,
                                                      Tree t, LearnedCost lc) {
  __shared__ SegLds<64> S[SEGW_WAVES];
  __shared__ uint32_t lcount[4], lbase[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t i = blockIdx.x * SEGW_WAVES + (uint32_t)wave;
  if (threadIdx.x < 4) lcount[threadIdx.x] = 0u;
  __syncthreads();
  const bool active = i < *nseg && i < nx.cap[2];  // both parities' lists have this capacity
  Seg sg;
  Split sp;
  float cb[2][6];  // children's centroid bounds: lo3, hi3
  int cls[2] = {-1, -1};
  uint32_t slot[2] = {0u, 0u};
  if (active) {
    sg = segs[i];
    sp = split_segment<64, SEGW_PER, true, LEARNED>(sg, idx, blo, bhi, idx_o, blo_o, bhi_o, t, lc, S[wave], lane,
                                                    &cb[0][0], &cb[0][3], &cb[1][0], &cb[1][3]);
    if (lane == 0) {
      const uint32_t mid = sg.first + sp.nleft - 1, depth = sg.meta >> 1;
      cls[0] = child_list(t, nx, mid, 0, depth, sg.first, sp.nleft);
      cls[1] = child_list(t, nx, mid, 1, depth, sg.first + sp.nleft, sg.count - sp.nleft);
#pragma unroll
      for (int c = 0; c < 2; c++)
        if (cls[c] >= 0) slot[c] = atomicAdd(&lcount[cls[c]], 1u);
    }
  }
  __syncthreads();
  if (threadIdx.x < 4 && lcount[threadIdx.x])
    lbase[threadIdx.x] = atomicAdd(list_counter(t, nx, (int)threadIdx.x), lcount[threadIdx.x]);
  __syncthreads();
  if (active && lane == 0) {
    const uint32_t mid = sg.first + sp.nleft - 1, depth = sg.meta >> 1;
#pragma unroll
    for (int c = 0; c < 2; c++)
      if (cls[c] >= 0)
        put_child(t, nx, cls[c], lbase[cls[c]] + slot[c], mid, (uint32_t)c, depth, c ? sg.first + sp.nleft : sg.first,
                  c ? sg.count - sp.nleft : sp.nleft, &cb[c][0], &cb[c][3]);
  }
}

// ------------------------------------------------------------ tiny segments --

static_assert(SMALL_MAX <= 64 * 32, "a small segment is split by one wave of at most 32 positions per lane");
static_assert(MID_MAX <= 256 * 16, "a mid segment is split by 256 threads of at most 16 positions each");

struct Bufs {
  uint32_t* idx[2];
  float4* blo[2];
  float4* bhi[2];
};

struct TinyLds {
  float4 lo[TINY_MAX], hi[TINY_MAX];
  uint32_t id[TINY_MAX];
};

constexpr int TINY_GROUPS = 64 / (int)TINY_MAX;
static_assert(TINY_GROUPS * TINY_MAX == 64, "tiny segments take 16-lane groups of a wave");

// Stable ranks of a lane's keys among its node's lanes [nf, nf + nc) of
// the row: r0 by (k0, position), r1 by (k1, k0, position), r2 by (k2, k1,
// k0, position); f0 and f1 count the keys below on axis 0 and 1 with ties
// broken by r2's order. The keys of lane (pos - R) mod 16 arrive by
// rotating the row one lane per step (a chain, which keeps few registers
// live); the lane itself counts for nothing.
__device__ inline void tiny_ranks(float k0, float k1, float k2, uint32_t pos, uint32_t nf, uint32_t nc,
                                  uint32_t& r0, uint32_t& r1, uint32_t& r2, uint32_t& f0, uint32_t& f1) {
  float a0 = k0, a1 = k1, a2 = k2;
#pragma unroll 1
  for (uint32_t R = 1; R < TINY_MAX; R++) {
    a0 = row_move<ROW_ROR + 1>(a0);
    a1 = row_move<ROW_ROR + 1>(a1);
    a2 = row_move<ROW_ROR + 1>(a2);
    const uint32_t j = (pos - R) & (TINY_MAX - 1u);
    // branch-free: bitwise operators on the comparisons
    const bool in = j - nf < nc;
    const bool lt0 = (a0 < k0) | ((a0 == k0) & (j < pos));
    const bool lt1 = (a1 < k1) | ((a1 == k1) & lt0);
    const bool lt2 = (a2 < k2) | ((a2 == k2) & lt1);
    r0 += (uint32_t)(in & lt0);
    r1 += (uint32_t)(in & lt1);
    r2 += (uint32_t)(in & lt2);
    f0 += (uint32_t)(in & ((a0 < k0) | ((a0 == k0) & lt2)));
    f1 += (uint32_t)(in & ((a1 < k1) | ((a1 == k1) & lt2)));
  }
}

// One step D of the prefix (pl, ph: lanes [nf, pos]) and suffix (ql, qh:
// lanes [pos, nf + nc)) boxes of a node's lanes.
template <int D>
__device__ inline void tiny_scan_step(float pl[3], float ph[3], float ql[3], float qh[3], bool busy, uint32_t pos,
                                      uint32_t nf, uint32_t nc) {
  float ul[3], uh[3], dl[3], dh[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    ul[k] = row_move<ROW_SHR + D>(pl[k]);
    uh[k] = row_move<ROW_SHR + D>(ph[k]);
    dl[k] = row_move<ROW_SHL + D>(ql[k]);
    dh[k] = row_move<ROW_SHL + D>(qh[k]);
  }
  if (busy && pos >= nf + (uint32_t)D) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      pl[k] = fminf(pl[k], ul[k]);
      ph[k] = fmaxf(ph[k], uh[k]);
    }
  }
  if (busy && pos + (uint32_t)D < nf + nc) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      ql[k] = fminf(ql[k], dl[k]);
      qh[k] = fmaxf(qh[k], dh[k]);
    }
  }
}

__device__ inline bool tiny_better(float c, int a, uint32_t k, float bc, int ba, uint32_t bk) {
  return c < bc || (c == bc && (a < ba || (a == ba && k < bk)));
}

// Whole subtrees of up to TINY_GROUPS tiny segments (2..TINY_MAX
// references each), one per 16-lane group of a wave: group g finishes
// q[g] (g < nq) with exact SAH splits, lane j of the group holding the
// reference at position j of its segment. The decisions are those of a
// sequential finisher: at each node the references are stably sorted by
// centroid (lo + hi) on x, then on y, then on z, each sort starting from
// the order the previous one left; every split position of each of the
// three orders is priced A_L*n_L + A_R*n_R; the first minimum over (axis,
// position) wins; and the node's references take the order of a stable
// sort on the winning axis applied to the z order (for z, the z order
// itself). A stable sort here is a rank: the number of the node's
// references with a smaller key tuple, (k_x, position) for the x order,
// (k_y, k_x, position) for y, (k_z, k_y, k_x, position) for z. All nodes of
// one depth are split together. vals receives the final order.
__device__ inline void tiny_groups(const TinyItem* q, int nq, const Bufs& B, const Tree& t, TinyLds* TG, int lane) {
  const int g = lane >> 4;
  const uint32_t pos = (uint32_t)lane & (TINY_MAX - 1u);
  TinyLds& T = TG[g];
  uint32_t first = 0u, c = 0u, np = 0u, nm = 0u, par = 0u;
  if (g < nq) {
    first = q[g].s.first;
    c = q[g].s.count;
    np = q[g].s.parent;
    nm = q[g].s.meta;
    par = q[g].par;
    if (c < 2u || c > TINY_MAX) {
      if (pos == 0u) atomicOr(&t.cnt->error, 16u);
      c = 0u;
    }
  }
  const bool act = pos < c;
  float4 l = make_float4(0.0f, 0.0f, 0.0f, 0.0f), h = l;
  uint32_t id = 0u;
  if (act) {
    l = B.blo[par][first + pos];
    h = B.bhi[par][first + pos];
    id = B.idx[par][first + pos];
  }
  // the node holding this lane's position: [nf, nf + nc) of the segment
  uint32_t nf = 0u, nc = act ? c : 0u;
  for (uint32_t it = 0;; it++) {
    if (nc == 1u) {
      t.child[2 * np + (nm & 1u)] = (first + nf) | REF_LEAF;
      t.parent_leaf[first + nf] = np;
      nc = 0u;
    }
    const bool busy = nc >= 2u;
    if (__ballot(busy) == 0) break;
    if (it >= TINY_MAX) {
      if (lane == 0) atomicOr(&t.cnt->error, 16u);
      break;
    }
    const float k0 = l.x + h.x, k1 = l.y + h.y, k2 = l.z + h.z;
    uint32_t r0 = 0u, r1 = 0u, r2 = 0u, f0 = 0u, f1 = 0u;
    tiny_ranks(k0, k1, k2, pos, nf, nc, r0, r1, r2, f0, f1);
    float bc = INFINITY;
    int ba = 0;
    uint32_t bk = 1u;
#pragma unroll 1
    for (int a = 0; a < 3; a++) {
      const uint32_t r = a == 0 ? r0 : (a == 1 ? r1 : r2);
      group_sync<true>();
      if (busy) {
        T.lo[nf + r] = l;
        T.hi[nf + r] = h;
      }
      group_sync<true>();
      float4 sl = l, sh = h;
      if (busy) {
        sl = T.lo[pos];
        sh = T.hi[pos];
      }
      // prefix box over the node's sorted positions [nf, pos], suffix over [pos, nf + nc)
      float pl[3] = {sl.x, sl.y, sl.z}, ph[3] = {sh.x, sh.y, sh.z};
      float ql[3] = {sl.x, sl.y, sl.z}, qh[3] = {sh.x, sh.y, sh.z};
      static_assert(TINY_MAX == 16, "the scans take 16-lane rows");
      tiny_scan_step<1>(pl, ph, ql, qh, busy, pos, nf, nc);
      tiny_scan_step<2>(pl, ph, ql, qh, busy, pos, nf, nc);
      tiny_scan_step<4>(pl, ph, ql, qh, busy, pos, nf, nc);
      tiny_scan_step<8>(pl, ph, ql, qh, busy, pos, nf, nc);
      float rl[3], rh[3];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        rl[k] = row_move<ROW_SHL + 1>(ql[k]);
        rh[k] = row_move<ROW_SHL + 1>(qh[k]);
      }
      // candidate: the first pos - nf + 1 sorted references left
      if (busy && pos + 1u < nf + nc) {
        const uint32_t k = pos - nf + 1u;
        const float cost = half_area(ph[0] - pl[0], ph[1] - pl[1], ph[2] - pl[2]) * (float)k +
                           half_area(rh[0] - rl[0], rh[1] - rl[1], rh[2] - rl[2]) * (float)(nc - k);
        if (cost < bc) {
          bc = cost;
          ba = a;
          bk = k;
        }
      }
    }
    // least (cost, axis, position) of each node reaches the node's first lane
    auto reduce_step = [&](float oc, int oa, uint32_t ok, uint32_t d) {
      if (busy && pos + d < nf + nc && tiny_better(oc, oa, ok, bc, ba, bk)) {
        bc = oc;
        ba = oa;
        bk = ok;
      }
    };
    reduce_step(row_move<ROW_SHL + 1>(bc), row_move<ROW_SHL + 1>(ba), row_move<ROW_SHL + 1>(bk), 1u);
    reduce_step(row_move<ROW_SHL + 2>(bc), row_move<ROW_SHL + 2>(ba), row_move<ROW_SHL + 2>(bk), 2u);
    reduce_step(row_move<ROW_SHL + 4>(bc), row_move<ROW_SHL + 4>(ba), row_move<ROW_SHL + 4>(bk), 4u);
    reduce_step(row_move<ROW_SHL + 8>(bc), row_move<ROW_SHL + 8>(ba), row_move<ROW_SHL + 8>(bk), 8u);
    ba = __shfl(ba, (int)nf, TINY_MAX);
    bk = __shfl(bk, (int)nf, TINY_MAX);
    const uint32_t rf = ba == 0 ? f0 : (ba == 1 ? f1 : r2);
    const uint32_t mid = first + nf + bk - 1u;
    if (busy && pos == nf) {
      t.child[2 * np + (nm & 1u)] = mid;
      t.parent_int[mid] = np;
      t.range[mid] = make_uint2(first + nf, first + nf + nc - 1u);
      set_depth(t, mid, nm >> 1);
    }
    group_sync<true>();
    if (busy) {
      T.lo[nf + rf] = l;
      T.hi[nf + rf] = h;
      T.id[nf + rf] = id;
    }
    group_sync<true>();
    if (busy) {
      l = T.lo[pos];
      h = T.hi[pos];
      id = T.id[pos];
      const uint32_t dm = ((nm >> 1) + 1u) << 1;
      if (pos - nf < bk) {
        nc = bk;
        nm = dm;
      } else {
        nf += bk;
        nc -= bk;
        nm = dm | 1u;
      }
      np = mid;
    }
  }
  if (act) t.vals[first + pos] = id;
  group_sync<true>();
}

// Whole subtrees of the tiny segments, TINY_GROUPS per wave.
constexpr int TINY_TB = 256;
constexpr int TINY_PER_BLOCK = TINY_GROUPS * (TINY_TB / 64);

__global__ __launch_bounds__(TINY_TB) void k_tiny(const TinyItem* __restrict__ items, uint32_t nitems, Bufs B,
                                                  Tree t) {
  __shared__ TinyLds T[TINY_TB / 64][TINY_GROUPS];
  __shared__ TinyItem Q[TINY_TB / 64][TINY_GROUPS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t first = blockIdx.x * TINY_PER_BLOCK + (uint32_t)wave * TINY_GROUPS;
  if (first >= nitems) return;
  const int nq = (int)min((uint32_t)TINY_GROUPS, nitems - first);
  if (lane < nq) Q[wave][lane] = items[first + lane];
  group_sync<true>();
  tiny_groups(Q[wave], nq, B, t, T[wave], lane);
}

// Writes the big-segment id of every position of the initial big segments
// (segment blockIdx.y, positions in blocks of TBS along x).
__global__ void k_fill_seg_ids(uint32_t* sid, const Seg* segs) {
  const Seg sg = segs[blockIdx.y];
  const uint32_t i = blockIdx.x * TBS + threadIdx.x;
  if (i < sg.count) sid[sg.first + i] = blockIdx.y;
}

}  // namespace

void build_sah_tree(uint32_t n, uint32_t* vals, const float4* slo, const float4* shi,
                    const std::vector<SahRoot>& trees, uint32_t* child, uint32_t* parent_int,
                    uint32_t* parent_leaf, uint2* range, uint8_t* depth, const LearnedCost* learned,
                    uint32_t fin_max, SahScratch& S, SahTreeOut& out, hipStream_t s) {
  if (fin_max > MEDIAN_MAX) fin_max = MEDIAN_MAX;
  LearnedCost lcost{};
  if (learned) lcost = *learned;
  if (trees.empty()) throw hip_error(hipErrorInvalidValue, "SAH build without a tree");
  if (trees.size() >= ROOT_TAG) throw hip_error(hipErrorInvalidValue, "too many trees");
  for (const SahRoot& r : trees)
    if (r.count < 2 || (uint64_t)r.first + r.count > n) throw hip_error(hipErrorInvalidValue, "bad SAH tree range");
  // segments of one list are disjoint and larger than the class below,
  // except trees' roots, which may be small at any size
  const uint32_t ntrees = (uint32_t)trees.size();
  const uint32_t cap[3] = {n / MID_MAX + 2, n / (SMALL_MAX + 1) + 2, n / (TINY_MAX + 1) + 2 + ntrees};
  const uint32_t tiny_cap = n / 2 + 2;  // tiny segments are disjoint and hold two or more references
  for (int p = 0; p < 2; p++) {
    S.idx[p].reserve((size_t)n * 4);
    S.blo[p].reserve((size_t)n * 16);
    S.bhi[p].reserve((size_t)n * 16);
    S.sid[p].reserve((size_t)n * 4);
  }
  for (int c = 0; c < 3; c++)
    for (int l = 0; l < 3; l++) S.segs[c][l].reserve((size_t)cap[c] * sizeof(Seg));
  S.tiny.reserve((size_t)tiny_cap * sizeof(TinyItem));
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

  // trees start as big, mid or small segments of the level loop (a tree's
  // root is split by bins even when tiny)
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
  if (cur[0]) {  // the big-segment kernels read every position's big-segment id
    hipLaunchKernelGGL(k_fill_u32, dim3(div_up(n, TBS)), dim3(TBS), 0, s, S.sid[0].as<uint32_t>(), (size_t)n, NIL);
    uint32_t longest = 0;
    for (const Seg& sg : init[0]) longest = std::max(longest, sg.count);
    hipLaunchKernelGGL(k_fill_seg_ids, dim3(div_up(longest, TBS), cur[0]), dim3(TBS), 0, s, S.sid[0].as<uint32_t>(),
                       S.segs[0][0].as<Seg>());
  }
  Counters hc{};
  for (int c = 0; c < 3; c++) hc.nseg[0][c] = cur[c];
  hc.ntiny = 0;
  hc.error = 0;
  RTG_HIP(hipMemcpyAsync(cnt, &hc, sizeof(hc), hipMemcpyHostToDevice, s));
  RTG_HIP(hipStreamSynchronize(s));  // the host staging above must outlive its copies

  Tree t{child, parent_int, parent_leaf, range, vals, cnt, S.roots.as<uint32_t>(), depth};
  // Level l reads the segment lists of parity l & 1 and fills those of the
  // other parity; its kernels take their segment counts from the device, so
  // the host queues level l + 1 before it knows whether level l left any
  // work, and reads level l's counters (through a pinned slot) meanwhile.
  // big-segment kernels run until a level is known to have left no big
  // segment (children are smaller than their parents, so none comes back)
  S.rb.ensure(sizeof(Counters));
  bool big_live = cur[0] > 0;
  // Grids: a level's kernels read their counts on the device, so a grid
  // need only cover an upper bound. The host knows the input counts of
  // level kin_level (read back, or level 0's); a segment has at most two
  // children, big ones only from big, mid ones from big or mid.
  uint32_t kin[3] = {cur[0], cur[1], cur[2]};
  uint32_t kin_level = 0;
  auto grid_bound = [&](uint32_t L, uint32_t b[3]) {
    uint64_t v[3] = {cap[0], cap[1], cap[2]};
    if (kin_level == L) {
      for (int c = 0; c < 3; c++) v[c] = kin[c];
    } else if (kin_level + 1 == L) {
      v[0] = 2ull * kin[0];
      v[1] = 2ull * ((uint64_t)kin[0] + kin[1]);
      v[2] = 2ull * ((uint64_t)kin[0] + kin[1] + kin[2]);
    }
    for (int c = 0; c < 3; c++) b[c] = (uint32_t)std::min<uint64_t>(v[c], cap[c]);
  };
  // level L: references (and big-segment ids) in buffers of parity L & 1,
  // written to the other parity; segment lists and counts in slot L % 3,
  // children to slot (L + 1) % 3, while slot (L + 2) % 3, last read by
  // level L - 1, is cleared for level L + 1
  auto enqueue_level = [&](uint32_t L) {
    const int in = (int)(L & 1u), o = in ^ 1;
    const int li = (int)(L % 3u), lo = (int)((L + 1u) % 3u), lz = (int)((L + 2u) % 3u);
    Lists nx;
    for (int c = 0; c < 3; c++) {
      nx.seg[c] = S.segs[c][lo].as<Seg>();
      nx.cap[c] = cap[c];
    }
    nx.nseg = cnt->nseg[lo];
    nx.tiny = S.tiny.as<TinyItem>();
    nx.tiny_cap = tiny_cap;
    nx.par = (uint32_t)o;
    nx.fin_max = fin_max;
    const Seg* bs = S.segs[0][li].as<Seg>();
    uint32_t gb[3];
    grid_bound(L, gb);
    if (big_live && gb[0]) {
      hipLaunchKernelGGL(k_init_gbins, dim3(div_up((size_t)gb[0] * BIN_WORDS, TBS)), dim3(TBS), 0, s,
                         S.bins.as<uint32_t>(), &cnt->nseg[li][0]);
      hipLaunchKernelGGL(k_big_bins, dim3(div_up(n, CHUNK)), dim3(TBS), 0, s, S.sid[in].as<uint32_t>(),
                         S.blo[in].as<float4>(), S.bhi[in].as<float4>(), n, bs, S.bins.as<uint32_t>());
      if (lcost.P)
        hipLaunchKernelGGL(k_big_split<true>, dim3(gb[0]), dim3(TBS), 0, s, bs, &cnt->nseg[li][0],
                           S.bins.as<uint32_t>(), S.splits.as<Split>(), nx, t, lcost);
      else
        hipLaunchKernelGGL(k_big_split<false>, dim3(gb[0]), dim3(TBS), 0, s, bs, &cnt->nseg[li][0],
                           S.bins.as<uint32_t>(), S.splits.as<Split>(), nx, t, lcost);
      hipLaunchKernelGGL(k_big_flags, dim3(div_up(n, TBS)), dim3(TBS), 0, s, S.sid[in].as<uint32_t>(),
                         S.blo[in].as<float4>(), S.bhi[in].as<float4>(), n, bs, S.splits.as<Split>(),
                         S.flags.as<uint32_t>());
      S.scanner.exclusive(S.flags.as<uint32_t>(), S.scan.as<uint32_t>(), n, s);
      hipLaunchKernelGGL(k_big_scatter, dim3(div_up(n, TBS)), dim3(TBS), 0, s, S.sid[in].as<uint32_t>(),
                         S.idx[in].as<uint32_t>(), S.blo[in].as<float4>(), S.bhi[in].as<float4>(), n, bs,
                         S.splits.as<Split>(), S.flags.as<uint32_t>(), S.scan.as<uint32_t>(), S.sid[o].as<uint32_t>(),
                         S.idx[o].as<uint32_t>(), S.blo[o].as<float4>(), S.bhi[o].as<float4>(), vals, cnt);
    }
    // (k_seg clears the next level's counts: it runs even with no mid segment)
    hipLaunchKernelGGL(lcost.P ? k_seg<true> : k_seg<false>, dim3(std::max(gb[1], 1u)), dim3(256), 0, s, S.segs[1][li].as<Seg>(),
                       &cnt->nseg[li][1], S.idx[in].as<uint32_t>(), S.blo[in].as<float4>(), S.bhi[in].as<float4>(),
                       S.idx[o].as<uint32_t>(), S.blo[o].as<float4>(), S.bhi[o].as<float4>(), nx, t, lcost,
                       cnt->nseg[lz]);
    if (gb[2])
      hipLaunchKernelGGL(lcost.P ? k_seg_wave<true> : k_seg_wave<false>, dim3(div_up(gb[2], SEGW_WAVES)), dim3(SEGW_TB), 0, s,
                       S.segs[2][li].as<Seg>(), &cnt->nseg[li][2], S.idx[in].as<uint32_t>(), S.blo[in].as<float4>(),
                       S.bhi[in].as<float4>(), S.idx[o].as<uint32_t>(), S.blo[o].as<float4>(), S.bhi[o].as<float4>(),
                       nx, t, lcost);
    RTG_HIP(hipGetLastError());
  };
  auto readback = [&](uint32_t level) {
    RTG_HIP(hipMemcpyAsync(S.rb.slot((int)level), cnt, sizeof(Counters), hipMemcpyDeviceToHost, s));
    RTG_HIP(hipEventRecord(S.rb.ev[level % LevelReadback::K], s));
  };
  uint32_t levels = 0;
  if (cur[0] + cur[1] + cur[2] > 0) {
    enqueue_level(0);
    readback(0);
    for (levels = 1;; levels++) {
      if (levels > 4096) {
        (void)hipStreamSynchronize(s);
        throw hip_error(hipErrorUnknown, "SAH build did not terminate");
      }
      enqueue_level(levels);  // speculative: without input it does nothing
      readback(levels);
      RTG_HIP(hipEventSynchronize(S.rb.ev[(levels - 1) % LevelReadback::K]));
      memcpy(&hc, S.rb.slot((int)(levels - 1)), sizeof(hc));
      if (hc.error) {
        (void)hipStreamSynchronize(s);
        throw hip_error(hipErrorUnknown, "SAH build failed: error bits " + std::to_string(hc.error));
      }
      const uint32_t* next = hc.nseg[levels % 3u];  // input of the level just queued
      for (int c = 0; c < 3; c++) kin[c] = next[c];
      kin_level = levels;
      if (next[0] + next[1] + next[2] == 0) break;
      if (next[0] == 0) big_live = false;
    }
  }
  // the tiny segments of all levels
  if (hc.ntiny > tiny_cap) throw hip_error(hipErrorUnknown, "SAH build: tiny segment list overflowed");
  if (hc.ntiny) {
    Bufs B;
    for (int p = 0; p < 2; p++) {
      B.idx[p] = S.idx[p].as<uint32_t>();
      B.blo[p] = S.blo[p].as<float4>();
      B.bhi[p] = S.bhi[p].as<float4>();
    }
    hipLaunchKernelGGL(k_tiny, dim3(div_up(hc.ntiny, TINY_PER_BLOCK)), dim3(TINY_TB), 0, s, S.tiny.as<TinyItem>(),
                       hc.ntiny, B, t);
    RTG_HIP(hipGetLastError());
  }
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
