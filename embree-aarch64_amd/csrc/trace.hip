// Ray traversal kernels: one ray per lane over the quantised BVH8.
//
// Reference anchors:
//   stream entry rules (rays with !(tnear <= tfar) are skipped, occluded rays
//     get tfar = -inf): kernels/common/rtcore.cpp:595-623, :848-874,
//     kernels/bvh/bvh_intersector_stream_filters.cpp:135-151, bvh_intersector1.cpp:191-193;
//   per-ray stack traversal, nearest child first, far siblings pushed with
//     their entry distance and skipped on pop once beyond the current hit:
//     kernels/bvh/bvh_intersector1.cpp:31-119, bvh_traverser1.h:519-643;
//   robust slab test (rdir scaled by 1 -/+ 3 ulp, zero_fix):
//     kernels/bvh/node_intersector1.h:122-190, common/math/vec3fa.h:172-176;
//   Moeller-Trumbore triangle test, same FP32/FMA sequence:
//     kernels/geometry/triangle_intersector_moeller.h:62-114, common/math/vec3.h:204,209,
//     common/math/math.h:367-372;
//   hit write-back: kernels/geometry/intersector_epilog.h:280-292.
// This file is compiled with -ffp-contract=off: every FMA is explicit.
//
// Differences from the reference, shared with the CPU oracle so results are
// traversal-order independent: t/u/v use IEEE division; the closest hit is
// the lexicographic minimum of (t, geomID, primID) over all valid hits of
// the original [tnear, tfar]; culling is conservative (robust slab test on
// boxes that contain the widened triangle boxes, interval ends widened by
// 2^-12 relative).
#include "trace.h"

#include <mutex>

// Scratch of the no-allocation queries (word 0: ray counter, word 1: error
// bits), one copy per GPU. In host code this name is a host-side stand-in,
// not a device address: kernels get the address noalloc_setup() looked up.
static __device__ __align__(256) uint32_t scratch[rtg::NOALLOC_SCRATCH_WORDS];

namespace rtg {

constexpr int TB = 256;
constexpr int STACK = (int)TRACE_STACK;  // entries per ray (the builder checks its trees against it)
// A node step pushes up to 7 entries after checking only that the stack
// holds at most STACK (fits()), so the arrays have room for 7 more.
constexpr int STACK_CAP = STACK + 8;

// Per-lane traversal stack: the top LDS_SLOTS entries live in LDS (slot-major,
// so lanes touching the same slot hit distinct banks), older entries spill
// to a private array only when the ring is full. Entry = (ref, entry distance).
// The spill array is a separate local of the kernel: were it a member here,
// a dynamically indexed store into it could alias sp and lo, and the
// compiler would keep the whole struct in scratch memory.
// The LDS pointer is typed in the local address space so that the ring is
// reached with ds_* instructions (a generic pointer makes every access a flat
// access through the vector memory pipeline).
typedef __attribute__((address_space(3))) uint64_t lds_u64;
__device__ inline uint64_t pack_entry(uint2 e) { return (uint64_t)e.x | ((uint64_t)e.y << 32); }
__device__ inline uint2 unpack_entry(uint64_t v) { return make_uint2((uint32_t)v, (uint32_t)(v >> 32)); }

// Second word of a closest-hit stack entry: the entry distance rounded down
// to a bfloat16 pattern (high half; -inf for a negative distance) and the
// least distance of the entries below it (low half), so that a pop can tell
// when every entry left is farther than the closest hit.
__device__ inline uint32_t pack_dist(uint32_t d16, float& smin) {
  const uint32_t y = (d16 << 16) | (__float_as_uint(smin) >> 16);
  smin = fminf(smin, __uint_as_float(d16 << 16));
  return y;
}
__device__ inline float entry_dist(uint32_t y) { return __uint_as_float(y & 0xffff0000u); }
__device__ inline float below_dist(uint32_t y) { return __uint_as_float(y << 16); }

template <int LDS_SLOTS>
struct RingStack {
  // A power-of-two ring finds entry i's slot as i & (LDS_SLOTS - 1); any
  // other size keeps the slot of entry sp in `top`. When the ring is full,
  // the oldest ring entry (lo = sp - LDS_SLOTS) sits in entry sp's slot.
  static constexpr bool POW2 = (LDS_SLOTS & (LDS_SLOTS - 1)) == 0;
  lds_u64* lds;    // this lane's slot 0; slot k is lds + k * TB
  uint64_t* spill; // STACK_CAP entries, private
  int sp, lo;      // [0, lo) spilled, [lo, sp) in the ring
  int top;         // sp mod LDS_SLOTS (other sizes than powers of two)
  __device__ inline int slot() const {
    if constexpr (POW2) return sp & (LDS_SLOTS - 1);
    else return top;
  }
  __device__ inline void init(uint64_t* base, uint64_t* spill_mem) {
    lds = (lds_u64*)base;
    spill = spill_mem;
    sp = lo = top = 0;
  }
  // Whether a node's entries (at most 7) may be put(): true while the stack
  // holds at most STACK entries, which a tree the builder accepted never
  // exceeds; past it the arrays' 7 spare entries keep a corrupt tree's
  // pushes in bounds until the error stops the ray.
  __device__ inline bool fits() const { return sp <= STACK; }
  __device__ inline void put(uint2 e) {
    if (sp - lo == LDS_SLOTS) {
      spill[lo] = lds[slot() * TB];
      lo++;
    }
    lds[slot() * TB] = pack_entry(e);
    sp++;
    if constexpr (!POW2) top = top == LDS_SLOTS - 1 ? 0 : top + 1;
  }
  __device__ inline bool push(uint2 e) {
    if (sp >= STACK) return false;
    put(e);
    return true;
  }
  __device__ inline uint2 pop() {
    sp--;
    if constexpr (!POW2) top = top == 0 ? LDS_SLOTS - 1 : top - 1;
    // an entry below the ring comes back through its ring slot, so that the
    // read below is always an LDS read
    if (sp < lo) {
      lo = sp;
      lds[slot() * TB] = spill[sp];
    }
    return unpack_entry(lds[slot() * TB]);
  }
  __device__ inline bool empty() const { return sp == 0; }
  __device__ inline void clear() { sp = lo = top = 0; }
};

// LDS_SLOTS == 0: the whole stack is a private array.
template <>
struct RingStack<0> {
  uint64_t* spill;
  int sp;
  __device__ inline void init(uint64_t*, uint64_t* spill_mem) {
    spill = spill_mem;
    sp = 0;
  }
  __device__ inline bool push(uint2 e) {
    if (sp >= STACK) return false;
    spill[sp++] = pack_entry(e);
    return true;
  }
  __device__ inline bool fits() const { return sp <= STACK; }
  __device__ inline void put(uint2 e) { spill[sp++] = pack_entry(e); }
  __device__ inline uint2 pop() { return unpack_entry(spill[--sp]); }
  __device__ inline bool empty() const { return sp == 0; }
  __device__ inline void clear() { sp = 0; }
};

// Rays of a persistent wave come from a private run of consecutive slots
// that the wave takes from the global counter CHUNK at a time, so a refill
// costs a global atomic only once every CHUNK rays (the atomic's latency
// stalls the whole wave). Every lane computes the same values (all inputs
// are wave-uniform), so the state needs no broadcast after the one atomic.
struct WavePool {
  static constexpr uint32_t CHUNK = 256;
  uint32_t lb = 0, le = 0;  // unassigned slots [lb, le) of the current run
  bool gdone = false;       // the global counter has passed M
  // Hands out nidle slots (idle lane `rank` gets the rank-th); returns this
  // lane's slot or 0xffffffff when the rays have run out.
  __device__ inline uint32_t take(uint32_t* next, uint32_t M, int nidle, uint32_t rank, int lane) {
    const uint32_t have = le - lb;
    uint32_t nb = 0, ne = 0;
    if (have < (uint32_t)nidle && !gdone) {
      uint32_t g = 0;
      if (lane == 0) g = atomicAdd(next, CHUNK);
      g = __shfl(g, 0, 64);
      if ((uint64_t)g + CHUNK >= M) gdone = true;
      if (g < M) {
        nb = g;
        ne = (uint64_t)g + CHUNK < M ? g + CHUNK : M;
      }
    }
    uint32_t slot = 0xffffffffu;
    if (rank < have) slot = lb + rank;
    else if (nb + (rank - have) < ne) slot = nb + (rank - have);
    if (have >= (uint32_t)nidle) {
      lb += (uint32_t)nidle;
    } else {
      const uint32_t need = (uint32_t)nidle - have;
      lb = nb + (need < ne - nb ? need : ne - nb);
      le = ne;
    }
    return slot;
  }
  __device__ inline bool empty() const { return gdone && lb >= le; }
};

__device__ inline float widen_up(float x) { return __builtin_isinf(x) ? x : x + fabsf(x) * CULL_MARGIN; }
__device__ inline float widen_down(float x) { return __builtin_isinf(x) ? x : x - fabsf(x) * CULL_MARGIN; }

struct TRay {
  float ox, oy, oz, dx, dy, dz;
  float rnx, rny, rnz, rfx, rfy, rfz;
  bool negx, negy, negz;
  float tnear, tfar, tlo;
};

__device__ inline float zero_fix(float d) { return fabsf(d) < MIN_RCP_INPUT ? MIN_RCP_INPUT : d; }

__device__ inline void setup_ray(TRay& r, float4 a, float4 b, float tfar) {
  const float round_down = 1.0f - 3.0f * 1.1920928955078125e-7f;
  const float round_up = 1.0f + 3.0f * 1.1920928955078125e-7f;
  r.ox = a.x; r.oy = a.y; r.oz = a.z; r.tnear = a.w;
  r.dx = b.x; r.dy = b.y; r.dz = b.z;
  r.tfar = tfar;
  const float rx = 1.0f / zero_fix(b.x), ry = 1.0f / zero_fix(b.y), rz = 1.0f / zero_fix(b.z);
  r.rnx = round_down * rx; r.rny = round_down * ry; r.rnz = round_down * rz;
  r.rfx = round_up * rx; r.rfy = round_up * ry; r.rfz = round_up * rz;
  r.negx = !(r.rnx >= 0.0f);
  r.negy = !(r.rny >= 0.0f);
  r.negz = !(r.rnz >= 0.0f);
  r.tlo = widen_down(a.w);
}

__device__ inline uint32_t spread3(uint32_t x) {
  x = (x | (x << 16)) & 0x030000FFu;
  x = (x | (x << 8)) & 0x0300F00Fu;
  x = (x | (x << 4)) & 0x030C30C3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

__device__ inline float byte_f(uint32_t w, int k) { return (float)((w >> (8 * k)) & 0xffu); }

struct NodeHit {
  uint32_t mask;
  uint32_t imask;
  uint32_t cb, tb;  // reference bases of meta_ref()
  uint32_t meta_lo, meta_hi;
  float tn[8];
  bool hit[8];
};

// Slab test of the 8 children on planes that decode as fmaf(2^(e-127), q, p)
// (the expression the builder verified against every widened child box),
// with rdir rounded down for near and up for far planes as the reference's
// robust TravRay does (kernels/bvh/node_intersector1.h:122-190).
__device__ inline void intersect_node(const QNode8* __restrict__ nodes, uint32_t i, const TRay& r, float thi,
                                      NodeHit& h) {
  const uint4* np = reinterpret_cast<const uint4*>(nodes + i);
  const uint4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3], w4 = np[4];
  const float px = __uint_as_float(w0.x), py = __uint_as_float(w0.y), pz = __uint_as_float(w0.z);
  const float sx = __uint_as_float((w0.w & 0xffu) << 23);
  const float sy = __uint_as_float(((w0.w >> 8) & 0xffu) << 23);
  const float sz = __uint_as_float(((w0.w >> 16) & 0xffu) << 23);
  h.imask = w0.w >> 24;
  h.cb = w1.x - 0x80u;
  h.tb = (w1.y | REF_LEAF) - 1u;
  // Defined here, so that the compiler keeps the node's 16 B loads whole
  // instead of loading these two words again where the pushes use them.
  asm volatile("" : "+v"(h.cb), "+v"(h.tb));
  h.meta_lo = w1.z;
  h.meta_hi = w1.w;
  // near / far plane bytes per axis (lo plane is near for a positive direction)
  const uint32_t nx0 = r.negx ? w2.z : w2.x, nx1 = r.negx ? w2.w : w2.y;
  const uint32_t fx0 = r.negx ? w2.x : w2.z, fx1 = r.negx ? w2.y : w2.w;
  const uint32_t ny0 = r.negy ? w3.z : w3.x, ny1 = r.negy ? w3.w : w3.y;
  const uint32_t fy0 = r.negy ? w3.x : w3.z, fy1 = r.negy ? w3.y : w3.w;
  const uint32_t nz0 = r.negz ? w4.z : w4.x, nz1 = r.negz ? w4.w : w4.y;
  const uint32_t fz0 = r.negz ? w4.x : w4.z, fz1 = r.negz ? w4.y : w4.w;
  // Plane distance (p + s q - o) rdir as one FMA per plane, q a + b with
  // a = s rdir (exact: s is a power of two) and b = (p - o) rdir (two
  // roundings) formed once per node. Against the decode-then-subtract form
  // the error moves from |plane - o| to |p - o|: at most u|b| + u/2 |t|
  // (u = 2^-23). Where |b| is close to |t| (origin away from the node) the
  // 3-ulp scaling of rdir covers it; where it cancels (origin near or in the
  // node), |p - o| is within the node's extent, below 2 max|coord|, and the
  // builder's box widening by max|coord| 2^-20 = 8u max|coord| covers it. So
  // no box holding a hit is culled (the hits of this form equal the
  // oracle's on every parity test), and each plane costs a byte conversion
  // and one FMA instead of an FMA, a subtraction and a product.
  const float anx = sx * r.rnx, afx = sx * r.rfx, any = sy * r.rny, afy = sy * r.rfy, anz = sz * r.rnz, afz = sz * r.rfz;
  const float bnx = (px - r.ox) * r.rnx, bfx = (px - r.ox) * r.rfx;
  const float bny = (py - r.oy) * r.rny, bfy = (py - r.oy) * r.rfy;
  const float bnz = (pz - r.oz) * r.rnz, bfz = (pz - r.oz) * r.rfz;
  uint32_t mask = 0;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int k = c & 3;
    const uint32_t mb = ((c < 4 ? h.meta_lo : h.meta_hi) >> (8 * k)) & 0xffu;
    const float qnx = byte_f(c < 4 ? nx0 : nx1, k), qfx = byte_f(c < 4 ? fx0 : fx1, k);
    const float qny = byte_f(c < 4 ? ny0 : ny1, k), qfy = byte_f(c < 4 ? fy0 : fy1, k);
    const float qnz = byte_f(c < 4 ? nz0 : nz1, k), qfz = byte_f(c < 4 ? fz0 : fz1, k);
    const float tnx = __builtin_fmaf(qnx, anx, bnx), tfx = __builtin_fmaf(qfx, afx, bfx);
    const float tny = __builtin_fmaf(qny, any, bny), tfy = __builtin_fmaf(qfy, afy, bfy);
    const float tnz = __builtin_fmaf(qnz, anz, bnz), tfz = __builtin_fmaf(qfz, afz, bfz);
    const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, r.tlo));
    const float tf = fminf(fminf(tfx, tfy), fminf(tfz, thi));
    h.tn[c] = tn;
    h.hit[c] = (tn <= tf) && (mb != 0u);
    mask |= h.hit[c] ? (1u << c) : 0u;
  }
  h.mask = mask;
}

// A child's reference from its meta byte m (internal children 0x80 | rank,
// leaves 1 + offset of the first record): m plus a base chosen by bit 7,
// cb = child_base - 0x80 or tb = (tri_base | REF_LEAF) - 1.
__device__ inline uint32_t meta_ref(uint32_t m, uint32_t cb, uint32_t tb) { return m + ((m & 0x80u) ? cb : tb); }

__device__ inline uint32_t child_ref(const NodeHit& h, int c) {
  const uint32_t mb = ((c < 4 ? h.meta_lo : h.meta_hi) >> (8 * (c & 3))) & 0xffu;
  return meta_ref(mb, h.cb, h.tb);
}

// Occlusion: pushes the hit children except the first in slot order, sets
// `next` to that one and returns whether any child was hit.
template <class Stack>
__device__ inline bool push_any_order(const NodeHit& h, Stack& stack, uint32_t* err, uint32_t& next) {
  if (!h.mask) return false;
  const uint32_t cb = h.cb, tb = h.tb;
  const uint64_t meta = (uint64_t)h.meta_lo | ((uint64_t)h.meta_hi << 32);
  const int c0 = __builtin_ctz(h.mask);
  uint32_t rem = h.mask & (h.mask - 1u);
  if (stack.fits()) {
    while (rem) {
      const int c = __builtin_ctz(rem);
      rem &= rem - 1u;
      stack.put(make_uint2(meta_ref((uint32_t)(meta >> (8 * c)) & 0xffu, cb, tb), 0u));
    }
  } else {
    atomicOr(err, 1u);
  }
  next = meta_ref((uint32_t)(meta >> (8 * c0)) & 0xffu, cb, tb);
  return true;
}

// Closest hit: pushes the hit children except the nearest, farthest first,
// each with its entry distance, sets `next` to the nearest child and returns
// whether any child was hit.
// A key is the entry distance's bit pattern with the child's meta byte in the
// low 8 bits (distinct among a node's children), or 0xffffffff for a child
// that is not hit. Entry distances are never below the ray's widened tnear,
// so for tnear >= 0 the bit patterns order like the distances (a negative
// distance, possible only with tnear < 0, orders as farther than every
// positive one: a different visiting order, the same result). The
// perturbation of the low 8 bits changes only the order among nearly equal
// children. A 19-comparator network (Batcher's odd-even merge sort) orders
// the keys decreasingly without branches: children not hit first, the
// nearest hit child last. The stored distance is the key's high half
// (rounded down; -inf for a negative distance), never above the true entry
// distance, so the pop test stays conservative.
template <class Stack>
__device__ inline bool push_near_to_far(const NodeHit& h, Stack& stack, float& smin, uint32_t* err,
                                        uint32_t& next) {
  constexpr uint32_t NO_HIT = 0xffffffffu;
  uint32_t k[8];
  // v_perm_b32 selector: byte 0 from the meta word (byte c & 3), bytes 1-3
  // from the distance
#pragma unroll
  for (int c = 0; c < 8; c++)
    k[c] = h.hit[c] ? __builtin_amdgcn_perm(__float_as_uint(h.tn[c]), c < 4 ? h.meta_lo : h.meta_hi,
                                            0x07060500u | (uint32_t)(c & 3))
                    : NO_HIT;
  auto ce = [&](int i, int j) {
    const uint32_t hi = max(k[i], k[j]), lo = min(k[i], k[j]);
    k[i] = hi;
    k[j] = lo;
  };
  ce(0, 1); ce(2, 3); ce(4, 5); ce(6, 7);
  ce(0, 2); ce(1, 3); ce(4, 6); ce(5, 7);
  ce(1, 2); ce(5, 6);
  ce(0, 4); ce(1, 5); ce(2, 6); ce(3, 7);
  ce(2, 4); ce(3, 5);
  ce(1, 2); ce(3, 4); ce(5, 6);
  if (k[7] == NO_HIT) return false;
  const uint32_t cb = h.cb, tb = h.tb;
  if (stack.fits()) {
#pragma unroll
    for (int j = 0; j < 7; j++) {
      if (k[j] != NO_HIT) {
        const uint32_t dk = ((int32_t)k[j] < 0) ? 0xff800000u : k[j];
        // (dk's high half, smin's high half)
        stack.put(make_uint2(meta_ref(k[j] & 0xffu, cb, tb), __builtin_amdgcn_perm(dk, __float_as_uint(smin), 0x07060302u)));
        const float d = __uint_as_float(dk & 0xffff0000u);
        smin = d < smin ? d : smin;
      }
    }
  } else {
    atomicOr(err, 1u);
  }
  next = meta_ref(k[7] & 0xffu, cb, tb);
  return true;
}

struct Hit {
  float t, u, v, ngx, ngy, ngz;
  uint32_t prim, geom;
  bool found;
};

__device__ inline float prod_diff(float a, float b, float c, float d) { return __builtin_fmaf(a, b, -(c * d)); }
__device__ inline float dot3(float ax, float ay, float az, float bx, float by, float bz) {
  return __builtin_fmaf(ax, bx, __builtin_fmaf(ay, by, az * bz));
}

// One triangle test. t is final; the hit's barycentrics are un/dn and vn/dn
// (divided only when the hit improves the current best).
struct TriTest {
  float t, un, vn, dn, ngx, ngy, ngz;
  uint32_t prim, geom, flags;
};

__device__ inline float sse_min(float a, float b) { return a < b ? a : b; }
__device__ inline float sse_max(float a, float b) { return a > b ? a : b; }

// PL = false: Moeller-Trumbore on the record (v0, e1 = v0-v1, e2 = v2-v0).
// PL = true: Pluecker (robust scenes) on the record (v0, v1, v2),
//   kernels/geometry/triangle_intersector_pluecker.h:55-110 and finalize
//   :29-36: edge functions of the origin-shifted vertices, eps = ulp*|U+V+W|,
//   stable_triangle_normal (common/math/vec3.h:210-222), den and T doubled,
//   t = T/den, u = U/UVW (U*0 when |UVW| < min_rcp_input), min/max of three
//   with SSE semantics; valid for tnear <= t <= tfar.
template <bool PL>
__device__ inline bool tri_test(const TRay& r, const TriRec* __restrict__ tp, TriTest& o) {
  const float4* p4 = reinterpret_cast<const float4*>(tp);
  const float4 a = p4[0], b = p4[1], c = p4[2];
  o.prim = __float_as_uint(a.w);
  o.geom = __float_as_uint(b.w);
  o.flags = __float_as_uint(c.w);
  if (!PL) {
    const float Cx = a.x - r.ox, Cy = a.y - r.oy, Cz = a.z - r.oz;
    // R = cross(C, D)
    const float Rx = prod_diff(Cy, r.dz, Cz, r.dy);
    const float Ry = prod_diff(Cz, r.dx, Cx, r.dz);
    const float Rz = prod_diff(Cx, r.dy, Cy, r.dx);
    // Ng = cross(e2, e1)
    o.ngx = prod_diff(c.y, b.z, c.z, b.y);
    o.ngy = prod_diff(c.z, b.x, c.x, b.z);
    o.ngz = prod_diff(c.x, b.y, c.y, b.x);
    const float den = dot3(o.ngx, o.ngy, o.ngz, r.dx, r.dy, r.dz);
    const float absDen = fabsf(den);
    const uint32_t sgn = __float_as_uint(den) & 0x80000000u;
    const float U = __uint_as_float(__float_as_uint(dot3(Rx, Ry, Rz, c.x, c.y, c.z)) ^ sgn);
    const float V = __uint_as_float(__float_as_uint(dot3(Rx, Ry, Rz, b.x, b.y, b.z)) ^ sgn);
    bool valid = (den != 0.0f) & (U >= 0.0f) & (V >= 0.0f) & (U + V <= absDen);
    const float T = __uint_as_float(__float_as_uint(dot3(o.ngx, o.ngy, o.ngz, Cx, Cy, Cz)) ^ sgn);
    valid &= (absDen * r.tnear < T) & (T <= absDen * r.tfar);
    if (valid) o.t = T / absDen;
    o.un = U;
    o.vn = V;
    o.dn = absDen;
    return valid;
  } else {
    const float v0x = a.x - r.ox, v0y = a.y - r.oy, v0z = a.z - r.oz;
    const float v1x = b.x - r.ox, v1y = b.y - r.oy, v1z = b.z - r.oz;
    const float v2x = c.x - r.ox, v2y = c.y - r.oy, v2z = c.z - r.oz;
    const float e0x = v2x - v0x, e0y = v2y - v0y, e0z = v2z - v0z;
    const float e1x = v0x - v1x, e1y = v0y - v1y, e1z = v0z - v1z;
    const float e2x = v1x - v2x, e2y = v1y - v2y, e2z = v1z - v2z;
    float sx = v2x + v0x, sy = v2y + v0y, sz = v2z + v0z;
    const float U = dot3(prod_diff(e0y, sz, e0z, sy), prod_diff(e0z, sx, e0x, sz), prod_diff(e0x, sy, e0y, sx),
                         r.dx, r.dy, r.dz);
    sx = v0x + v1x; sy = v0y + v1y; sz = v0z + v1z;
    const float V = dot3(prod_diff(e1y, sz, e1z, sy), prod_diff(e1z, sx, e1x, sz), prod_diff(e1x, sy, e1y, sx),
                         r.dx, r.dy, r.dz);
    sx = v1x + v2x; sy = v1y + v2y; sz = v1z + v2z;
    const float W = dot3(prod_diff(e2y, sz, e2z, sy), prod_diff(e2z, sx, e2x, sz), prod_diff(e2x, sy, e2y, sx),
                         r.dx, r.dy, r.dz);
    const float UVW = U + V + W;
    const float eps = 1.1920928955078125e-7f * fabsf(UVW);
    bool valid = (sse_min(sse_min(U, V), W) >= -eps) || (sse_max(sse_max(U, V), W) <= eps);
    // stable_triangle_normal(e0, e1, e2)
    const float ab_x = e0z * e1y, ab_y = e0x * e1z, ab_z = e0y * e1x;
    const float bc_x = e1z * e2y, bc_y = e1x * e2z, bc_z = e1y * e2x;
    o.ngx = fabsf(ab_x) < fabsf(bc_x) ? __builtin_fmaf(e0y, e1z, -ab_x) : __builtin_fmaf(e1y, e2z, -bc_x);
    o.ngy = fabsf(ab_y) < fabsf(bc_y) ? __builtin_fmaf(e0z, e1x, -ab_y) : __builtin_fmaf(e1z, e2x, -bc_y);
    o.ngz = fabsf(ab_z) < fabsf(bc_z) ? __builtin_fmaf(e0x, e1y, -ab_z) : __builtin_fmaf(e1x, e2y, -bc_z);
    const float d = dot3(o.ngx, o.ngy, o.ngz, r.dx, r.dy, r.dz);
    const float den = d + d;
    const float tt = dot3(v0x, v0y, v0z, o.ngx, o.ngy, o.ngz);
    const float T = tt + tt;
    o.t = T / den;
    valid &= (r.tnear <= o.t) & (o.t <= r.tfar) & (den != 0.0f);
    const bool tiny = fabsf(UVW) < MIN_RCP_INPUT;
    o.un = tiny ? U * 0.0f : U;
    o.vn = tiny ? V * 0.0f : V;
    o.dn = tiny ? 1.0f : UVW;
    return valid;
  }
}

template <bool ALIGNED>
__device__ inline void load_ray(const char* rp, float4& a, float4& b, float& tfar) {
  if (ALIGNED) {
    a = *reinterpret_cast<const float4*>(rp);
    b = *reinterpret_cast<const float4*>(rp + 16);
  } else {
    const float* f = reinterpret_cast<const float*>(rp);
    a = make_float4(f[0], f[1], f[2], f[3]);
    b = make_float4(f[4], f[5], f[6], f[7]);
  }
  tfar = *reinterpret_cast<const float*>(rp + 32);
}

template <bool OCCLUDED, bool ALIGNED, bool COUNT, int LDS_SLOTS, bool PL>
__global__ __launch_bounds__(TB) void k_trace(DevScene sc, char* __restrict__ rays, uint32_t M, uint64_t stride,
                                              uint32_t inst_id, uint32_t* __restrict__ cnt_nodes,
                                              uint32_t* __restrict__ cnt_tris, uint32_t* __restrict__ err,
                                              const uint32_t* __restrict__ perm, uint32_t* node_visits) {
  const uint32_t slot = blockIdx.x * TB + threadIdx.x;
  if (slot >= M) return;
  const uint32_t i = perm ? perm[slot] : slot;  // lanes take rays in spatially sorted order when given
  char* rp = rays + (uint64_t)i * stride;
  float4 a, b;
  float tfar;
  load_ray<ALIGNED>(rp, a, b, tfar);
  uint32_t nvis = 0, ntri = 0;
  if (!(a.w <= tfar)) {
    if (COUNT) { cnt_nodes[i] = 0; cnt_tris[i] = 0; }
    return;
  }
  TRay r;
  setup_ray(r, a, b, tfar);
  Hit best;
  best.found = false;
  best.t = INFINITY;
  best.prim = best.geom = 0xffffffffu;
  best.u = best.v = best.ngx = best.ngy = best.ngz = 0.0f;
  float lim = widen_up(tfar);

  __shared__ uint64_t lds_stack[LDS_SLOTS > 0 ? LDS_SLOTS * TB : 1];
  uint64_t spill_mem[STACK_CAP];
  RingStack<LDS_SLOTS> stack;
  stack.init(lds_stack + threadIdx.x, spill_mem);
  float smin = INFINITY;  // least entry distance on the stack (closest hit)
  uint32_t cur = sc.root;
  bool occluded = false;
  // every node and triangle record is reached at most once per ray, so the
  // loop runs at most num_nodes + num_tris times; the cap turns a corrupt
  // structure into an error instead of a hang
  // (node and triangle counts are below 2^31 each, so the bound fits 32 bits)
  const uint32_t max_iter = sc.num_nodes + sc.num_tris + 16u;
  uint32_t iter = 0;
  while (true) {
    if (++iter > max_iter) {
      atomicOr(err, 2u);
      break;
    }
    if (!(cur & REF_LEAF)) {
      if (cur >= sc.num_nodes) {
        atomicOr(err, 4u);
        break;
      }
      NodeHit h;
      intersect_node(sc.nodes, cur, r, lim, h);
      if (COUNT) {
        nvis++;
        if (node_visits) atomicAdd(&node_visits[cur], 1u);
      }
      if (OCCLUDED ? push_any_order(h, stack, err, cur) : push_near_to_far(h, stack, smin, err, cur)) {
        continue;
      }
    } else {
      uint32_t ti = cur & ~REF_LEAF;
      while (true) {
        if (ti >= sc.num_tris) {
          atomicOr(err, 4u);
          break;
        }
        TriTest h;
        const bool valid = tri_test<PL>(r, sc.tris + ti, h);
        if (COUNT) ntri++;
        if (valid) {
          if (OCCLUDED) {
            occluded = true;
            break;
          }
          const bool better = !best.found || h.t < best.t ||
                              (h.t == best.t && (h.geom < best.geom || (h.geom == best.geom && h.prim < best.prim)));
          if (better) {
            best.found = true;
            best.t = h.t;
            best.u = h.un / h.dn;
            best.v = h.vn / h.dn;
            best.ngx = h.ngx; best.ngy = h.ngy; best.ngz = h.ngz;
            best.prim = h.prim; best.geom = h.geom;
          }
        }
        if (h.flags & TRI_LAST) break;
        ti++;
      }
      if (OCCLUDED && occluded) break;
      if (!OCCLUDED && best.found) lim = widen_up(fminf(best.t, tfar));
    }
    // pop the next entry that can still hold a closer hit
    bool got = false;
    while (!stack.empty()) {
      const uint2 e = stack.pop();
      if (!OCCLUDED) smin = below_dist(e.y);
      if (OCCLUDED || entry_dist(e.y) <= lim) {
        cur = e.x;
        got = true;
        break;
      }
      if (smin > lim) {
        // every entry left is farther than the closest hit so far
        stack.clear();
        smin = INFINITY;
        break;
      }
    }
    if (!got) break;
  }
  if (COUNT) { cnt_nodes[i] = nvis; cnt_tris[i] = ntri; }
  if (OCCLUDED) {
    if (occluded) *reinterpret_cast<float*>(rp + 32) = -INFINITY;
  } else if (best.found) {
    float* f = reinterpret_cast<float*>(rp);
    uint32_t* u = reinterpret_cast<uint32_t*>(rp);
    f[8] = best.t;
    f[12] = best.ngx;
    f[13] = best.ngy;
    f[14] = best.ngz;
    f[15] = best.u;
    f[16] = best.v;
    u[17] = best.prim;
    u[18] = best.geom;
    u[19] = inst_id;
  }
}

// Persistent variant: a fixed grid of waves pulls rays from a global counter.
// A lane whose ray is finished takes a new one as soon as at least
// `refill_min` lanes of its wave are idle, so incoherent rays of very
// different lengths keep the SIMD lanes busy (Aila and Laine, "Understanding
// the efficiency of ray traversal on GPUs", HPG 2009). Each loop iteration is
// one traversal step (one node, or one leaf plus the pop that follows) for
// every active lane. Every wave leaves the loop once the counter has passed
// M and its lanes are idle.
// Resident 256-lane blocks per compute unit the persistent kernel is
// compiled for (one wave per SIMD each): the register budget follows from it.
// 6 (at most 80 VGPRs, no spills) measured fastest on the 2^20-triangle soup
// with 2^24 incoherent rays: 1655 Mrays/s closest hit against 1585 at 7 (72
// VGPRs, five of them spilled to scratch in the node step); 5 compiles to the
// same code as 6. A 16-entry LDS stack (32 KB per block) leaves room for 5.
constexpr int TRACE_MIN_BLOCKS = 6;

// BATCH: the ray slots 0..M-1 are the concatenation of the streams in
// `items` (sorted by `first`); a lane finds its stream by binary search and
// starts at that stream's scene root.
template <bool OCCLUDED, bool ALIGNED, bool COUNT, int LDS_SLOTS, bool PL, bool BATCH = false>
__global__ __launch_bounds__(TB, LDS_SLOTS > 12 ? 5 : TRACE_MIN_BLOCKS) void k_trace_p(DevScene sc, char* __restrict__ rays, uint32_t M, uint64_t stride,
                                                uint32_t inst_id, uint32_t* __restrict__ cnt_nodes,
                                                uint32_t* __restrict__ cnt_tris, uint32_t* __restrict__ err,
                                                const uint32_t* __restrict__ perm, uint32_t* __restrict__ next,
                                                int refill_min, uint32_t* node_visits,
                                                const BatchItem* __restrict__ items, uint32_t nitems) {
  const int lane = threadIdx.x & 63;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  __shared__ uint64_t lds_stack[LDS_SLOTS > 0 ? LDS_SLOTS * TB : 1];
  uint64_t spill_mem[STACK_CAP];
  RingStack<LDS_SLOTS> stack;
  stack.init(lds_stack + threadIdx.x, spill_mem);
  float smin = INFINITY;  // least entry distance on the stack (closest hit)

  TRay r;
  Hit best;
  float lim = 0.0f;
  uint32_t cur = 0, ray = 0, nvis = 0, ntri = 0;
  char* lrp = nullptr;      // BATCH: this lane's record
  uint32_t lroot = 0;       // BATCH: its scene's root
  bool active = false, occluded = false, pool_empty = false;
  WavePool pool;
  // a ray touches each node and triangle record at most once
  // (node and triangle counts are below 2^31 each, so the bound fits 32 bits)
  const uint32_t max_iter = sc.num_nodes + sc.num_tris + 16u;
  uint32_t iter = 0;

  while (true) {
    // ---- refill idle lanes (wave-uniform decision)
    if (!pool_empty) {
      const uint64_t idle = __ballot(!active);
      const int nidle = __popcll(idle);
      if (nidle >= refill_min) {
        const uint32_t slot = pool.take(next, M, nidle, (uint32_t)__popcll(idle & lt), lane);
        pool_empty = pool.empty();
        if (!active) {
          if (slot != 0xffffffffu) {
            char* src;
            if (BATCH) {
              uint32_t lo = 0, hi = nitems - 1;
              while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (items[mid].first <= slot) lo = mid; else hi = mid - 1;
              }
              const BatchItem& it = items[lo];
              ray = slot - it.first;
              lrp = it.rays + (uint64_t)ray * it.stride;
              lroot = it.root;
              src = lrp;
            } else {
              ray = perm ? perm[slot] : slot;
              src = rays + (uint64_t)ray * stride;
            }
            float4 a, b;
            float tfar;
            load_ray<ALIGNED>(src, a, b, tfar);
            if (a.w <= tfar) {
              setup_ray(r, a, b, tfar);
              best.found = false;
              best.t = INFINITY;
              best.prim = best.geom = 0xffffffffu;
              best.u = best.v = best.ngx = best.ngy = best.ngz = 0.0f;
              lim = widen_up(tfar);
              cur = BATCH ? lroot : sc.root;
              stack.init(lds_stack + threadIdx.x, spill_mem);
              smin = INFINITY;
              occluded = false;
              nvis = ntri = 0;
              iter = 0;
              active = true;
            } else if (COUNT) {
              cnt_nodes[ray] = 0;
              cnt_tris[ray] = 0;
            }
          }
        }
      }
    }
    if (__ballot(active) == 0) {
      if (pool_empty) break;
      continue;
    }
    if (!active) continue;

    // ---- one traversal step
    bool done = false;
    if (++iter > max_iter) {
      atomicOr(err, 2u);
      done = true;
    } else if (!(cur & REF_LEAF)) {
      if (cur >= sc.num_nodes) {
        atomicOr(err, 4u);
        done = true;
      } else {
        NodeHit h;
        intersect_node(sc.nodes, cur, r, lim, h);
        if (COUNT) {
          nvis++;
          if (node_visits) atomicAdd(&node_visits[cur], 1u);
        }
        if (OCCLUDED ? push_any_order(h, stack, err, cur) : push_near_to_far(h, stack, smin, err, cur)) {
          continue;  // next step descends into cur
        }
      }
    } else {
      // one triangle per step: a lane in a long leaf does not hold the
      // wave's other lanes in a triangle loop
      const uint32_t ti = cur & ~REF_LEAF;
      if (ti >= sc.num_tris) {
        atomicOr(err, 4u);
        done = true;
      } else {
        TriTest h;
        const bool valid = tri_test<PL>(r, sc.tris + ti, h);
        if (COUNT) ntri++;
        if (valid) {
          if (OCCLUDED) {
            occluded = true;
          } else {
            const bool better = !best.found || h.t < best.t ||
                                (h.t == best.t && (h.geom < best.geom || (h.geom == best.geom && h.prim < best.prim)));
            if (better) {
              best.found = true;
              best.t = h.t;
              best.u = h.un / h.dn;
              best.v = h.vn / h.dn;
              best.ngx = h.ngx; best.ngy = h.ngy; best.ngz = h.ngz;
              best.prim = h.prim; best.geom = h.geom;
              lim = widen_up(fminf(best.t, r.tfar));
            }
          }
        }
        if (OCCLUDED && occluded) {
          done = true;
        } else if (!(h.flags & TRI_LAST)) {
          cur = (ti + 1) | REF_LEAF;  // the leaf's next record
          continue;
        }
      }
    }
    if (!done) {
      // pop the next entry that can still hold a closer hit
      bool got = false;
      while (!stack.empty()) {
        const uint2 e = stack.pop();
        if (!OCCLUDED) smin = below_dist(e.y);
        if (OCCLUDED || entry_dist(e.y) <= lim) {
          cur = e.x;
          got = true;
          break;
        }
        if (smin > lim) {
          // every entry left is farther than the closest hit so far
          stack.clear();
          smin = INFINITY;
          break;
        }
      }
      done = !got;
    }
    if (done) {
      char* rp = BATCH ? lrp : rays + (uint64_t)ray * stride;
      if (COUNT) { cnt_nodes[ray] = nvis; cnt_tris[ray] = ntri; }
      if (OCCLUDED) {
        if (occluded) *reinterpret_cast<float*>(rp + 32) = -INFINITY;
      } else if (best.found) {
        float* f = reinterpret_cast<float*>(rp);
        uint32_t* u = reinterpret_cast<uint32_t*>(rp);
        f[8] = best.t;
        f[12] = best.ngx;
        f[13] = best.ngy;
        f[14] = best.ngz;
        f[15] = best.u;
        f[16] = best.v;
        u[17] = best.prim;
        u[18] = best.geom;
        u[19] = inst_id;
      }
      active = false;
    }
  }
}

// Persistent traversal with a second cursor for triangles: a lane keeps
// one leaf in progress (`tri`) next to its node cursor (`cur`), and each
// loop iteration does one node step and one triangle step, so the wave's
// triangle branch no longer runs while those lanes' node work waits (and the
// other way round). A leaf met while a leaf is still in progress waits in
// `cur`. Node entries are popped after the triangle step, so the cull test
// sees the newest hit distance; nodes tested while a leaf is pending use the
// older one (speculative; hits do not depend on it).
template <bool OCCLUDED, bool ALIGNED, bool COUNT, bool PL>
__global__ __launch_bounds__(TB, TRACE_MIN_BLOCKS) void k_trace_d(DevScene sc, char* __restrict__ rays, uint32_t M,
                                                                    uint64_t stride, uint32_t inst_id,
                                                                    uint32_t* __restrict__ cnt_nodes,
                                                                    uint32_t* __restrict__ cnt_tris,
                                                                    uint32_t* __restrict__ err,
                                                                    const uint32_t* __restrict__ perm,
                                                                    uint32_t* __restrict__ next, int refill_min,
                                                                    uint32_t* node_visits) {
  constexpr uint32_t NONE_REF = 0xffffffffu;
  const int lane = threadIdx.x & 63;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  __shared__ uint64_t lds_stack[8 * TB];
  uint64_t spill_mem[STACK_CAP];
  RingStack<8> stack;
  stack.init(lds_stack + threadIdx.x, spill_mem);
  float smin = INFINITY;  // least entry distance on the stack (closest hit)

  TRay r;
  Hit best;
  float lim = 0.0f;
  uint32_t cur = NONE_REF, tri = NONE_REF, ray = 0, nvis = 0, ntri = 0;
  bool active = false, occluded = false, pool_empty = false;
  WavePool pool;
  // (node and triangle counts are below 2^31 each, so the bound fits 32 bits)
  const uint32_t max_iter = sc.num_nodes + sc.num_tris + 16u;
  uint32_t iter = 0;

  while (true) {
    if (!pool_empty) {
      const uint64_t idle = __ballot(!active);
      const int nidle = __popcll(idle);
      if (nidle >= refill_min) {
        const uint32_t slot = pool.take(next, M, nidle, (uint32_t)__popcll(idle & lt), lane);
        pool_empty = pool.empty();
        if (!active && slot != 0xffffffffu) {
          ray = perm ? perm[slot] : slot;
          float4 a, b;
          float tfar;
          load_ray<ALIGNED>(rays + (uint64_t)ray * stride, a, b, tfar);
          if (a.w <= tfar) {
            setup_ray(r, a, b, tfar);
            best.found = false;
            best.t = INFINITY;
            best.prim = best.geom = 0xffffffffu;
            best.u = best.v = best.ngx = best.ngy = best.ngz = 0.0f;
            lim = widen_up(tfar);
            cur = sc.root;
            tri = NONE_REF;
            stack.init(lds_stack + threadIdx.x, spill_mem);
            smin = INFINITY;
            occluded = false;
            nvis = ntri = 0;
            iter = 0;
            active = true;
          } else if (COUNT) {
            cnt_nodes[ray] = 0;
            cnt_tris[ray] = 0;
          }
        }
      }
    }
    if (__ballot(active) == 0) {
      if (pool_empty) break;
      continue;
    }
    if (!active) continue;

    bool done = false;
    if (++iter > max_iter) {
      atomicOr(err, 2u);
      done = true;
    }
    // a waiting leaf starts once the triangle cursor is free
    if (!done && tri == NONE_REF && cur != NONE_REF && (cur & REF_LEAF)) {
      tri = cur & ~REF_LEAF;
      cur = NONE_REF;
    }
    // node step
    if (!done && cur != NONE_REF && !(cur & REF_LEAF)) {
      if (cur >= sc.num_nodes) {
        atomicOr(err, 4u);
        done = true;
      } else {
        NodeHit h;
        intersect_node(sc.nodes, cur, r, lim, h);
        if (COUNT) {
          nvis++;
          if (node_visits) atomicAdd(&node_visits[cur], 1u);
        }
        if (!(OCCLUDED ? push_any_order(h, stack, err, cur) : push_near_to_far(h, stack, smin, err, cur))) {
          cur = NONE_REF;
        }
      }
    }
    // triangle step
    if (!done && tri != NONE_REF) {
      if (tri >= sc.num_tris) {
        atomicOr(err, 4u);
        done = true;
      } else {
        TriTest h;
        const bool valid = tri_test<PL>(r, sc.tris + tri, h);
        if (COUNT) ntri++;
        if (valid) {
          if (OCCLUDED) {
            occluded = true;
          } else {
            const bool better = !best.found || h.t < best.t ||
                                (h.t == best.t && (h.geom < best.geom || (h.geom == best.geom && h.prim < best.prim)));
            if (better) {
              best.found = true;
              best.t = h.t;
              best.u = h.un / h.dn;
              best.v = h.vn / h.dn;
              best.ngx = h.ngx; best.ngy = h.ngy; best.ngz = h.ngz;
              best.prim = h.prim; best.geom = h.geom;
              lim = widen_up(fminf(best.t, r.tfar));
            }
          }
        }
        if (OCCLUDED && occluded) done = true;
        else tri = (h.flags & TRI_LAST) ? NONE_REF : tri + 1u;
      }
    }
    // the node cursor takes the next entry that can still hold a closer hit
    if (!done && cur == NONE_REF) {
      while (!stack.empty()) {
        const uint2 e = stack.pop();
        if (!OCCLUDED) smin = below_dist(e.y);
        if (OCCLUDED || entry_dist(e.y) <= lim) {
          cur = e.x;
          break;
        }
        if (smin > lim) {
          // every entry left is farther than the closest hit so far
          stack.clear();
          smin = INFINITY;
          break;
        }
      }
      if (cur == NONE_REF && tri == NONE_REF) done = true;
    }
    if (done) {
      char* rp = rays + (uint64_t)ray * stride;
      if (COUNT) { cnt_nodes[ray] = nvis; cnt_tris[ray] = ntri; }
      if (OCCLUDED) {
        if (occluded) *reinterpret_cast<float*>(rp + 32) = -INFINITY;
      } else if (best.found) {
        float* f = reinterpret_cast<float*>(rp);
        uint32_t* u = reinterpret_cast<uint32_t*>(rp);
        f[8] = best.t;
        f[12] = best.ngx;
        f[13] = best.ngy;
        f[14] = best.ngz;
        f[15] = best.u;
        f[16] = best.v;
        u[17] = best.prim;
        u[18] = best.geom;
        u[19] = inst_id;
      }
      active = false;
    }
  }
}

// ------------------------------------------------------------- two-level --

// Two-level traversal of a scene with instances. The top tree's leaves are
// InstRec records; reaching one pushes a REF_INST_EXIT entry, moves the ray
// into the object's space with the record's world-to-object transform
// (kernels/geometry/instance_intersector.cpp:52-71 for closest hit, :73-98
// for occlusion) and continues at the object tree's root; popping the exit
// entry restores the world-space ray from its record. Both trees share one
// stack, so the current hit distance culls object entries and top entries
// alike (t has the same meaning in both spaces). Persistent and refilled as
// k_trace_p. Hits are ordered by (t, instID, geomID, primID).
// ctx_inst != RTC_INVALID_GEOMETRY_ID: the query context's one-level
// instance stack is already in use, so every instance push fails and only
// the scene's own triangles are tested (kernels/common/instance_stack.h:21-38).
template <bool OCCLUDED, bool ALIGNED, bool CNT = false>
__global__ __launch_bounds__(TB, 5) void k_trace_inst(DevTop top, char* __restrict__ rays, uint32_t M,
                                                      uint64_t stride, uint32_t ctx_inst, uint32_t* __restrict__ err,
                                                      uint32_t* __restrict__ next, int refill_min,
                                                      uint32_t* __restrict__ top_visits = nullptr,
                                                      uint32_t* __restrict__ obj_visits = nullptr,
                                                      uint32_t target_obj = 0xffffffffu) {
  const int lane = threadIdx.x & 63;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  __shared__ uint64_t lds_stack[8 * TB];
  // the world-space ray and its reciprocals while the lane is in a moved
  // object (restored from here on the way out, not reloaded and divided again)
  __shared__ float4 lds_world[3 * TB];
  uint64_t spill_mem[STACK_CAP];
  RingStack<8> stack;
  stack.init(lds_stack + threadIdx.x, spill_mem);
  float smin = INFINITY;  // least entry distance on the stack (closest hit)

  TRay r;
  Hit best;
  uint32_t best_inst = 0;
  float lim = 0.0f;
  uint32_t cur = 0, ray = 0, inst = 0;
  bool in_obj = false;  // in an object tree (else the top tree)
  bool moved = false;   // r is in an instance's object space
  const QNode8* onodes = nullptr;
  const TriRec* otris = nullptr;
  uint32_t onn = 0, otn = 0;
  uint32_t oidx = 0xffffffffu;  // the object the lane is in (counted variant)
  bool orobust = false;
  bool active = false, occluded = false, pool_empty = false;
  WavePool pool;
  uint64_t iter = 0;

  while (true) {
    // ---- refill idle lanes (wave-uniform decision)
    if (!pool_empty) {
      const uint64_t idle = __ballot(!active);
      const int nidle = __popcll(idle);
      if (nidle >= refill_min) {
        const uint32_t slot = pool.take(next, M, nidle, (uint32_t)__popcll(idle & lt), lane);
        pool_empty = pool.empty();
        if (!active) {
          if (slot != 0xffffffffu) {
            ray = slot;
            float4 a, b;
            float tfar;
            load_ray<ALIGNED>(rays + (uint64_t)ray * stride, a, b, tfar);
            if (a.w <= tfar) {
              setup_ray(r, a, b, tfar);
              best.found = false;
              best.t = INFINITY;
              best.prim = best.geom = 0xffffffffu;
              best.u = best.v = best.ngx = best.ngy = best.ngz = 0.0f;
              best_inst = 0xffffffffu;
              lim = widen_up(tfar);
              cur = top.root;
              in_obj = false;
              moved = false;
              stack.init(lds_stack + threadIdx.x, spill_mem);
              smin = INFINITY;
              occluded = false;
              iter = 0;
              active = true;
            }
          }
        }
      }
    }
    if (__ballot(active) == 0) {
      if (pool_empty) break;
      continue;
    }
    if (!active) continue;

    // ---- one traversal step
    bool done = false;
    if (++iter > top.max_iter) {
      atomicOr(err, 2u);
      done = true;
    } else if (!(cur & REF_LEAF)) {
      const QNode8* nodes = in_obj ? onodes : top.nodes;
      const uint32_t nn = in_obj ? onn : top.num_nodes;
      if (cur >= nn) {
        atomicOr(err, 4u);
        done = true;
      } else {
        if (CNT) {
          if (!in_obj) {
            if (top_visits) atomicAdd(top_visits + cur, 1u);
          } else if (obj_visits && oidx == target_obj) {
            atomicAdd(obj_visits + cur, 1u);
          }
        }
        NodeHit h;
        intersect_node(nodes, cur, r, lim, h);
        if (OCCLUDED ? push_any_order(h, stack, err, cur) : push_near_to_far(h, stack, smin, err, cur)) {
          continue;
        }
      }
    } else if (in_obj) {
      // one triangle of the current object
      const uint32_t ti = cur & ~REF_LEAF;
      if (ti >= otn) {
        atomicOr(err, 4u);
        done = true;
      } else {
        TriTest h;
        const bool valid = orobust ? tri_test<true>(r, otris + ti, h) : tri_test<false>(r, otris + ti, h);
        if (valid) {
          if (OCCLUDED) {
            occluded = true;
          } else {
            const bool better =
                !best.found || h.t < best.t ||
                (h.t == best.t &&
                 (inst < best_inst ||
                  (inst == best_inst &&
                   (h.geom < best.geom || (h.geom == best.geom && h.prim < best.prim)))));
            if (better) {
              // the hit goes to the record at once, so that only the fields
              // the comparison needs stay in registers
              best.found = true;
              best.t = h.t;
              best.prim = h.prim;
              best.geom = h.geom;
              best_inst = inst;
              float* f = reinterpret_cast<float*>(rays + (uint64_t)ray * stride);
              uint32_t* u = reinterpret_cast<uint32_t*>(f);
              f[8] = h.t;
              f[12] = h.ngx;
              f[13] = h.ngy;
              f[14] = h.ngz;
              f[15] = h.un / h.dn;
              f[16] = h.vn / h.dn;
              u[17] = h.prim;
              u[18] = h.geom;
              u[19] = inst;
              lim = widen_up(fminf(best.t, r.tfar));
            }
          }
        }
        if (OCCLUDED && occluded) {
          done = true;
        } else if (!(h.flags & TRI_LAST)) {
          cur = (ti + 1) | REF_LEAF;
          continue;
        }
      }
    } else {
      // an instance record of the top tree
      const uint32_t ii = cur & ~REF_LEAF;
      if (ii >= top.num_insts) {
        atomicOr(err, 4u);
        done = true;
      } else {
        const float4* q = reinterpret_cast<const float4*>(top.insts + ii);
        const float4 m0 = q[0], m1 = q[1], m2 = q[2];
        const uint4 w = reinterpret_cast<const uint4*>(top.insts + ii)[3];
        const bool identity = (w.z & INST_IDENTITY) != 0u;
        // the rest of a multi-record leaf resumes after this instance
        if (!(w.z & TRI_LAST) && !stack.push(make_uint2((ii + 1u) | REF_LEAF, pack_dist(0xff80u, smin))))
          atomicOr(err, 1u);
        if ((identity || ctx_inst == 0xffffffffu) && w.y < top.num_objects) {
          const DevScene o = top.objects[w.y];
          if (o.num_nodes > 0) {
            if (!stack.push(make_uint2(REF_INST_EXIT, pack_dist(0xff80u, smin)))) atomicOr(err, 1u);
            if (!identity) {
              lds_world[threadIdx.x] = make_float4(r.ox, r.oy, r.oz, r.dx);
              lds_world[TB + threadIdx.x] = make_float4(r.dy, r.dz, r.rnx, r.rny);
              lds_world[2 * TB + threadIdx.x] = make_float4(r.rnz, r.rfx, r.rfy, r.rfz);
              // columns vx = m[0..2], vy = m[3..5], vz = m[6..8], p = m[9..11]
              const float ox = __builtin_fmaf(r.ox, m0.x, __builtin_fmaf(r.oy, m0.w, __builtin_fmaf(r.oz, m1.z, m2.y)));
              const float oy = __builtin_fmaf(r.ox, m0.y, __builtin_fmaf(r.oy, m1.x, __builtin_fmaf(r.oz, m1.w, m2.z)));
              const float oz = __builtin_fmaf(r.ox, m0.z, __builtin_fmaf(r.oy, m1.y, __builtin_fmaf(r.oz, m2.x, m2.w)));
              const float dx = __builtin_fmaf(r.dx, m0.x, __builtin_fmaf(r.dy, m0.w, r.dz * m1.z));
              const float dy = __builtin_fmaf(r.dx, m0.y, __builtin_fmaf(r.dy, m1.x, r.dz * m1.w));
              const float dz = __builtin_fmaf(r.dx, m0.z, __builtin_fmaf(r.dy, m1.y, r.dz * m2.x));
              setup_ray(r, make_float4(ox, oy, oz, r.tnear), make_float4(dx, dy, dz, 0.0f), r.tfar);
              moved = true;
            }
            in_obj = true;
            inst = identity ? ctx_inst : w.x;
            if (CNT) oidx = w.y;
            onodes = o.nodes;
            otris = o.tris;
            onn = o.num_nodes;
            otn = o.num_tris;
            orobust = o.robust != 0u;
            cur = o.root;
            continue;
          }
        }
      }
    }
    if (!done) {
      // pop the next entry that can still hold a closer hit; an exit entry
      // returns the ray to world space
      bool got = false;
      while (!stack.empty()) {
        const uint2 e = stack.pop();
        if (!OCCLUDED) smin = below_dist(e.y);
        if (e.x == REF_INST_EXIT) {
          in_obj = false;
          if (moved) {
            const float4 a = lds_world[threadIdx.x], b = lds_world[TB + threadIdx.x];
            const float4 c = lds_world[2 * TB + threadIdx.x];
            r.ox = a.x; r.oy = a.y; r.oz = a.z; r.dx = a.w;
            r.dy = b.x; r.dz = b.y; r.rnx = b.z; r.rny = b.w;
            r.rnz = c.x; r.rfx = c.y; r.rfy = c.z; r.rfz = c.w;
            r.negx = !(r.rnx >= 0.0f);
            r.negy = !(r.rny >= 0.0f);
            r.negz = !(r.rnz >= 0.0f);
            moved = false;
          }
          continue;
        }
        if (OCCLUDED || entry_dist(e.y) <= lim) {
          cur = e.x;
          got = true;
          break;
        }
        if (smin > lim) {
          // every entry left is farther than the closest hit so far
          stack.clear();
          smin = INFINITY;
          break;
        }
      }
      done = !got;
    }
    if (done) {
      char* rp = rays + (uint64_t)ray * stride;
      if (OCCLUDED) {
        if (occluded) *reinterpret_cast<float*>(rp + 32) = -INFINITY;
      }
      active = false;
    }
  }
}

__global__ void k_zero_u32(uint32_t* p) { *p = 0u; }

// Sort key of a ray: 30-bit Morton code of its origin in the scene box
// (origins outside are clamped to the box faces); inactive rays sort last.
__global__ __launch_bounds__(TB) void k_ray_keys(const char* __restrict__ rays, uint32_t M, uint64_t stride, float lx,
                                                 float ly, float lz, float sx, float sy, float sz,
                                                 uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t i = blockIdx.x * TB + threadIdx.x;
  if (i >= M) return;
  const float* f = reinterpret_cast<const float*>(rays + (uint64_t)i * stride);
  const float ox = f[0], oy = f[1], oz = f[2], tnear = f[3], tfar = f[8];
  uint32_t key = 0xffffffffu;
  if (tnear <= tfar) {
    const uint32_t qx = (uint32_t)fminf(fmaxf((ox - lx) * sx, 0.0f), 1023.0f);
    const uint32_t qy = (uint32_t)fminf(fmaxf((oy - ly) * sy, 0.0f), 1023.0f);
    const uint32_t qz = (uint32_t)fminf(fmaxf((oz - lz) * sz, 0.0f), 1023.0f);
    key = spread3(qx) | (spread3(qy) << 1) | (spread3(qz) << 2);
  }
  keys[i] = key;
  vals[i] = i;
}

struct LaunchArgs {
  DevScene sc;
  char* rays;
  uint32_t M;
  uint64_t stride;
  uint32_t inst_id;
  uint32_t *cnt_nodes, *cnt_tris, *err;
  const uint32_t* perm;
  uint32_t* counter;
  int refill;
  uint32_t* node_visits;
  bool persistent;
  dim3 grid, pgrid;
  int lds_slots;
  hipStream_t s;
  bool dual;
};

template <bool OCC, bool AL, bool CNT, int SL, bool PL>
void launch_one(const LaunchArgs& a) {
  if (a.persistent && a.dual)  // (its ring has 8 slots whatever SL is)
    hipLaunchKernelGGL((k_trace_d<OCC, AL, CNT, PL>), a.pgrid, dim3(TB), 0, a.s, a.sc, a.rays, a.M, a.stride,
                       a.inst_id, a.cnt_nodes, a.cnt_tris, a.err, a.perm, a.counter, a.refill, a.node_visits);
  else if (a.persistent)
    hipLaunchKernelGGL((k_trace_p<OCC, AL, CNT, SL, PL>), a.pgrid, dim3(TB), 0, a.s, a.sc, a.rays, a.M, a.stride,
                       a.inst_id, a.cnt_nodes, a.cnt_tris, a.err, a.perm, a.counter, a.refill, a.node_visits,
                       nullptr, 0u);
  else
    hipLaunchKernelGGL((k_trace<OCC, AL, CNT, SL, PL>), a.grid, dim3(TB), 0, a.s, a.sc, a.rays, a.M, a.stride,
                       a.inst_id, a.cnt_nodes, a.cnt_tris, a.err, a.perm, a.node_visits);
}

template <bool OCC, bool AL, bool CNT, bool PL>
void launch_slots(const LaunchArgs& a) {
  if (a.lds_slots <= 0) launch_one<OCC, AL, CNT, 0, PL>(a);
  else if (a.lds_slots <= 8) launch_one<OCC, AL, CNT, 8, PL>(a);
  else if (a.lds_slots <= 12) launch_one<OCC, AL, CNT, 12, PL>(a);
  else launch_one<OCC, AL, CNT, 16, PL>(a);
}

template <bool PL>
void launch_pl(const LaunchArgs& a, bool count, bool occluded, bool aligned) {
  if (count) {
    if (occluded) launch_slots<true, false, true, PL>(a);
    else launch_slots<false, false, true, PL>(a);
  } else if (occluded) {
    if (aligned) launch_slots<true, true, false, PL>(a);
    else launch_slots<true, false, false, PL>(a);
  } else {
    if (aligned) launch_slots<false, true, false, PL>(a);
    else launch_slots<false, false, false, PL>(a);
  }
}

void launch_trace(const DevScene& sc, void* rays, uint32_t M, uint64_t stride, bool occluded, uint32_t inst_id,
                  uint32_t* cnt_nodes, uint32_t* cnt_tris, uint32_t* err, hipStream_t s, const TraceParams& tp,
                  RaySorter* sorter, uint32_t* node_visits, uint32_t* counter) {
  if (M == 0 || sc.num_nodes == 0) return;
  const bool aligned = ((uintptr_t)rays % 16 == 0) && (stride % 16 == 0);
  const dim3 grid(div_up(M, TB)), block(TB);
  char* rp = (char*)rays;
  const bool count = cnt_nodes != nullptr;
  const uint32_t* perm = nullptr;
  if (tp.sort_rays && sorter && M >= tp.sort_min) {
    sorter->keys.reserve((size_t)M * 4);
    sorter->vals.reserve((size_t)M * 4);
    hipLaunchKernelGGL(k_ray_keys, grid, block, 0, s, rp, M, stride, sc.lo[0], sc.lo[1], sc.lo[2], sc.scale[0],
                       sc.scale[1], sc.scale[2], sorter->keys.as<uint32_t>(), sorter->vals.as<uint32_t>());
    sorter->sort.sort_pairs(sorter->keys.as<uint32_t>(), sorter->vals.as<uint32_t>(), M, 0, 32, s);
    perm = sorter->vals.as<uint32_t>();
  }
  const bool persistent = tp.persistent && (sorter || counter);
  const int rmin = occluded ? tp.refill_min_occ : tp.refill_min;
  const int refill = rmin < 1 ? 1 : (rmin > 64 ? 64 : rmin);  // 64 idle lanes always refill
  dim3 pgrid(1);
  if (persistent) {
    if (!counter) {
      sorter->counter.reserve(256);
      counter = sorter->counter.as<uint32_t>();
    }
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, counter);
    const uint32_t want = (uint32_t)tp.cus * (uint32_t)tp.blocks_per_cu;
    pgrid = dim3(want < grid.x ? want : grid.x);
  }
  LaunchArgs la{sc, rp, M, stride, inst_id, cnt_nodes, cnt_tris, err, perm,
                persistent ? counter : nullptr, refill, count ? node_visits : nullptr,
                persistent, grid, pgrid, tp.lds_slots, s, tp.dual == 1 || (tp.dual == 2 && occluded)};
  if (sc.robust) launch_pl<true>(la, count, occluded, aligned);
  else launch_pl<false>(la, count, occluded, aligned);
  RTG_HIP(hipGetLastError());
}

// (LDS stack rings of 8 or of 12 entries per lane: lds_slots up to 8, or more)
template <bool OCC, bool AL, bool PL>
void launch_batch_one(const DevScene& sc, const BatchItem* items, uint32_t nitems, uint32_t M, uint32_t* err,
                      uint32_t* counter, int refill, dim3 pgrid, hipStream_t s, int lds_slots) {
  if (lds_slots <= 8)
    hipLaunchKernelGGL((k_trace_p<OCC, AL, false, 8, PL, true>), pgrid, dim3(TB), 0, s, sc, (char*)nullptr, M,
                       (uint64_t)0, 0xffffffffu, (uint32_t*)nullptr, (uint32_t*)nullptr, err,
                       (const uint32_t*)nullptr, counter, refill, (uint32_t*)nullptr, items, nitems);
  else
    hipLaunchKernelGGL((k_trace_p<OCC, AL, false, 12, PL, true>), pgrid, dim3(TB), 0, s, sc, (char*)nullptr, M,
                       (uint64_t)0, 0xffffffffu, (uint32_t*)nullptr, (uint32_t*)nullptr, err,
                       (const uint32_t*)nullptr, counter, refill, (uint32_t*)nullptr, items, nitems);
}

void launch_trace_batch(const DevScene& sc, const BatchItem* items, uint32_t nitems, uint32_t M, bool occluded,
                        bool aligned, uint32_t* err, uint32_t* counter, hipStream_t s, const TraceParams& tp) {
  if (M == 0 || nitems == 0 || sc.num_nodes == 0) return;
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, counter);
  const int rmin = occluded ? tp.refill_min_occ : tp.refill_min;
  const int refill = rmin < 1 ? 1 : (rmin > 64 ? 64 : rmin);
  const uint32_t want = (uint32_t)tp.cus * (uint32_t)tp.blocks_per_cu;
  const uint32_t blocks = div_up(M, TB);
  const dim3 pgrid(want < blocks ? want : blocks);
  if (sc.robust) {
    if (occluded) {
      if (aligned) launch_batch_one<true, true, true>(sc, items, nitems, M, err, counter, refill, pgrid, s, tp.lds_slots);
      else launch_batch_one<true, false, true>(sc, items, nitems, M, err, counter, refill, pgrid, s, tp.lds_slots);
    } else {
      if (aligned) launch_batch_one<false, true, true>(sc, items, nitems, M, err, counter, refill, pgrid, s, tp.lds_slots);
      else launch_batch_one<false, false, true>(sc, items, nitems, M, err, counter, refill, pgrid, s, tp.lds_slots);
    }
  } else {
    if (occluded) {
      if (aligned) launch_batch_one<true, true, false>(sc, items, nitems, M, err, counter, refill, pgrid, s, tp.lds_slots);
      else launch_batch_one<true, false, false>(sc, items, nitems, M, err, counter, refill, pgrid, s, tp.lds_slots);
    } else {
      if (aligned) launch_batch_one<false, true, false>(sc, items, nitems, M, err, counter, refill, pgrid, s, tp.lds_slots);
      else launch_batch_one<false, false, false>(sc, items, nitems, M, err, counter, refill, pgrid, s, tp.lds_slots);
    }
  }
  RTG_HIP(hipGetLastError());
}

void launch_trace_inst(const DevTop& top, void* rays, uint32_t M, uint64_t stride, bool occluded, uint32_t ctx_inst,
                       uint32_t* err, hipStream_t s, const TraceParams& tp, RaySorter* sorter, uint32_t* counter) {
  if (M == 0 || top.num_nodes == 0) return;
  const bool aligned = ((uintptr_t)rays % 16 == 0) && (stride % 16 == 0);
  const uint32_t blocks = div_up(M, TB);
  const uint32_t want = (uint32_t)tp.cus * (uint32_t)tp.blocks_per_cu;
  const dim3 grid(want < blocks ? want : blocks), block(TB);
  const int rmin = occluded ? tp.refill_min_occ_inst : tp.refill_min;
  const int refill = rmin < 1 ? 1 : (rmin > 64 ? 64 : rmin);
  if (!counter) {
    sorter->counter.reserve(256);
    counter = sorter->counter.as<uint32_t>();
  }
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, counter);
  char* rp = (char*)rays;
  if (occluded) {
    if (aligned) hipLaunchKernelGGL((k_trace_inst<true, true>), grid, block, 0, s, top, rp, M, stride, ctx_inst, err, counter, refill);
    else hipLaunchKernelGGL((k_trace_inst<true, false>), grid, block, 0, s, top, rp, M, stride, ctx_inst, err, counter, refill);
  } else {
    if (aligned) hipLaunchKernelGGL((k_trace_inst<false, true>), grid, block, 0, s, top, rp, M, stride, ctx_inst, err, counter, refill);
    else hipLaunchKernelGGL((k_trace_inst<false, false>), grid, block, 0, s, top, rp, M, stride, ctx_inst, err, counter, refill);
  }
  RTG_HIP(hipGetLastError());
}

void launch_trace_inst_counted(const DevTop& top, void* rays, uint32_t M, uint64_t stride, uint32_t* err,
                               uint32_t* counter, uint32_t* top_visits, uint32_t* obj_visits, uint32_t target_obj,
                               hipStream_t s, const TraceParams& tp) {
  if (M == 0 || top.num_nodes == 0) return;
  const bool aligned = ((uintptr_t)rays % 16 == 0) && (stride % 16 == 0);
  const uint32_t blocks = div_up(M, TB);
  const uint32_t want = (uint32_t)tp.cus * (uint32_t)tp.blocks_per_cu;
  const dim3 grid(want < blocks ? want : blocks), block(TB);
  const int refill = tp.refill_min < 1 ? 1 : (tp.refill_min > 64 ? 64 : tp.refill_min);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, counter);
  char* rp = (char*)rays;
  if (aligned)
    hipLaunchKernelGGL((k_trace_inst<false, true, true>), grid, block, 0, s, top, rp, M, stride, 0xffffffffu, err,
                       counter, refill, top_visits, obj_visits, target_obj);
  else
    hipLaunchKernelGGL((k_trace_inst<false, false, true>), grid, block, 0, s, top, rp, M, stride, 0xffffffffu, err,
                       counter, refill, top_visits, obj_visits, target_obj);
  RTG_HIP(hipGetLastError());
}

namespace {
constexpr int NOALLOC_MAX_GPUS = 64;
std::mutex g_noalloc_mtx;
uint32_t* g_noalloc_ptr[NOALLOC_MAX_GPUS] = {};

__global__ void k_fill_words(uint32_t* p, uint32_t n, uint32_t v) {
  const uint32_t i = blockIdx.x * TB + threadIdx.x;
  if (i < n) p[i] = v;
}
}  // namespace

void noalloc_setup(int ordinal) {
  if (ordinal < 0 || ordinal >= NOALLOC_MAX_GPUS) throw hip_error(hipErrorInvalidDevice, "GPU ordinal out of range");
  std::lock_guard<std::mutex> g(g_noalloc_mtx);
  if (g_noalloc_ptr[ordinal]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr)
    throw hip_error(e != hipSuccess ? e : hipErrorInvalidSymbol,
                    std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_noalloc_ptr[ordinal] = static_cast<uint32_t*>(p);
}

uint32_t* noalloc_scratch(int ordinal) {
  if (ordinal < 0 || ordinal >= NOALLOC_MAX_GPUS) return nullptr;
  std::lock_guard<std::mutex> g(g_noalloc_mtx);
  return g_noalloc_ptr[ordinal];
}

void fill_words(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_words, dim3(div_up(n, TB)), dim3(TB), 0, s, p, n, v);
  RTG_HIP(hipGetLastError());
}

}  // namespace rtg

