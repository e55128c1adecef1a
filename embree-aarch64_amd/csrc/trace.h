// Traversal kernel launcher.
#pragma once
#include "bvh_types.h"
#include "gpu_util.h"
#include "sort.h"

namespace rtg {

// Launch-shape parameters of the traversal (chosen per device, tunable).
struct TraceParams {
  // LDS stack entries per lane (0: private stack, 8, 12 or 16; the kernel
  // with a triangle cursor always has 8). On the 2^20-triangle soup with
  // 2^24 incoherent rays, closest hit: 12 gave 2398 Mrays/s, 8 2345 and 16
  // 2089 (16 fits 5 blocks of 256 lanes per CU, 12 and 8 fit 6).
  int lds_slots = 12;
  int sort_rays = 0;        // trace in origin-Morton order (a permutation; records stay in place)
  uint32_t sort_min = 65536;  // smallest stream that is reordered
  int persistent = 1;       // persistent waves pulling rays from a counter
  // Idle lanes of a wave that trigger a refill. A refill stalls the wave
  // for one ray-record load (and, once per 256 rays, the global counter
  // atomic); on the 2^20-triangle soup with 2^24 incoherent rays 10 gave
  // 1936 Mrays/s closest hit (8-14 within 1 %, 24: 1854, 32: 1818) and 24
  // gave 3795 Mrays/s occlusion (48: 3339).
  int refill_min = 10;      // closest hit
  int refill_min_occ = 24;  // occlusion rays, which finish sooner
  // Occlusion rays through a two-level scene: 10,000 instances, 12.6 M
  // shadow rays gave 1169 Mrays/s at 8-12 against 1080 at 24 and 1129 at 16.
  int refill_min_occ_inst = 12;
  // Persistent kernel with a separate triangle cursor per lane: 0 never,
  // 1 always, 2 for occlusion only. On the 2^20-triangle soup with 2^24
  // incoherent rays it gave 3934 vs 3789 Mrays/s occlusion but 1724 vs 1939
  // closest hit (speculative node visits before a pending leaf's hits).
  int dual = 2;
  int cus = 256;            // compute units (set from the device)
  int blocks_per_cu = 8;    // resident 256-thread blocks per compute unit for the persistent grid
};

// One ray stream of a batched traversal (device memory).
struct BatchItem {
  char* rays;        // records, `stride` bytes apart
  uint64_t stride;
  uint32_t first;    // slot of its first ray in the concatenation of all streams
  uint32_t count;
  uint32_t root;     // root node of its scene
  uint32_t pad;
};

// Per-thread buffers of the ray reordering.
struct RaySorter {
  DevBuf keys, vals, counter;
  RadixSorter sort;
};

// Traces M ray records (RTCRayHit for closest hit, RTCRay for occlusion)
// spaced `stride` bytes apart in device memory, in place. cnt_nodes /
// cnt_tris (optional, M entries each) receive per-ray node visits and
// triangle tests. err receives bit 0 on a traversal stack overflow.
// node_visits (optional, with the counters): one atomic increment per node
// visit, indexed like the QNode8 array.
// counter (optional): the persistent kernel's ray counter (one device word);
// without it the sorter's is used, and with neither the grid covers every ray.
void launch_trace(const DevScene& sc, void* rays, uint32_t M, uint64_t stride, bool occluded, uint32_t inst_id,
                  uint32_t* cnt_nodes, uint32_t* cnt_tris, uint32_t* err, hipStream_t s,
                  const TraceParams& tp, RaySorter* sorter, uint32_t* node_visits = nullptr,
                  uint32_t* counter = nullptr);

// One persistent launch over several ray streams of scenes that share sc's
// node and triangle arrays (a batched build). items (device memory) sorted
// by first; M = total rays. Records traced in place. counter: one device
// word, cleared by a kernel (the sequence can be captured in a hipGraph).
void launch_trace_batch(const DevScene& sc, const BatchItem* items, uint32_t nitems, uint32_t M, bool occluded,
                        bool aligned, uint32_t* err, uint32_t* counter, hipStream_t s, const TraceParams& tp);

// Two-level traversal of a scene with instances (top tree over InstRec
// leaves, object trees in top.objects); otherwise as launch_trace. ctx_inst
// is the query context's instID[0]; counter (one device word) or, if null,
// the sorter's supplies the ray counter.
void launch_trace_inst(const DevTop& top, void* rays, uint32_t M, uint64_t stride, bool occluded, uint32_t ctx_inst,
                       uint32_t* err, hipStream_t s, const TraceParams& tp, RaySorter* sorter,
                       uint32_t* counter = nullptr);

// Closest-hit two-level traversal that also counts the rays that enter each
// node of the top tree (top_visits, one word per top node, may be null) and
// of object target_obj's tree (obj_visits, one word per node of that object,
// summed over all its instances; may be null).
void launch_trace_inst_counted(const DevTop& top, void* rays, uint32_t M, uint64_t stride, uint32_t* err,
                               uint32_t* counter, uint32_t* top_visits, uint32_t* obj_visits, uint32_t target_obj,
                               hipStream_t s, const TraceParams& tp);

// Scratch of the no-allocation queries: word 0 is the persistent kernel's
// ray counter, word 1 the traversal error bits; one fixed array per GPU
// (NOALLOC_SCRATCH_WORDS words), held until the process ends.
constexpr uint32_t NOALLOC_SCRATCH_WORDS = 64;
constexpr uint32_t NOALLOC_MAX_RAYS = 1u << 31;  // the counter word's range, with room for every wave's overshoot
// Looks up the device address of this GPU's copy of the scratch array; the
// GPU `ordinal` must be current. Once per GPU (later calls return at once);
// throws hip_error when the lookup fails.
void noalloc_setup(int ordinal);
// Device address of GPU `ordinal`'s scratch array, or nullptr before noalloc_setup.
uint32_t* noalloc_scratch(int ordinal);
// Writes v to n words at p on stream s (a kernel, so the call can be captured).
void fill_words(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s);

}  // namespace rtg
