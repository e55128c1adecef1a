// GPU BVH8 builder: primitive references -> Morton codes -> LDS radix sort ->
// binary radix tree (LBVH) -> bottom-up refit with an SAH leaf/inner cost ->
// top-down collapse into quantised 8-wide nodes and Moeller-ready triangle
// records.
#pragma once
#include <vector>

#include "builder_sah.h"
#include "mlp.h"
#include "bvh_types.h"
#include "gpu_util.h"
#include "sort.h"
#include "spatial.h"

namespace rtg {

enum BuilderKind { BUILDER_MORTON = 0, BUILDER_SAH = 1 };

struct BuildParams {
  int builder = BUILDER_SAH;   // binary tree: top-down binned SAH (default, as the reference's MEDIUM), or Morton radix tree
  uint32_t max_leaf = 8;       // max triangles per leaf (<= 15)
  uint32_t width = 8;          // children per node (2..8; the node record holds 8 slots)
  uint32_t median_below = 0;   // binned-SAH segments of at most this many references (<= 64) split at the median
  float cost_node = 1.0f;      // SAH cost of visiting one 8-wide node, per unit half area
  float cost_tri = 0.6f;       // SAH cost of one triangle test, per unit half area
  int quality = 1;             // RTCBuildQuality (LOW: Morton tree, HIGH: binned SAH, MEDIUM: `builder`)
  int robust = 0;              // leaves hold (v0, v1, v2) for the Pluecker test
  uint32_t min_leaf = 1;       // subtrees of at most this many primitives become leaves
  uint32_t block = 1;          // a leaf's cost counts blocks of this many primitives
  float split_budget = 0.0f;   // spatial splits: extra references up to this fraction of the triangles (0: none)
  const LearnedCost* learned = nullptr;  // binned-SAH split cost from a trained model (device data)
};

struct BuildStats {
  uint32_t num_prims = 0;      // triangle records: the valid triangles, more when spatial splits cut some
  uint32_t num_split = 0;      // records added by spatial splits
  uint32_t num_nodes = 0;
  uint32_t max_depth = 0;
  uint32_t max_stack = 0;      // traversal stack entries the deepest path can need
  float bounds_lo[3] = {0, 0, 0};
  float bounds_hi[3] = {0, 0, 0};
  double sah = 0;              // reference SAH statistic (leaf cost per block of 4)
  double sah_tri = 0;          // same with leaf cost per triangle
  double ms_total = 0;
  double ms_phase[8] = {0};    // primrefs, morton, sort, tree, refit, collapse
};

// One scene of a build: its root node and bounds.
struct SceneSlot {
  float bounds_lo[3] = {INFINITY, INFINITY, INFINITY};
  float bounds_hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  uint32_t num_prims = 0;
  uint32_t root = 0;  // QNode8 index (scene s of a batch has root s)
};

// Owns the device arrays of one build: one scene, or every scene of a
// batched build (they share the node and triangle arrays).
struct DeviceBVH {
  DevBuf nodes, tris, info;
  DevBuf insts;            // leaf records of an instance build (then tris is empty)
  bool instances = false;
  uint64_t generation = 0; // counts builds into these arrays (a build may move them)
  std::vector<SceneSlot> slots;
  DevScene view(uint32_t slot = 0) const {
    DevScene s;
    s.nodes = nodes.as<QNode8>();
    s.tris = tris.as<TriRec>();
    s.num_nodes = stats.num_nodes;
    s.num_tris = stats.num_prims;
    s.robust = robust;
    const SceneSlot& sl = slot < slots.size() ? slots[slot] : SceneSlot();
    s.root = sl.root;
    for (int k = 0; k < 3; k++) {
      const float ext = sl.bounds_hi[k] - sl.bounds_lo[k];
      s.lo[k] = sl.bounds_lo[k];
      s.scale[k] = ext > 0.0f ? 1023.99f / ext : 0.0f;
    }
    return s;
  }
  bool empty(uint32_t slot = 0) const { return stats.num_nodes == 0 || slot >= slots.size() || slots[slot].num_prims == 0; }
  BuildStats stats;  // of the whole build
  uint32_t robust = 0;
  // Exchanges the trees of two builds (a new generation for both).
  void swap_tree(DeviceBVH& o) {
    nodes.swap(o.nodes);
    tris.swap(o.tris);
    info.swap(o.info);
    insts.swap(o.insts);
    std::swap(instances, o.instances);
    std::swap(slots, o.slots);
    std::swap(stats, o.stats);
    std::swap(robust, o.robust);
    generation++;
    o.generation++;
  }
};

// Scratch reused across builds on one device.
struct BuildScratch {
  DevBuf geoms, lo4, hi4, keys, vals, slo, shi, child, parent_int, parent_leaf, range, nlo, nhi, cost,
      isleaf, split8, visit, queue, counters, sacc, depth, lvl_hist, lvl_order, leaf_src;
  RadixSorter sorter;
  SahScratch sah;
  SpatialScratch spatial;
  LevelReadback rb;  // collapse level counters
  PinnedBuf lvl_host;  // refit level histogram and offsets
};

// A binary tree over caller boxes with the SAH-optimal collapse decisions
// (for rtcBuildBVH): inner node i has children child[2i], child[2i+1]
// (REF_LEAF | position for one primitive), covers positions range[i],
// split8/isleaf as the collapse reads them; boxes are 8 floats (lo xyz, w,
// hi xyz, w); for a primitive box pbox[8p + 7] holds, as float bits, the
// input index of the primitive at position p.
struct UserTree {
  uint32_t n = 0;        // primitives placed (valid boxes)
  uint32_t root = 0;     // binary root reference
  uint32_t width = 8;
  std::vector<uint32_t> child, range;
  std::vector<uint8_t> split8, isleaf;
  std::vector<float> nbox, pbox;
};

// Builds over the given geometries (device pointers in GeomDesc). With
// nscenes > 1 every GeomDesc names its scene, the tree is the binned-SAH
// forest of all scenes in one pass, and scene s gets root node s.
// Throws hip_error on a HIP failure.
// With `user` set the build stops after the binary tree and hands it over
// (one scene of GEOM_INSTANCES boxes; no nodes or records are written).
void build_bvh(const std::vector<GeomDesc>& geoms, const BuildParams& p, BuildScratch& scratch,
               DeviceBVH& out, hipStream_t s, uint32_t nscenes = 1, UserTree* user = nullptr);

}  // namespace rtg
