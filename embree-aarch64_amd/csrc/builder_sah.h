// Top-down binned-SAH binary tree over the valid primitive references, built
// level by level on the GPU. Its output has the same form as the radix tree
// of the Morton builder (child / parent / range arrays over final primitive
// positions), so the bottom-up SAH refit and the 8-wide collapse run on
// either tree unchanged.
#pragma once
#include <vector>

#include "bvh_types.h"
#include "gpu_util.h"
#include "mlp.h"
#include "sort.h"

namespace rtg {

// Pinned host slots and events through which the host reads a level's
// counters while the next level already runs.
struct LevelReadback {
  static constexpr int K = 2;
  void* host = nullptr;
  size_t bytes = 0;  // per slot
  hipEvent_t ev[K] = {};
  void ensure(size_t b) {
    if (!ev[0])
      for (hipEvent_t& e : ev) RTG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    if (b <= bytes) return;
    if (host) (void)hipHostFree(host);
    host = nullptr;
    bytes = 0;
    RTG_HIP(hipHostMalloc(&host, K * b, hipHostMallocDefault));
    bytes = b;
  }
  void* slot(int k) const { return (char*)host + (size_t)(k % K) * bytes; }
  ~LevelReadback() {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    if (host) (void)hipHostFree(host);
  }
  LevelReadback() = default;
  LevelReadback(const LevelReadback&) = delete;
  LevelReadback& operator=(const LevelReadback&) = delete;
};

struct SahScratch {
  DevBuf idx[2], blo[2], bhi[2], sid[2], segs[3][3], tiny, bins, splits, flags, scan, counters, roots;
  Scanner scanner;
  LevelReadback rb;
};

// One tree of the forest: positions [first, first+count), count >= 2.
struct SahRoot {
  uint32_t first, count;
  float clo[3], chi[3];  // bounds of lo+hi over its references
};

struct SahTreeOut {
  std::vector<uint32_t> roots;  // inner node id of each tree's root
  uint32_t levels = 0;          // top-down levels run
};

// Inputs: n valid references, `vals` holding their primref indices and
// slo/shi their boxes in that order, and the trees to build (disjoint
// position ranges of at least two references). Outputs: vals reordered so
// that every inner node covers a contiguous position range; child/
// parent_int/parent_leaf/range filled for the inner nodes (an inner node's
// id is its split position, so ids lie in [0, n-2]; ids no tree uses keep
// parent_int = REF_EMPTY).
// `learned` (may be null) replaces the half-area term of the split cost.
// Segments of at most median_below (<= 64) references are split at their
// median position all the way down instead of by binned SAH.
void build_sah_tree(uint32_t n, uint32_t* vals, const float4* slo, const float4* shi,
                    const std::vector<SahRoot>& trees, uint32_t* child, uint32_t* parent_int,
                    uint32_t* parent_leaf, uint2* range, uint8_t* depth, const LearnedCost* learned,
                    uint32_t median_below, SahScratch& S, SahTreeOut& out, hipStream_t s);

}  // namespace rtg
