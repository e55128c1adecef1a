// Traversal kernel launcher.
#pragma once
#include "bvh_types.h"
#include "gpu_util.h"

namespace rtg {

// Traces M ray records (RTCRayHit for closest hit, RTCRay for occlusion)
// spaced `stride` bytes apart in device memory, in place. cnt_nodes /
// cnt_tris (optional, M entries each) receive per-ray node visits and
// triangle tests. err receives bit 0 on a traversal stack overflow.
void launch_trace(const DevScene& sc, void* rays, uint32_t M, uint64_t stride, bool occluded, uint32_t inst_id,
                  uint32_t* cnt_nodes, uint32_t* cnt_tris, uint32_t* err, hipStream_t s);

}  // namespace rtg
