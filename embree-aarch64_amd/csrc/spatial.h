// Spatial splits of the HIGH-quality build: before the binned-SAH tree is
// built, triangles whose boxes are much larger than the triangles
// themselves are cut into several references, each with the box of its
// piece of the triangle, so that the tree can separate what one box would
// have held together. This takes the place of the reference's spatial-split
// binning (kernels/builders/heuristic_spatial_array.h,
// BVH8Triangle4SceneBuilderFastSpatialSAH at RTC_BUILD_QUALITY_HIGH,
// kernels/bvh/bvh8_factory.cpp:867-870) with a pre-pass that suits a
// GPU's level-parallel build (after Karras and Aila, "Fast Parallel
// Construction of High-Quality Bounding Volume Hierarchies", HPG 2013,
// section 5). A triangle may then sit in several leaves; the triangle test
// is the same, so hits are the same.
#pragma once
#include "bvh_types.h"
#include "gpu_util.h"
#include "sort.h"

namespace rtg {

constexpr uint32_t SPLIT_MAX_PIECES = 16;  // references one triangle may become

struct SpatialScratch {
  DevBuf prio, count, offset, lo4, hi4, vals, sample;
  PinnedBuf sample_host;
  Scanner scanner;
};

// n valid references in sorted order: vals[j] indexes lo4/hi4 (boxes, .w
// = geometry index and primitive index, as the primitive references hold
// them). Splits up to budget * n extra references, given out in proportion
// to (2^-level g)^(1/3): g the half area the first cut saves (the box's less
// its pieces'), counted only when at least 30 % of the box's, level the
// depth of the coarsest plane of the scene's dyadic grid that the cut uses
// (0: the scene's middle). On
// return vals/lo4/hi4 hold the references (vals[r] = r, pieces of one
// triangle adjacent, in the triangles' order) and the result is their
// count. glo/ghi: the scene box.
uint32_t spatial_presplit(uint32_t n, float budget, const GeomDesc* geoms, const float glo[3], const float ghi[3],
                          DevBuf& vals, DevBuf& lo4, DevBuf& hi4, SpatialScratch& S, hipStream_t s);

}  // namespace rtg
