/* User BVH builder (reference: include/embree3/rtcore_builder.h:1-125,
 * kernels/common/rtcore_builder.cpp:344-430).
 *
 * The caller supplies primitive boxes and node/leaf callbacks; the library
 * builds the tree on the GPU (Morton radix tree for RTC_BUILD_QUALITY_LOW,
 * binned SAH otherwise) with the SAH-optimal collapse to the requested
 * branching factor, then walks it on the calling thread and calls
 * createNode / setNodeBounds / createLeaf / setNodeChildren for every node.
 * Types, field order and callback signatures are the reference's (ABI). */
#pragma once

#include "rtcore_scene.h"

RTC_NAMESPACE_BEGIN

/* Opaque BVH type (reference rtcore_builder.h:11) */
typedef struct RTCBVHTy* RTCBVH;

/* Input build primitives for the builder (reference rtcore_builder.h:14-20) */
struct RTC_ALIGN(32) RTCBuildPrimitive
{
  float lower_x, lower_y, lower_z;
  unsigned int geomID;
  float upper_x, upper_y, upper_z;
  unsigned int primID;
};

/* Opaque thread local allocator type (reference rtcore_builder.h:23) */
typedef struct RTCThreadLocalAllocatorTy* RTCThreadLocalAllocator;

/* Callbacks (reference rtcore_builder.h:26-39) */
typedef void* (*RTCCreateNodeFunction) (RTCThreadLocalAllocator allocator, unsigned int childCount, void* userPtr);
typedef void (*RTCSetNodeChildrenFunction) (void* nodePtr, void** children, unsigned int childCount, void* userPtr);
typedef void (*RTCSetNodeBoundsFunction) (void* nodePtr, const struct RTCBounds** bounds, unsigned int childCount, void* userPtr);
typedef void* (*RTCCreateLeafFunction) (RTCThreadLocalAllocator allocator, const struct RTCBuildPrimitive* primitives, size_t primitiveCount, void* userPtr);
typedef void (*RTCSplitPrimitiveFunction) (const struct RTCBuildPrimitive* primitive, unsigned int dimension, float position, struct RTCBounds* leftBounds, struct RTCBounds* rightBounds, void* userPtr);

/* Build flags (reference rtcore_builder.h:42-46) */
enum RTCBuildFlags
{
  RTC_BUILD_FLAG_NONE    = 0,
  RTC_BUILD_FLAG_DYNAMIC = (1 << 0),
};

enum RTCBuildConstants
{
  RTC_BUILD_MAX_PRIMITIVES_PER_LEAF = 32
};

/* Input for builders (reference rtcore_builder.h:54-79) */
struct RTCBuildArguments
{
  size_t byteSize;

  enum RTCBuildQuality buildQuality;
  enum RTCBuildFlags buildFlags;
  unsigned int maxBranchingFactor;
  unsigned int maxDepth;
  unsigned int sahBlockSize;
  unsigned int minLeafSize;
  unsigned int maxLeafSize;
  float traversalCost;
  float intersectionCost;

  RTCBVH bvh;
  struct RTCBuildPrimitive* primitives;
  size_t primitiveCount;
  size_t primitiveArrayCapacity;

  RTCCreateNodeFunction createNode;
  RTCSetNodeChildrenFunction setNodeChildren;
  RTCSetNodeBoundsFunction setNodeBounds;
  RTCCreateLeafFunction createLeaf;
  RTCSplitPrimitiveFunction splitPrimitive;
  RTCProgressMonitorFunction buildProgress;
  void* userPtr;
};

/* Default build settings (reference rtcore_builder.h:82-108) */
RTC_FORCEINLINE struct RTCBuildArguments rtcDefaultBuildArguments()
{
  struct RTCBuildArguments args;
  args.byteSize = sizeof(args);
  args.buildQuality = RTC_BUILD_QUALITY_MEDIUM;
  args.buildFlags = RTC_BUILD_FLAG_NONE;
  args.maxBranchingFactor = 2;
  args.maxDepth = 32;
  args.sahBlockSize = 1;
  args.minLeafSize = 1;
  args.maxLeafSize = RTC_BUILD_MAX_PRIMITIVES_PER_LEAF;
  args.traversalCost = 1.0f;
  args.intersectionCost = 1.0f;
  args.bvh = NULL;
  args.primitives = NULL;
  args.primitiveCount = 0;
  args.primitiveArrayCapacity = 0;
  args.createNode = NULL;
  args.setNodeChildren = NULL;
  args.setNodeBounds = NULL;
  args.createLeaf = NULL;
  args.splitPrimitive = NULL;
  args.buildProgress = NULL;
  args.userPtr = NULL;
  return args;
}

/* Creates a new BVH (reference rtcore_builder.h:111, rtcore_builder.cpp:344). */
RTC_API RTCBVH rtcNewBVH(RTCDevice device);

/* Builds a BVH and returns the root the callbacks created (reference
 * rtcore_builder.h:114, rtcore_builder.cpp:355-395). The primitive array is
 * reordered in place, and leaves receive pointers into it, as the
 * reference's binned-SAH builder does. Differences: maxBranchingFactor is
 * capped at 8 (the GPU collapse's node width); maxDepth is not enforced
 * (leaf sizes are); at RTC_BUILD_QUALITY_HIGH with a splitPrimitive
 * callback, primitives are first cut by that callback while the array has
 * room (primitiveArrayCapacity), the cut that most reduces the summed box
 * half areas first, at the coarsest plane of a dyadic grid over the
 * primitives' bounds across each box's largest axis, and the tree is built
 * over the pieces (a primitive may then sit in several leaves, as with the
 * reference's spatial splits, rtcore_builder.cpp:379-380); primitives with
 * non-finite or inverted bounds are left out of the tree; the callbacks run
 * on the calling thread. Memory from rtcThreadLocalAlloc stays valid until
 * the next build on the same RTCBVH or its release. */
RTC_API void* rtcBuildBVH(const struct RTCBuildArguments* args);

/* Allocates memory owned by the BVH (reference rtcore_builder.h:117). */
RTC_API void* rtcThreadLocalAlloc(RTCThreadLocalAllocator allocator, size_t bytes, size_t align);

/* Reference counting (reference rtcore_builder.h:120-123). */
RTC_API void rtcRetainBVH(RTCBVH bvh);
RTC_API void rtcReleaseBVH(RTCBVH bvh);

RTC_NAMESPACE_END
