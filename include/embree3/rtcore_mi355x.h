/* Extensions of the MI355X triangle ray-query library.
 *
 * None of these has a counterpart in the reference API. They exist for
 * callers that keep rays resident in GPU memory (a renderer's wavefront
 * queue, a benchmark, a training loop) and for the workload generators the
 * tests and benchmarks use. Every function reports failures through the
 * device error slot like the standard API. */
#pragma once

#include "rtcore.h"

RTC_NAMESPACE_BEGIN

/* Device memory on the device's GPU (hipMalloc / hipFree). */
RTC_API void* rtcgDeviceMalloc(RTCDevice device, size_t bytes);
RTC_API void rtcgDeviceFree(RTCDevice device, void* ptr);

/* Synchronous copy between any two host or device pointers (hipMemcpyDefault). */
RTC_API void rtcgMemcpy(RTCDevice device, void* dst, const void* src, size_t bytes);

/* Same contract as rtcIntersect1M / rtcOccluded1M, but the records live in
 * device memory and the call is asynchronous on `hipStream` (a hipStream_t;
 * NULL selects the calling thread's stream of this device). */
RTC_API void rtcgIntersect1MAsync(RTCScene scene, const struct RTCIntersectContext* context, struct RTCRayHit* d_rayhit,
                                  unsigned int M, size_t byteStride, void* hipStream);
RTC_API void rtcgOccluded1MAsync(RTCScene scene, const struct RTCIntersectContext* context, struct RTCRay* d_ray,
                                 unsigned int M, size_t byteStride, void* hipStream);

/* Message of the calling thread's most recent error (any device); the
 * pointer stays valid until the thread's next error. */
RTC_API const char* rtcgGetErrorString(void);

/* Waits for the calling thread's stream of this device. */
RTC_API void rtcgSynchronize(RTCDevice device);

/* Timing harness: `steps` times, copies M records from d_src to d_rays
 * (untimed) and traces d_rays with HIP events recorded around the trace
 * kernel on the stream it runs on; per-step milliseconds go to ms_out. */
RTC_API void rtcgTimeTrace(RTCScene scene, void* d_rays, const void* d_src, unsigned int M, size_t byteStride,
                           int occluded, int steps, float* ms_out);

/* Timing harness for the scene build: rebuilds the committed scene `steps`
 * times and reports per-build milliseconds (wall clock, build fully
 * synchronised). */
RTC_API void rtcgTimeBuild(RTCScene scene, int steps, float* ms_out);

/* Counting trace: like rtcgIntersect1MAsync/rtcgOccluded1MAsync (synchronous)
 * but also writes per-ray node visits and triangle tests to the device
 * arrays d_nodes[M] and d_tris[M]. */
RTC_API void rtcgTraceCounted(RTCScene scene, void* d_rays, unsigned int M, size_t byteStride, int occluded,
                              unsigned int* d_nodes, unsigned int* d_tris);

/* Built acceleration structure summary. */
struct RTCGSceneStats
{
  unsigned int num_prims;      /* valid triangles in the BVH */
  unsigned int num_nodes;      /* 8-wide nodes */
  unsigned int max_depth;      /* node levels */
  unsigned int node_bytes;     /* bytes per node record */
  unsigned int tri_bytes;      /* bytes per triangle record */
  unsigned int max_stack;      /* traversal stack entries the deepest path can need */
  double sah;                  /* SAH statistic, leaf cost per block of 4 triangles */
  double sah_tri;              /* SAH statistic, leaf cost per triangle */
  double build_ms;             /* last commit's build time */
};
RTC_API void rtcgGetSceneStats(RTCScene scene, struct RTCGSceneStats* stats);

/* Copies the device BVH to host memory: nodes (80 B each) and triangle
 * records (48 B each); sizes from rtcgGetSceneStats. Either pointer may be NULL. */
RTC_API void rtcgReadBVH(RTCScene scene, void* nodes, void* tris);

/* ---- workload generators (host memory, deterministic) ---------------- */

/* UV sphere of the reference tutorial generator with numPhi rings:
 * numTheta = 2*numPhi, vertices numTheta*(numPhi+1), triangles
 * 2*numTheta*(numPhi-1). Returns the triangle count; pass NULL buffers to
 * query sizes (vertex count written to *numVertices). */
RTC_API unsigned int rtcgGenSphere(unsigned int numPhi, float cx, float cy, float cz, float radius,
                                   float* vertices, unsigned int* indices, unsigned int* numVertices);

/* Pinhole camera rays (closest-hit records, tnear 0, tfar +inf) for a
 * width x height image from `org` looking down -z with the given vertical
 * field of view in degrees. */
RTC_API void rtcgGenPinholeRays(unsigned int width, unsigned int height, float ox, float oy, float oz,
                                float fovDeg, struct RTCRayHit* rays);

/* Random triangle soup: centroids uniform in [0,1]^3, vertex offsets
 * uniform in [-h,h]^3; 3 unique vertices per triangle (indexed). */
RTC_API void rtcgGenSoup(unsigned int numTris, unsigned long long seed, float h, float* vertices,
                         unsigned int* indices);

/* Incoherent rays: origins uniform in [lo,hi]^3, directions uniform on the
 * sphere; tnear 0 and the given tfar. */
RTC_API void rtcgGenRandomRays(unsigned int numRays, unsigned long long seed, float lo, float hi, float tfar,
                               struct RTCRayHit* rays);

RTC_NAMESPACE_END
