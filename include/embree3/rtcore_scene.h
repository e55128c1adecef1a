/* Scene objects and ray queries of the triangle ray-query C API. */
#pragma once

#include "rtcore_device.h"
#include "rtcore_geometry.h"

RTC_NAMESPACE_BEGIN

struct RTCRay;
struct RTCRayHit;

/* Scene flags (reference: rtcore_scene.h:18-25). ROBUST selects the
 * Pluecker triangle test; DYNAMIC and COMPACT are accepted and do not change
 * the GPU layout. */
enum RTCSceneFlags
{
  RTC_SCENE_FLAG_NONE                    = 0,
  RTC_SCENE_FLAG_DYNAMIC                 = (1 << 0),
  RTC_SCENE_FLAG_COMPACT                 = (1 << 1),
  RTC_SCENE_FLAG_ROBUST                  = (1 << 2),
  RTC_SCENE_FLAG_CONTEXT_FILTER_FUNCTION = (1 << 3)
};

/* Creates a scene (reference: rtcore_scene.h:28, rtcore.cpp:187). */
RTC_API RTCScene rtcNewScene(RTCDevice device);

/* Device of a scene, retained (reference: rtcore_scene.h:32). */
RTC_API RTCDevice rtcGetSceneDevice(RTCScene hscene);

/* Reference counting (reference: rtcore_scene.h:35-38). */
RTC_API void rtcRetainScene(RTCScene scene);
RTC_API void rtcReleaseScene(RTCScene scene);

/* Attaches a geometry; IDs are handed out in attach order 0,1,...
 * (reference: rtcore_scene.h:42, rtcore.cpp:1718-1731). */
RTC_API unsigned int rtcAttachGeometry(RTCScene scene, RTCGeometry geometry);

/* Attaches with a caller-chosen ID (reference: rtcore_scene.h:45). */
RTC_API void rtcAttachGeometryByID(RTCScene scene, RTCGeometry geometry, unsigned int geomID);

/* Detaches (reference: rtcore_scene.h:48). */
RTC_API void rtcDetachGeometry(RTCScene scene, unsigned int geomID);

/* Geometry lookup, not retained (reference: rtcore_scene.h:51). */
RTC_API RTCGeometry rtcGetGeometry(RTCScene scene, unsigned int geomID);

/* Builds the GPU BVH of all enabled, committed geometries (reference:
 * rtcore_scene.h:55, rtcore.cpp:259-267, kernels/common/scene.cpp:655-812).
 * Blocking; the BVH stays resident in device memory. */
RTC_API void rtcCommitScene(RTCScene scene);

/* Joins a commit (reference: rtcore_scene.h:58). The build runs on the GPU,
 * so a joining thread just performs the same blocking commit. */
RTC_API void rtcJoinCommitScene(RTCScene scene);

typedef bool (*RTCProgressMonitorFunction)(void* ptr, double n);

/* Progress monitor (reference: rtcore_scene.h:65, scene.cpp:965-972): called
 * between build phases with the completed fraction; returning false cancels
 * the commit with RTC_ERROR_CANCELLED. */
RTC_API void rtcSetSceneProgressMonitorFunction(RTCScene scene, RTCProgressMonitorFunction progress, void* ptr);

/* Build quality (reference: rtcore_scene.h:68): LOW selects the Morton/LBVH
 * builder, MEDIUM and HIGH the binned-SAH builder. */
RTC_API void rtcSetSceneBuildQuality(RTCScene scene, enum RTCBuildQuality quality);

/* Scene flags (reference: rtcore_scene.h:71-74). */
RTC_API void rtcSetSceneFlags(RTCScene scene, enum RTCSceneFlags flags);
RTC_API enum RTCSceneFlags rtcGetSceneFlags(RTCScene scene);

/* Scene bounds after commit (reference: rtcore_scene.h:77). */
RTC_API void rtcGetSceneBounds(RTCScene scene, struct RTCBounds* bounds_o);

/* Single-ray closest hit (reference: rtcore_scene.h:96, rtcore.cpp:479). */
RTC_API void rtcIntersect1(RTCScene scene, struct RTCIntersectContext* context, struct RTCRayHit* rayhit);

/* Stream closest hit over M AoS records spaced byteStride bytes apart
 * (reference: rtcore_scene.h:108, rtcore.cpp:595-623). The records may live
 * in host memory (copied over PCIe) or in device memory of this GPU
 * (traced in place). Rays with !(tnear <= tfar) are left untouched; on a hit
 * tfar, Ng, u, v, primID, geomID and instID[0] are written; on a miss nothing
 * is written. */
RTC_API void rtcIntersect1M(RTCScene scene, struct RTCIntersectContext* context, struct RTCRayHit* rayhit, unsigned int M, size_t byteStride);

/* Single-ray occlusion (reference: rtcore_scene.h:120, rtcore.cpp:733). */
RTC_API void rtcOccluded1(RTCScene scene, struct RTCIntersectContext* context, struct RTCRay* ray);

/* Stream occlusion (reference: rtcore_scene.h:132, rtcore.cpp:848-874). An
 * occluded ray gets tfar = -inf; nothing else is written. */
RTC_API void rtcOccluded1M(RTCScene scene, struct RTCIntersectContext* context, struct RTCRay* ray, unsigned int M, size_t byteStride);

/* Packet and pointer-SoA forms (reference: rtcore_scene.h:99-141,
 * rtcore.cpp:500-593, 750-846). Lanes with valid[i] == 0 (packets) or
 * !(tnear <= tfar) are left untouched; the others are traced as by
 * rtcIntersect1M / rtcOccluded1M. Records in host memory. */
RTC_API void rtcIntersect4(const int* valid, RTCScene scene, struct RTCIntersectContext* context, struct RTCRayHit4* rayhit);
RTC_API void rtcIntersect8(const int* valid, RTCScene scene, struct RTCIntersectContext* context, struct RTCRayHit8* rayhit);
RTC_API void rtcIntersect16(const int* valid, RTCScene scene, struct RTCIntersectContext* context, struct RTCRayHit16* rayhit);
RTC_API void rtcIntersect1Mp(RTCScene scene, struct RTCIntersectContext* context, struct RTCRayHit** rayhit, unsigned int M);
RTC_API void rtcIntersectNM(RTCScene scene, struct RTCIntersectContext* context, struct RTCRayHitN* rayhit, unsigned int N, unsigned int M, size_t byteStride);
RTC_API void rtcIntersectNp(RTCScene scene, struct RTCIntersectContext* context, const struct RTCRayHitNp* rayhit, unsigned int N);
RTC_API void rtcOccluded4(const int* valid, RTCScene scene, struct RTCIntersectContext* context, struct RTCRay4* ray);
RTC_API void rtcOccluded8(const int* valid, RTCScene scene, struct RTCIntersectContext* context, struct RTCRay8* ray);
RTC_API void rtcOccluded16(const int* valid, RTCScene scene, struct RTCIntersectContext* context, struct RTCRay16* ray);
RTC_API void rtcOccluded1Mp(RTCScene scene, struct RTCIntersectContext* context, struct RTCRay** ray, unsigned int M);
RTC_API void rtcOccludedNM(RTCScene scene, struct RTCIntersectContext* context, struct RTCRayN* ray, unsigned int N, unsigned int M, size_t byteStride);
RTC_API void rtcOccludedNp(RTCScene scene, struct RTCIntersectContext* context, const struct RTCRayNp* ray, unsigned int N);

RTC_NAMESPACE_END
