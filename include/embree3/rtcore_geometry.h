/* Geometry objects of the triangle ray-query C API.
 *
 * Only RTC_GEOMETRY_TYPE_TRIANGLE is implemented: vertex buffer slot 0 in
 * RTC_FORMAT_FLOAT3, index buffer in RTC_FORMAT_UINT3
 * (reference: kernels/common/scene_triangle_mesh.cpp:35-80). Creating any
 * other type records RTC_ERROR_INVALID_OPERATION and returns NULL, the
 * reference's behaviour for a geometry type compiled out
 * (kernels/common/rtcore.cpp:1139-1250). */
#pragma once

#include "rtcore_buffer.h"

RTC_NAMESPACE_BEGIN

typedef struct RTCSceneTy* RTCScene;
typedef struct RTCGeometryTy* RTCGeometry;

/* Geometry types (reference: rtcore_geometry.h:18-51). */
enum RTCGeometryType
{
  RTC_GEOMETRY_TYPE_TRIANGLE = 0,
  RTC_GEOMETRY_TYPE_QUAD     = 1,
  RTC_GEOMETRY_TYPE_GRID     = 2,

  RTC_GEOMETRY_TYPE_SUBDIVISION = 8,

  RTC_GEOMETRY_TYPE_CONE_LINEAR_CURVE   = 15,
  RTC_GEOMETRY_TYPE_ROUND_LINEAR_CURVE  = 16,
  RTC_GEOMETRY_TYPE_FLAT_LINEAR_CURVE   = 17,

  RTC_GEOMETRY_TYPE_ROUND_BEZIER_CURVE  = 24,
  RTC_GEOMETRY_TYPE_FLAT_BEZIER_CURVE   = 25,
  RTC_GEOMETRY_TYPE_NORMAL_ORIENTED_BEZIER_CURVE  = 26,

  RTC_GEOMETRY_TYPE_ROUND_BSPLINE_CURVE = 32,
  RTC_GEOMETRY_TYPE_FLAT_BSPLINE_CURVE  = 33,
  RTC_GEOMETRY_TYPE_NORMAL_ORIENTED_BSPLINE_CURVE  = 34,

  RTC_GEOMETRY_TYPE_ROUND_HERMITE_CURVE = 40,
  RTC_GEOMETRY_TYPE_FLAT_HERMITE_CURVE  = 41,
  RTC_GEOMETRY_TYPE_NORMAL_ORIENTED_HERMITE_CURVE  = 42,

  RTC_GEOMETRY_TYPE_SPHERE_POINT = 50,
  RTC_GEOMETRY_TYPE_DISC_POINT = 51,
  RTC_GEOMETRY_TYPE_ORIENTED_DISC_POINT = 52,

  RTC_GEOMETRY_TYPE_ROUND_CATMULL_ROM_CURVE = 58,
  RTC_GEOMETRY_TYPE_FLAT_CATMULL_ROM_CURVE  = 59,
  RTC_GEOMETRY_TYPE_NORMAL_ORIENTED_CATMULL_ROM_CURVE  = 60,

  RTC_GEOMETRY_TYPE_USER     = 120,
  RTC_GEOMETRY_TYPE_INSTANCE = 121
};

/* Creates a geometry (reference: rtcore_geometry.h:135, rtcore.cpp:1139). */
RTC_API RTCGeometry rtcNewGeometry(RTCDevice device, enum RTCGeometryType type);

/* Reference counting (reference: rtcore_geometry.h:138-141). */
RTC_API void rtcRetainGeometry(RTCGeometry geometry);
RTC_API void rtcReleaseGeometry(RTCGeometry geometry);

/* Marks the geometry ready for the next scene commit (reference:
 * rtcore_geometry.h:144, rtcore.cpp:1708). */
RTC_API void rtcCommitGeometry(RTCGeometry geometry);

/* Enable / disable (reference: rtcore_geometry.h:148-151). A disabled
 * geometry contributes no primitives at the next scene commit. */
RTC_API void rtcEnableGeometry(RTCGeometry geometry);
RTC_API void rtcDisableGeometry(RTCGeometry geometry);

/* Motion blur time steps (reference: rtcore_geometry.h:155). Only one time
 * step is supported; any other count records RTC_ERROR_INVALID_OPERATION. */
RTC_API void rtcSetGeometryTimeStepCount(RTCGeometry geometry, unsigned int timeStepCount);

/* Ray mask of the geometry (reference: rtcore_geometry.h:164). Stored; the
 * traversal ignores masks, matching the reference's default build without
 * EMBREE_RAY_MASK (CMakeLists.txt option, default OFF). */
RTC_API void rtcSetGeometryMask(RTCGeometry geometry, unsigned int mask);

/* Per-geometry build quality (reference: rtcore_geometry.h:167). */
RTC_API void rtcSetGeometryBuildQuality(RTCGeometry geometry, enum RTCBuildQuality quality);

/* Binds a buffer object (reference: rtcore_geometry.h:174, rtcore.cpp:1436). */
RTC_API void rtcSetGeometryBuffer(RTCGeometry geometry, enum RTCBufferType type, unsigned int slot, enum RTCFormat format, RTCBuffer buffer, size_t byteOffset, size_t byteStride, size_t itemCount);

/* Binds caller memory without copying (reference: rtcore_geometry.h:177,
 * rtcore.cpp:1455). The memory must stay valid until the scene is committed
 * again or released; it is read at rtcCommitScene. */
RTC_API void rtcSetSharedGeometryBuffer(RTCGeometry geometry, enum RTCBufferType type, unsigned int slot, enum RTCFormat format, const void* ptr, size_t byteOffset, size_t byteStride, size_t itemCount);

/* Allocates and binds a new buffer (reference: rtcore_geometry.h:180,
 * rtcore.cpp:1470-1490; vertex buffers get 16 bytes of padding). */
RTC_API void* rtcSetNewGeometryBuffer(RTCGeometry geometry, enum RTCBufferType type, unsigned int slot, enum RTCFormat format, size_t byteStride, size_t itemCount);

/* Buffer data pointer (reference: rtcore_geometry.h:183). */
RTC_API void* rtcGetGeometryBufferData(RTCGeometry geometry, enum RTCBufferType type, unsigned int slot);

/* Marks a buffer modified (reference: rtcore_geometry.h:186). */
RTC_API void rtcUpdateGeometryBuffer(RTCGeometry geometry, enum RTCBufferType type, unsigned int slot);

/* Filter callbacks (reference: rtcore_geometry.h:190-193). Host callbacks
 * cannot run inside GPU traversal: setting a non-NULL filter records
 * RTC_ERROR_INVALID_OPERATION; NULL is accepted. */
RTC_API void rtcSetGeometryIntersectFilterFunction(RTCGeometry geometry, RTCFilterFunctionN filter);
RTC_API void rtcSetGeometryOccludedFilterFunction(RTCGeometry geometry, RTCFilterFunctionN filter);

/* Instances (reference: rtcore_geometry.h:224-233; kernels/common/scene_instance.cpp,
 * kernels/geometry/instance_intersector.cpp:52-98). An RTC_GEOMETRY_TYPE_INSTANCE
 * geometry places a committed scene into its parent with an affine
 * local-to-world transform; the parent's commit builds a top tree over its
 * instances (and its own triangles) and the traversal moves each ray into
 * the instance's object space. One instance level (RTC_MAX_INSTANCE_LEVEL_COUNT
 * 1): instances inside an instanced scene are not entered. Hits report
 * instID[0] = the instance's geomID in the parent and Ng in object space.
 * An instanced scene committed again after its parent's commit stays, for
 * the parent, as it was at the parent's commit (the parent holds that tree)
 * until the parent is committed again. One time step only. */
RTC_API void rtcSetGeometryInstancedScene(RTCGeometry geometry, RTCScene scene);

/* Transform of time step 0 (reference: rtcore_geometry.h:227, rtcore.cpp:1008-1039):
 * RTC_FORMAT_FLOAT3X4_ROW_MAJOR, RTC_FORMAT_FLOAT3X4_COLUMN_MAJOR or
 * RTC_FORMAT_FLOAT4X4_COLUMN_MAJOR; other formats record RTC_ERROR_INVALID_OPERATION. */
RTC_API void rtcSetGeometryTransform(RTCGeometry geometry, unsigned int timeStep, enum RTCFormat format, const void* xfm);

/* Reads the transform back in the given format (reference: rtcore_geometry.h:233,
 * rtcore.cpp:1041-1069). */
RTC_API void rtcGetGeometryTransform(RTCGeometry geometry, float time, enum RTCFormat format, void* xfm);

/* User data pointer (reference: rtcore_geometry.h:196-199). */
RTC_API void rtcSetGeometryUserData(RTCGeometry geometry, void* ptr);
RTC_API void* rtcGetGeometryUserData(RTCGeometry geometry);

RTC_NAMESPACE_END
