/* Device object of the triangle ray-query C API.
 *
 * A device owns one HIP device context (device 0 unless the config string
 * names another with "gpu=<n>"), its streams and scratch memory. */
#pragma once

#include "rtcore_common.h"

RTC_NAMESPACE_BEGIN

typedef struct RTCDeviceTy* RTCDevice;

/* Creates a device from a config string (reference: rtcore_device.h:14,
 * kernels/common/rtcore.cpp:27, kernels/common/state.cpp:249-420 for the
 * key=value syntax). Recognised keys: verbose; gpu (HIP device ordinal);
 * tri_builder or its alias builder (state.cpp:309-310): default (the build
 * quality chooses: LOW the Morton tree, MEDIUM and HIGH the binned-SAH tree),
 * sah and dynamic (binned SAH), morton or lbvh (Morton radix tree), at every
 * quality; any other name fails rtcCommitScene with
 * RTC_ERROR_INVALID_ARGUMENT, as the reference's factory does
 * (bvh8_factory.cpp:878); bvh_leaf / max_leaf, bvh_width, cost_node,
 * cost_tri, sah_block (leaves priced per block of that many triangles, as
 * the reference's sahBlockSize; tree shape); spatial_budget (extra triangle
 * references spatial splits may add at RTC_BUILD_QUALITY_HIGH, as a fraction
 * of the triangles; default 0.3, 0 turns them off); trace_stack, refill_min, refill_min_occ,
 * refill_min_occ_inst, blocks_per_cu, persistent, dual, ray_sort
 * (traversal launch settings, as rtcgAutotune writes them). The reference's
 * CPU keys (threads, isa, set_affinity, ...) and unknown keys are ignored as
 * the reference ignores unknown keys. The learned split cost is chosen per
 * scene (rtcgSetSceneSplitModel), not by a device key. */
RTC_API RTCDevice rtcNewDevice(const char* config);

/* Reference counting (reference: rtcore_device.h:17-20, rtcore.cpp:37-68). */
RTC_API void rtcRetainDevice(RTCDevice device);
RTC_API void rtcReleaseDevice(RTCDevice device);

/* Device properties (reference: rtcore_device.h:23-52). */
enum RTCDeviceProperty
{
  RTC_DEVICE_PROPERTY_VERSION       = 0,
  RTC_DEVICE_PROPERTY_VERSION_MAJOR = 1,
  RTC_DEVICE_PROPERTY_VERSION_MINOR = 2,
  RTC_DEVICE_PROPERTY_VERSION_PATCH = 3,

  RTC_DEVICE_PROPERTY_NATIVE_RAY4_SUPPORTED  = 32,
  RTC_DEVICE_PROPERTY_NATIVE_RAY8_SUPPORTED  = 33,
  RTC_DEVICE_PROPERTY_NATIVE_RAY16_SUPPORTED = 34,
  RTC_DEVICE_PROPERTY_RAY_STREAM_SUPPORTED   = 35,

  RTC_DEVICE_PROPERTY_BACKFACE_CULLING_CURVES_ENABLED = 63,
  RTC_DEVICE_PROPERTY_RAY_MASK_SUPPORTED          = 64,
  RTC_DEVICE_PROPERTY_BACKFACE_CULLING_ENABLED    = 65,
  RTC_DEVICE_PROPERTY_FILTER_FUNCTION_SUPPORTED   = 66,
  RTC_DEVICE_PROPERTY_IGNORE_INVALID_RAYS_ENABLED = 67,
  RTC_DEVICE_PROPERTY_COMPACT_POLYS_ENABLED       = 68,

  RTC_DEVICE_PROPERTY_TRIANGLE_GEOMETRY_SUPPORTED    = 96,
  RTC_DEVICE_PROPERTY_QUAD_GEOMETRY_SUPPORTED        = 97,
  RTC_DEVICE_PROPERTY_SUBDIVISION_GEOMETRY_SUPPORTED = 98,
  RTC_DEVICE_PROPERTY_CURVE_GEOMETRY_SUPPORTED       = 99,
  RTC_DEVICE_PROPERTY_USER_GEOMETRY_SUPPORTED        = 100,
  RTC_DEVICE_PROPERTY_POINT_GEOMETRY_SUPPORTED       = 101,

  RTC_DEVICE_PROPERTY_TASKING_SYSTEM        = 128,
  RTC_DEVICE_PROPERTY_JOIN_COMMIT_SUPPORTED = 129,
  RTC_DEVICE_PROPERTY_PARALLEL_COMMIT_SUPPORTED = 130
};

/* Property query (reference: rtcore_device.h:55, kernels/common/device.cpp:445-515). */
RTC_API ssize_t rtcGetDeviceProperty(RTCDevice device, enum RTCDeviceProperty prop);

/* Property update (reference: rtcore_device.h:58). No property is writable;
 * every call records RTC_ERROR_INVALID_ARGUMENT as the reference does. */
RTC_API void rtcSetDeviceProperty(RTCDevice device, const enum RTCDeviceProperty prop, ssize_t value);

/* Error codes (reference: rtcore_device.h:61-70). */
enum RTCError
{
  RTC_ERROR_NONE              = 0,
  RTC_ERROR_UNKNOWN           = 1,
  RTC_ERROR_INVALID_ARGUMENT  = 2,
  RTC_ERROR_INVALID_OPERATION = 3,
  RTC_ERROR_OUT_OF_MEMORY     = 4,
  RTC_ERROR_UNSUPPORTED_CPU   = 5,
  RTC_ERROR_CANCELLED         = 6
};

/* Returns and clears the calling thread's first recorded error
 * (reference: rtcore_device.h:73, kernels/common/device.cpp:258-285). */
RTC_API enum RTCError rtcGetDeviceError(RTCDevice device);

typedef void (*RTCErrorFunction)(void* userPtr, enum RTCError code, const char* str);

/* Error callback (reference: rtcore_device.h:79). */
RTC_API void rtcSetDeviceErrorFunction(RTCDevice device, RTCErrorFunction error, void* userPtr);

typedef bool (*RTCMemoryMonitorFunction)(void* ptr, ssize_t bytes, bool post);

/* Memory monitor (reference: rtcore_device.h:85, device.cpp:318-326): called
 * before every device allocation of the scene build with the byte count;
 * returning false aborts the commit with RTC_ERROR_OUT_OF_MEMORY. */
RTC_API void rtcSetDeviceMemoryMonitorFunction(RTCDevice device, RTCMemoryMonitorFunction memoryMonitor, void* userPtr);

RTC_NAMESPACE_END
