/* Buffer objects of the triangle ray-query C API (host memory). */
#pragma once

#include "rtcore_device.h"

RTC_NAMESPACE_BEGIN

/* Buffer types (reference: rtcore_buffer.h:11-33). Triangle geometries use
 * RTC_BUFFER_TYPE_INDEX and RTC_BUFFER_TYPE_VERTEX slot 0. */
enum RTCBufferType
{
  RTC_BUFFER_TYPE_INDEX            = 0,
  RTC_BUFFER_TYPE_VERTEX           = 1,
  RTC_BUFFER_TYPE_VERTEX_ATTRIBUTE = 2,
  RTC_BUFFER_TYPE_NORMAL           = 3,
  RTC_BUFFER_TYPE_TANGENT          = 4,
  RTC_BUFFER_TYPE_NORMAL_DERIVATIVE = 5,

  RTC_BUFFER_TYPE_GRID                 = 8,

  RTC_BUFFER_TYPE_FACE                 = 16,
  RTC_BUFFER_TYPE_LEVEL                = 17,
  RTC_BUFFER_TYPE_EDGE_CREASE_INDEX    = 18,
  RTC_BUFFER_TYPE_EDGE_CREASE_WEIGHT   = 19,
  RTC_BUFFER_TYPE_VERTEX_CREASE_INDEX  = 20,
  RTC_BUFFER_TYPE_VERTEX_CREASE_WEIGHT = 21,
  RTC_BUFFER_TYPE_HOLE                 = 22,

  RTC_BUFFER_TYPE_FLAGS = 32
};

typedef struct RTCBufferTy* RTCBuffer;

/* Host buffer owned by the library (reference: rtcore_buffer.h:37). */
RTC_API RTCBuffer rtcNewBuffer(RTCDevice device, size_t byteSize);

/* Host buffer wrapping caller memory, not copied (reference: rtcore_buffer.h:40). */
RTC_API RTCBuffer rtcNewSharedBuffer(RTCDevice device, void* ptr, size_t byteSize);

/* Data pointer (reference: rtcore_buffer.h:43). */
RTC_API void* rtcGetBufferData(RTCBuffer buffer);

/* Reference counting (reference: rtcore_buffer.h:46-49). */
RTC_API void rtcRetainBuffer(RTCBuffer buffer);
RTC_API void rtcReleaseBuffer(RTCBuffer buffer);

RTC_NAMESPACE_END
