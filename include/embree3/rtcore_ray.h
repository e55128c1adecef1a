/* Ray and hit records of the triangle ray-query C API.
 *
 * The byte layout is the ABI the stream calls read and write in place:
 * RTCRay 48 B, RTCHit 32 B, RTCRayHit 80 B, all 16-byte aligned
 * (reference: include/embree3/rtcore_ray.h:11-49). The GPU kernels read
 * org/tnear/dir/time/tfar from offset 0..35 and write tfar plus the hit
 * fields at offsets 32 and 48..79. */
#pragma once

#include "rtcore_common.h"

RTC_NAMESPACE_BEGIN

struct RTC_ALIGN(16) RTCRay
{
  float org_x;
  float org_y;
  float org_z;
  float tnear;

  float dir_x;
  float dir_y;
  float dir_z;
  float time;

  float tfar;
  unsigned int mask;
  unsigned int id;
  unsigned int flags;
};

struct RTC_ALIGN(16) RTCHit
{
  float Ng_x;
  float Ng_y;
  float Ng_z;

  float u;
  float v;

  unsigned int primID;
  unsigned int geomID;
  unsigned int instID[RTC_MAX_INSTANCE_LEVEL_COUNT];
};

struct RTCRayHit
{
  struct RTCRay ray;
  struct RTCHit hit;
};

RTC_NAMESPACE_END
