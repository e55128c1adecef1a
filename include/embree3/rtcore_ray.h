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

/* Packets of K rays in SoA form (reference: rtcore_ray.h:50-160): every
 * field is an array of K lanes, in the order of RTCRay then RTCHit, so field
 * f of lane i is element f*K + i. The same order with a runtime K is the
 * opaque RTCRayHitN of rtcIntersectNM. */
#define RTCG_PACKET_TYPES(K, ALIGN)                                  \
  struct RTC_ALIGN(ALIGN) RTCRay##K                                    \
  {                                                                    \
    float org_x[K]; float org_y[K]; float org_z[K]; float tnear[K];    \
    float dir_x[K]; float dir_y[K]; float dir_z[K]; float time[K];     \
    float tfar[K]; unsigned int mask[K]; unsigned int id[K];           \
    unsigned int flags[K];                                             \
  };                                                                   \
  struct RTC_ALIGN(ALIGN) RTCHit##K                                    \
  {                                                                    \
    float Ng_x[K]; float Ng_y[K]; float Ng_z[K]; float u[K];           \
    float v[K]; unsigned int primID[K]; unsigned int geomID[K];        \
    unsigned int instID[RTC_MAX_INSTANCE_LEVEL_COUNT][K];              \
  };                                                                   \
  struct RTCRayHit##K                                                  \
  {                                                                    \
    struct RTCRay##K ray;                                              \
    struct RTCHit##K hit;                                              \
  };

RTCG_PACKET_TYPES(4, 16)
RTCG_PACKET_TYPES(8, 32)
RTCG_PACKET_TYPES(16, 64)
#undef RTCG_PACKET_TYPES

/* N rays in pointer-SoA form: one array per field (reference: rtcore_ray.h:163-205). */
struct RTCRayNp
{
  float* org_x;
  float* org_y;
  float* org_z;
  float* tnear;
  float* dir_x;
  float* dir_y;
  float* dir_z;
  float* time;
  float* tfar;
  unsigned int* mask;
  unsigned int* id;
  unsigned int* flags;
};

struct RTCHitNp
{
  float* Ng_x;
  float* Ng_y;
  float* Ng_z;
  float* u;
  float* v;
  unsigned int* primID;
  unsigned int* geomID;
  unsigned int* instID[RTC_MAX_INSTANCE_LEVEL_COUNT];
};

struct RTCRayHitNp
{
  struct RTCRayNp ray;
  struct RTCHitNp hit;
};

struct RTCHitN;
struct RTCRayHitN;

RTC_NAMESPACE_END
