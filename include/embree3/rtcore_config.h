/* Build-time configuration of the MI355X triangle ray-query library.
 *
 * Same macro names and meaning as the reference's generated configuration
 * header (reference: kernels/rtcore_config.h.in:6-57); the values are fixed
 * here because this library has exactly one build: Embree 3.12.1 API level,
 * one instance level, no API namespace, symbols exported with C linkage. */
#pragma once

#define RTC_VERSION_MAJOR 3
#define RTC_VERSION_MINOR 12
#define RTC_VERSION_PATCH 1
#define RTC_VERSION 31201
#define RTC_VERSION_STRING "3.12.1-mi355x"

#define RTC_MAX_INSTANCE_LEVEL_COUNT 1

#define EMBREE_MIN_WIDTH 0
#define RTC_MIN_WIDTH EMBREE_MIN_WIDTH

#define RTC_NAMESPACE_BEGIN
#define RTC_NAMESPACE_END
#define RTC_NAMESPACE_USE

#if defined(__cplusplus)
#  define RTC_API_EXTERN_C extern "C"
#else
#  define RTC_API_EXTERN_C
#endif

#define RTC_API_IMPORT RTC_API_EXTERN_C
#define RTC_API_EXPORT RTC_API_EXTERN_C __attribute__((visibility("default")))

#if defined(RTC_EXPORT_API)
#  define RTC_API RTC_API_EXPORT
#else
#  define RTC_API RTC_API_IMPORT
#endif
