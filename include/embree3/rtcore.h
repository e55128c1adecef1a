/* Umbrella header of the MI355X triangle ray-query library, a drop-in for
 * the triangle subset of the reference's include/embree3/rtcore.h. */
#pragma once

#include "rtcore_config.h"
#include "rtcore_common.h"
#include "rtcore_device.h"
#include "rtcore_buffer.h"
#include "rtcore_ray.h"
#include "rtcore_geometry.h"
#include "rtcore_scene.h"
#include "rtcore_builder.h"
