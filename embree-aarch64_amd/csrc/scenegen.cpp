// Deterministic workload generators for tests and benchmarks (host code).
// Compiled with -ffp-contract=off so the float formulas below round the
// same way on every host.
#define RTC_EXPORT_API
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "embree3/rtcore.h"
#include "embree3/rtcore_mi355x.h"

namespace {

// Counter-based generator: SplitMix64 finaliser keyed by (seed, stream, index),
// so every element is independent of thread count and evaluation order.
inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline float u01(uint64_t seed, uint64_t stream, uint64_t i) {
  const uint64_t h = splitmix64(splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + i);
  return (float)(h >> 40) * (1.0f / 16777216.0f);
}

void init_rayhit(RTCRayHit& r) {
  memset(&r, 0, sizeof(r));
  r.ray.tnear = 0.0f;
  r.ray.tfar = INFINITY;
  r.ray.mask = 0xffffffffu;
  r.ray.flags = 0;
  r.ray.time = 0.0f;
  r.hit.geomID = RTC_INVALID_GEOMETRY_ID;
  r.hit.primID = RTC_INVALID_GEOMETRY_ID;
  r.hit.instID[0] = RTC_INVALID_GEOMETRY_ID;
}

}  // namespace

// UV sphere with the topology of the reference tutorial generator
// (tutorials/common/scenegraph/geometry_creation.cpp:121-175): rings phi =
// 0..numPhi, numTheta = 2*numPhi points per ring, cap fans at both poles.
// Angles are formed as phi*pi/numPhi and theta*2*pi/numTheta in FP32 (the
// tutorial multiplies by rcp(numPhi), which is one ulp low on x86).
RTC_API unsigned int rtcgGenSphere(unsigned int numPhi, float cx, float cy, float cz, float radius,
                                              float* vertices, unsigned int* indices, unsigned int* numVertices) {
  const unsigned numTheta = 2 * numPhi;
  const unsigned nv = numTheta * (numPhi + 1);
  const unsigned nt = numPhi >= 2 ? 2 * numTheta * (numPhi - 1) : 0;
  if (numVertices) *numVertices = nv;
  if (!vertices || !indices) return nt;
  const float pi = (float)M_PI;
  unsigned t = 0;
  for (unsigned phi = 0; phi <= numPhi; phi++) {
    for (unsigned th = 0; th < numTheta; th++) {
      const float pf = phi * pi / numPhi;
      const float tf = th * 2.0f * pi / numTheta;
      float* v = vertices + 3 * (size_t)(phi * numTheta + th);
      // a zero centre is not added, so signed zeros of the unit sphere survive
      v[0] = radius * sinf(pf) * sinf(tf);
      v[1] = radius * cosf(pf);
      v[2] = radius * sinf(pf) * cosf(tf);
      if (cx != 0.0f) v[0] = cx + v[0];
      if (cy != 0.0f) v[1] = cy + v[1];
      if (cz != 0.0f) v[2] = cz + v[2];
    }
    if (phi == 0) continue;
    for (unsigned th = 1; th <= numTheta; th++) {
      const unsigned p00 = (phi - 1) * numTheta + th - 1;
      const unsigned p01 = (phi - 1) * numTheta + th % numTheta;
      const unsigned p10 = phi * numTheta + th - 1;
      const unsigned p11 = phi * numTheta + th % numTheta;
      unsigned* ix = indices + 3 * (size_t)t;
      if (phi == 1) {
        ix[0] = p10; ix[1] = numTheta - 1; ix[2] = p11;
        t++;
      } else if (phi == numPhi) {
        ix[0] = numPhi * numTheta; ix[1] = p00; ix[2] = p01;
        t++;
      } else {
        ix[0] = p10; ix[1] = p00; ix[2] = p11;
        ix[3] = p01; ix[4] = p11; ix[5] = p00;
        t += 2;
      }
    }
  }
  return t;
}

// Pinhole camera: pixel centres (x+0.5, y+0.5), image plane at distance 1
// with half-height tan(fov/2), directions normalised.
RTC_API void rtcgGenPinholeRays(unsigned int width, unsigned int height, float ox, float oy, float oz,
                                           float fovDeg, RTCRayHit* rays) {
  const float tanh_ = tanf(0.5f * fovDeg * (float)M_PI / 180.0f);
  for (unsigned y = 0; y < height; y++)
    for (unsigned x = 0; x < width; x++) {
      RTCRayHit& r = rays[(size_t)y * width + x];
      init_rayhit(r);
      const float px = (2 * (x + 0.5f) / width - 1) * tanh_;
      const float py = (1 - 2 * (y + 0.5f) / height) * tanh_;
      const float n = sqrtf(px * px + py * py + 1);
      r.ray.org_x = ox; r.ray.org_y = oy; r.ray.org_z = oz;
      r.ray.dir_x = px / n; r.ray.dir_y = py / n; r.ray.dir_z = -1 / n;
      r.ray.id = y * width + x;
    }
}

RTC_API void rtcgGenSoup(unsigned int numTris, unsigned long long seed, float h, float* vertices,
                                    unsigned int* indices) {
  for (unsigned i = 0; i < numTris; i++) {
    float c[3];
    for (int k = 0; k < 3; k++) c[k] = u01(seed, 0, 3ull * i + k);
    for (int j = 0; j < 3; j++) {
      for (int k = 0; k < 3; k++) {
        const float o = (2.0f * u01(seed, 1, 9ull * i + 3 * j + k) - 1.0f) * h;
        vertices[9 * (size_t)i + 3 * j + k] = c[k] + o;
      }
      indices[3 * (size_t)i + j] = 3 * i + j;
    }
  }
}

RTC_API void rtcgGenRandomRays(unsigned int numRays, unsigned long long seed, float lo, float hi, float tfar,
                                          RTCRayHit* rays) {
  const float twopi = 2.0f * (float)M_PI;
  for (unsigned i = 0; i < numRays; i++) {
    RTCRayHit& r = rays[i];
    init_rayhit(r);
    r.ray.org_x = lo + (hi - lo) * u01(seed, 2, 3ull * i + 0);
    r.ray.org_y = lo + (hi - lo) * u01(seed, 2, 3ull * i + 1);
    r.ray.org_z = lo + (hi - lo) * u01(seed, 2, 3ull * i + 2);
    const float z = 2.0f * u01(seed, 3, 2ull * i) - 1.0f;
    const float phi = twopi * u01(seed, 3, 2ull * i + 1);
    const float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    r.ray.dir_x = rr * cosf(phi);
    r.ray.dir_y = rr * sinf(phi);
    r.ray.dir_z = z;
    r.ray.tfar = tfar;
    r.ray.id = i;
  }
}
