// Deterministic workload generators for tests and benchmarks (host code).
// Compiled with -ffp-contract=off so the float formulas below round the
// same way on every host.
#define RTC_EXPORT_API
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

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

namespace {

// Gradient noise after Perlin's "Improved Noise" (SIGGRAPH 2002): a
// seeded permutation of 0..255 hashes each lattice corner to one of eight
// gradient directions (+-1,+-1), (+-1,0), (0,+-1); the corner's value is the
// dot product of its gradient with the offset to the point, blended with the
// quintic fade 6t^5 - 15t^4 + 10t^3. Zero at every lattice point; |n| <= 1.
struct Perlin {
  uint8_t perm[512];
  Perlin(uint64_t seed, uint32_t oct) {
    for (int i = 0; i < 256; i++) perm[i] = (uint8_t)i;
    for (int i = 255; i > 0; i--) {  // Fisher-Yates with the counter-based generator
      const int j = (int)(u01(seed, 64 + oct, (uint64_t)i) * (float)(i + 1));
      const uint8_t t = perm[i];
      perm[i] = perm[j < i ? j : i];
      perm[j < i ? j : i] = t;
    }
    for (int i = 0; i < 256; i++) perm[256 + i] = perm[i];
  }
  static float fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }
  static float grad(uint32_t h, float x, float y) {
    switch (h & 7u) {
      case 0: return x + y;
      case 1: return -x + y;
      case 2: return x - y;
      case 3: return -x - y;
      case 4: return x;
      case 5: return -x;
      case 6: return y;
      default: return -y;
    }
  }
  float operator()(float x, float y) const {
    const float fx = floorf(x), fy = floorf(y);
    const uint32_t X = (uint32_t)(int32_t)fx & 255u, Y = (uint32_t)(int32_t)fy & 255u;
    x -= fx;
    y -= fy;
    const float u = fade(x), v = fade(y);
    const uint32_t a = perm[X] + Y, b = perm[X + 1] + Y;
    const float n00 = grad(perm[a], x, y), n10 = grad(perm[b], x - 1.0f, y);
    const float n01 = grad(perm[a + 1], x, y - 1.0f), n11 = grad(perm[b + 1], x - 1.0f, y - 1.0f);
    const float nx0 = n00 + u * (n10 - n00), nx1 = n01 + u * (n11 - n01);
    return 0.5f * (nx0 + v * (nx1 - nx0));
  }
};

void normalize3(float v[3]) {
  const float l = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (l > 0.0f) {
    v[0] /= l;
    v[1] /= l;
    v[2] /= l;
  }
}

void cross3(const float a[3], const float b[3], float r[3]) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

}  // namespace

// n x n vertex grid over [0,1]^2 (x, y), height z = amplitude * sum over
// `octaves` of Perlin gradient noise at frequency 4*2^o with weight 2^-o
// (one permutation per octave); two triangles per cell, 2*(n-1)^2 in all.
RTC_API unsigned int rtcgGenHeightfield(unsigned int n, unsigned long long seed, float amplitude, unsigned int octaves,
                                        float* vertices, unsigned int* indices) {
  if (n < 2) return 0;
  const unsigned nt = 2 * (n - 1) * (n - 1);
  if (!vertices || !indices) return nt;
  const float inv = 1.0f / (float)(n - 1);
  std::vector<Perlin> noise;
  for (unsigned o = 0; o < octaves; o++) noise.emplace_back(seed, o);
  for (unsigned j = 0; j < n; j++)
    for (unsigned i = 0; i < n; i++) {
      const float x = (float)i * inv, y = (float)j * inv;
      float h = 0.0f, w = 1.0f, f = 4.0f;
      for (unsigned o = 0; o < octaves; o++) {
        h += w * noise[o](x * f, y * f);
        w *= 0.5f;
        f *= 2.0f;
      }
      float* v = vertices + 3 * ((size_t)j * n + i);
      v[0] = x;
      v[1] = y;
      v[2] = amplitude * h;
    }
  size_t t = 0;
  for (unsigned j = 0; j + 1 < n; j++)
    for (unsigned i = 0; i + 1 < n; i++) {
      const unsigned a = j * n + i, b = a + 1, c = a + n, d = c + 1;
      unsigned* ix = indices + 3 * t;
      ix[0] = a; ix[1] = b; ix[2] = d;
      ix[3] = a; ix[4] = d; ix[5] = c;
      t += 2;
    }
  return nt;
}

// Pinhole camera at org looking at target (up vector up, vertical field of
// view fovDeg), one ray per pixel centre, row-major from the top row.
RTC_API void rtcgGenCameraRays(unsigned int width, unsigned int height, const float* org, const float* target,
                               const float* up, float fovDeg, RTCRayHit* rays) {
  float fw[3] = {target[0] - org[0], target[1] - org[1], target[2] - org[2]};
  normalize3(fw);
  float rt[3], u[3];
  cross3(fw, up, rt);
  normalize3(rt);
  cross3(rt, fw, u);
  const float th = tanf(0.5f * fovDeg * (float)M_PI / 180.0f);
  const float aspect = (float)width / (float)height;
  for (unsigned y = 0; y < height; y++)
    for (unsigned x = 0; x < width; x++) {
      RTCRayHit& r = rays[(size_t)y * width + x];
      init_rayhit(r);
      const float px = (2.0f * ((float)x + 0.5f) / (float)width - 1.0f) * th * aspect;
      const float py = (1.0f - 2.0f * ((float)y + 0.5f) / (float)height) * th;
      float d[3] = {fw[0] + px * rt[0] + py * u[0], fw[1] + px * rt[1] + py * u[1], fw[2] + px * rt[2] + py * u[2]};
      normalize3(d);
      r.ray.org_x = org[0]; r.ray.org_y = org[1]; r.ray.org_z = org[2];
      r.ray.dir_x = d[0]; r.ray.dir_y = d[1]; r.ray.dir_z = d[2];
      r.ray.id = y * width + x;
    }
}

// Next-bounce rays of a path tracer: for every hit record in `hits`, a ray
// from the hit point (tnear = epsilon) in a cosine-weighted direction about
// the geometric normal turned towards the incoming ray
// (tutorials/common/math/sampling.h:52-78 cosineSampleHemisphere, frame(N)
// as common/math/linearspace3.h:113-120). Misses spawn nothing. Returns the
// number of rays written to out (at most M).
RTC_API unsigned int rtcgGenBounceRays(const RTCRayHit* hits, unsigned int M, unsigned long long seed, float epsilon,
                                       RTCRayHit* out) {
  const float twopi = 2.0f * (float)M_PI;
  unsigned k = 0;
  for (unsigned i = 0; i < M; i++) {
    const RTCRayHit& h = hits[i];
    if (h.hit.geomID == RTC_INVALID_GEOMETRY_ID) continue;
    float N[3] = {h.hit.Ng_x, h.hit.Ng_y, h.hit.Ng_z};
    normalize3(N);
    if (N[0] * h.ray.dir_x + N[1] * h.ray.dir_y + N[2] * h.ray.dir_z > 0.0f) {
      N[0] = -N[0]; N[1] = -N[1]; N[2] = -N[2];
    }
    const float t = h.ray.tfar;
    const float p[3] = {h.ray.org_x + t * h.ray.dir_x, h.ray.org_y + t * h.ray.dir_y, h.ray.org_z + t * h.ray.dir_z};
    // frame(N): dx from the larger of (0, Nz, -Ny) and (-Nz, 0, Nx)
    float dx0[3] = {0.0f, N[2], -N[1]}, dx1[3] = {-N[2], 0.0f, N[0]};
    float* dx = (dx0[1] * dx0[1] + dx0[2] * dx0[2] > dx1[0] * dx1[0] + dx1[2] * dx1[2]) ? dx0 : dx1;
    normalize3(dx);
    float dy[3];
    cross3(N, dx, dy);
    normalize3(dy);
    const float u = u01(seed, 20, 2ull * i), v = u01(seed, 20, 2ull * i + 1);
    const float phi = twopi * u, ct = sqrtf(v), st = sqrtf(1.0f - v);
    const float l[3] = {cosf(phi) * st, sinf(phi) * st, ct};
    RTCRayHit& r = out[k++];
    init_rayhit(r);
    r.ray.org_x = p[0]; r.ray.org_y = p[1]; r.ray.org_z = p[2];
    r.ray.dir_x = dx[0] * l[0] + dy[0] * l[1] + N[0] * l[2];
    r.ray.dir_y = dx[1] * l[0] + dy[1] * l[1] + N[1] * l[2];
    r.ray.dir_z = dx[2] * l[0] + dy[2] * l[1] + N[2] * l[2];
    r.ray.tnear = epsilon;
    r.ray.id = i;
  }
  return k;
}

namespace {

float lattice3(uint64_t seed, uint32_t oct, int32_t x, int32_t y, int32_t z) {
  const uint64_t key = ((uint64_t)(uint32_t)x * 73856093ull) ^ ((uint64_t)(uint32_t)y * 19349663ull << 21) ^
                       ((uint64_t)(uint32_t)z * 83492791ull << 42);
  return 2.0f * u01(seed, 32 + oct, key) - 1.0f;
}

float value_noise3(uint64_t seed, uint32_t oct, float x, float y, float z) {
  const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
  const int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
  float tx = x - fx, ty = y - fy, tz = z - fz;
  tx = tx * tx * (3.0f - 2.0f * tx);
  ty = ty * ty * (3.0f - 2.0f * ty);
  tz = tz * tz * (3.0f - 2.0f * tz);
  float c[2][2];
  for (int dz = 0; dz < 2; dz++)
    for (int dy = 0; dy < 2; dy++) {
      const float a = lattice3(seed, oct, ix, iy + dy, iz + dz), b = lattice3(seed, oct, ix + 1, iy + dy, iz + dz);
      c[dz][dy] = a + (b - a) * tx;
    }
  const float e = c[0][0] + (c[0][1] - c[0][0]) * ty, f = c[1][0] + (c[1][1] - c[1][0]) * ty;
  return e + (f - e) * tz;
}

}  // namespace

// Unit UV sphere (reference tutorial topology, numPhi rings) displaced
// radially by amplitude * 3-octave value noise of the surface point
// (frequency 2*2^o, weight 2^-o). Returns the triangle count; NULL buffers
// query the sizes as rtcgGenSphere does.
RTC_API unsigned int rtcgGenDisplacedSphere(unsigned int numPhi, unsigned long long seed, float amplitude,
                                            float* vertices, unsigned int* indices, unsigned int* numVertices) {
  const unsigned nt = rtcgGenSphere(numPhi, 0.0f, 0.0f, 0.0f, 1.0f, vertices, indices, numVertices);
  if (!vertices || !indices) return nt;
  const unsigned nv = 2 * numPhi * (numPhi + 1);
  for (unsigned i = 0; i < nv; i++) {
    float* v = vertices + 3 * (size_t)i;
    float h = 0.0f, w = 1.0f, f = 2.0f;
    for (unsigned o = 0; o < 3; o++) {
      h += w * value_noise3(seed, o, v[0] * f, v[1] * f, v[2] * f);
      w *= 0.5f;
      f *= 2.0f;
    }
    const float s = 1.0f + amplitude * h;
    v[0] *= s;
    v[1] *= s;
    v[2] *= s;
  }
  return nt;
}

// Rays from uniform points on the sphere of the given radius towards
// uniform points inside the unit ball (closest-hit records).
RTC_API void rtcgGenSphereRays(unsigned int numRays, unsigned long long seed, float radius, RTCRayHit* rays) {
  const float twopi = 2.0f * (float)M_PI;
  for (unsigned i = 0; i < numRays; i++) {
    RTCRayHit& r = rays[i];
    init_rayhit(r);
    const float z = 2.0f * u01(seed, 40, 2ull * i) - 1.0f, phi = twopi * u01(seed, 40, 2ull * i + 1);
    const float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    const float o[3] = {radius * rr * cosf(phi), radius * rr * sinf(phi), radius * z};
    // target: direction uniform on the sphere, radius cbrt(u)
    const float z2 = 2.0f * u01(seed, 41, 3ull * i) - 1.0f, phi2 = twopi * u01(seed, 41, 3ull * i + 1);
    const float r2 = cbrtf(u01(seed, 41, 3ull * i + 2)), rr2 = sqrtf(fmaxf(0.0f, 1.0f - z2 * z2));
    float d[3] = {r2 * rr2 * cosf(phi2) - o[0], r2 * rr2 * sinf(phi2) - o[1], r2 * z2 - o[2]};
    normalize3(d);
    r.ray.org_x = o[0]; r.ray.org_y = o[1]; r.ray.org_z = o[2];
    r.ray.dir_x = d[0]; r.ray.dir_y = d[1]; r.ray.dir_z = d[2];
    r.ray.id = i;
  }
}
