// Spatial pre-splitting of triangles for the HIGH-quality build (see
// spatial.h). One thread per triangle throughout: its priority, its number
// of pieces within the budget, and the pieces themselves, each the box of
// the triangle clipped (Sutherland-Hodgman) to a region of its box cut by
// axis-aligned planes.
#include "spatial.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace rtg {
namespace {

constexpr int TB = 256;

struct Grid {
  float lo[3];
  float scale[3];  // 2^32 / extent (0 on a flat axis)
};

__device__ inline void load_tri(const GeomDesc* __restrict__ geoms, float4 l, float4 h, float v[9]) {
  const GeomDesc g = geoms[__float_as_uint(l.w)];
  const uint32_t prim = __float_as_uint(h.w);
  const uint32_t* ix = (const uint32_t*)(g.indices + (uint64_t)prim * g.istride);
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const float* p = (const float*)(g.vertices + (uint64_t)ix[j] * g.vstride);
    v[3 * j] = p[0];
    v[3 * j + 1] = p[1];
    v[3 * j + 2] = p[2];
  }
}

// Position of x on the scene's dyadic grid as a 32-bit fraction.
__device__ inline uint32_t grid_q(float x, const Grid& gr, int k) {
  const float u = (x - gr.lo[k]) * gr.scale[k];
  return u <= 0.0f ? 0u : (u >= 4294967040.0f ? 0xffffff00u : (uint32_t)u);
}

// Depth of the coarsest grid plane strictly between lo and hi on axis k
// (32 when none).
__device__ inline int plane_level(float lo, float hi, const Grid& gr, int k) {
  if (gr.scale[k] == 0.0f) return 32;
  const uint32_t a = grid_q(lo, gr, k), b = grid_q(hi, gr, k);
  return a == b ? 32 : __clz(a ^ b);
}

// That plane's coordinate, or the midpoint where rounding puts it outside.
__device__ inline float plane_at(float lo, float hi, const Grid& gr, int k) {
  const float mid = 0.5f * (lo + hi);
  if (gr.scale[k] == 0.0f) return mid;
  const uint32_t a = grid_q(lo, gr, k), b = grid_q(hi, gr, k);
  if (a == b) return mid;
  const int d = __clz(a ^ b);
  const uint32_t q = (b >> (31 - d)) << (31 - d);
  const float x = gr.lo[k] + (float)q / gr.scale[k];
  return (x > lo && x < hi) ? x : mid;
}

// Box of the part of triangle v inside the closed box [rl, rh]: the bounds
// of the clipped polygon's vertices, which are the triangle's vertices in
// the box, the points where its edges cross the box's faces inside the box,
// and the points where the box's edges pierce the triangle. Points are
// taken within tol of the box; the box is then held to [rl, rh] and to the
// triangle's box (tl, th) and widened by eps within the latter. Returns
// false (and the box region) when no point qualifies.
__device__ inline bool piece_box(const float v[9], const float rl[3], const float rh[3], const float tl[3],
                                 const float th[3], float tol, float eps, float bl[3], float bh[3]) {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  auto add = [&](float x, float y, float z) {
    if (x >= rl[0] - tol && x <= rh[0] + tol && y >= rl[1] - tol && y <= rh[1] + tol && z >= rl[2] - tol &&
        z <= rh[2] + tol) {
      lo[0] = fminf(lo[0], x); lo[1] = fminf(lo[1], y); lo[2] = fminf(lo[2], z);
      hi[0] = fmaxf(hi[0], x); hi[1] = fmaxf(hi[1], y); hi[2] = fmaxf(hi[2], z);
    }
  };
#pragma unroll
  for (int i = 0; i < 3; i++) add(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
#pragma unroll
  for (int e = 0; e < 3; e++) {
    const float* a = v + 3 * e;
    const float* b = v + 3 * ((e + 1) % 3);
#pragma unroll
    for (int k = 0; k < 3; k++) {
#pragma unroll
      for (int f = 0; f < 2; f++) {
        const float c = f ? rh[k] : rl[k];
        const float da = a[k] - c, db = b[k] - c;
        if ((da < 0.0f) != (db < 0.0f) && da != db) {
          const float t = da / (da - db);
          float p[3];
#pragma unroll
          for (int m = 0; m < 3; m++) p[m] = m == k ? c : a[m] + t * (b[m] - a[m]);
          add(p[0], p[1], p[2]);
        }
      }
    }
  }
  // box edges through the triangle
  const float e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
  const float e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
  const float nr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  const float dd = nr[0] * v[0] + nr[1] * v[1] + nr[2] * v[2];
  int dom = 0;
#pragma unroll
  for (int k = 1; k < 3; k++)
    if (fabsf(nr[k]) > fabsf(nr[dom])) dom = k;
  const int du = (dom + 1) % 3, dw = (dom + 2) % 3;
  const float sgn = nr[dom] >= 0.0f ? 1.0f : -1.0f;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if (nr[k] == 0.0f) continue;
    const int m = (k + 1) % 3, n2 = (k + 2) % 3;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      float p[3];
      p[m] = (c & 1) ? rh[m] : rl[m];
      p[n2] = (c & 2) ? rh[n2] : rl[n2];
      p[k] = (dd - nr[m] * p[m] - nr[n2] * p[n2]) / nr[k];
      if (!(p[k] >= rl[k] - tol && p[k] <= rh[k] + tol)) continue;
      // inside the triangle: the three edge functions in the plane that
      // drops the normal's dominant axis share the normal's sign
      bool in = true;
#pragma unroll
      for (int e = 0; e < 3; e++) {
        const float* a = v + 3 * e;
        const float* b = v + 3 * ((e + 1) % 3);
        const float f = ((b[du] - a[du]) * (p[dw] - a[dw]) - (b[dw] - a[dw]) * (p[du] - a[du])) * sgn;
        in = in && f >= -tol * (fabsf(b[du] - a[du]) + fabsf(b[dw] - a[dw]));
      }
      if (in) add(p[0], p[1], p[2]);
    }
  }
  const bool found = lo[0] <= hi[0];
#pragma unroll
  for (int m = 0; m < 3; m++) {
    float l = found ? fmaxf(lo[m], rl[m]) : rl[m];
    float h = found ? fminf(hi[m], rh[m]) : rh[m];
    l = fmaxf(fmaxf(l, tl[m]) - eps, tl[m]);
    h = fminf(fminf(h, th[m]) + eps, th[m]);
    if (l > h) {
      l = tl[m];
      h = th[m];
    }
    bl[m] = l;
    bh[m] = h;
  }
  return found;
}

// Priority of cutting a triangle: (2^-level g)^(1/3), g the half area the
// first cut saves (the box's less its two pieces' boxes, the cut across the
// box's largest axis at the coarsest grid plane inside it) when that is at
// least MIN_GAIN of the box's and the plane is at most max_level deep, else
// 0; level the depth of that plane.
constexpr float MIN_GAIN = 0.3f;

__global__ __launch_bounds__(TB) void k_split_priority(const uint32_t* __restrict__ vals,
                                                       const float4* __restrict__ lo4, const float4* __restrict__ hi4,
                                                       uint32_t n, const GeomDesc* __restrict__ geoms, Grid gr,
                                                       int max_level, float tol_rel, float* __restrict__ prio) {
  const uint32_t j = blockIdx.x * TB + threadIdx.x;
  if (j >= n) return;
  const uint32_t o = vals[j];
  const float4 l = lo4[o], h = hi4[o];
  const float tl[3] = {l.x, l.y, l.z}, th[3] = {h.x, h.y, h.z};
  int ax = 0;
  for (int m = 1; m < 3; m++)
    if (th[m] - tl[m] > th[ax] - tl[ax]) ax = m;
  const int level = plane_level(tl[ax], th[ax], gr, ax);
  float p = 0.0f;
  if (level <= max_level) {
    float v[9];
    load_tri(geoms, l, h, v);
    const float x = plane_at(tl[ax], th[ax], gr, ax);
    float rl[3] = {tl[0], tl[1], tl[2]}, rh[3] = {th[0], th[1], th[2]};
    float al[3], ah[3], bl[3], bh[3];
    rh[ax] = x;
    const float tol = tol_rel * fmaxf(fmaxf(fabsf(tl[0]), fabsf(th[0])),
                                      fmaxf(fmaxf(fabsf(tl[1]), fabsf(th[1])), fmaxf(fabsf(tl[2]), fabsf(th[2]))));
    piece_box(v, rl, rh, tl, th, tol, 0.0f, al, ah);
    rh[ax] = th[ax];
    rl[ax] = x;
    piece_box(v, rl, rh, tl, th, tol, 0.0f, bl, bh);
    const float abox = half_area(th[0] - tl[0], th[1] - tl[1], th[2] - tl[2]);
    const float gain = abox - half_area(ah[0] - al[0], ah[1] - al[1], ah[2] - al[2]) -
                       half_area(bh[0] - bl[0], bh[1] - bl[1], bh[2] - bl[2]);
    if (gain >= MIN_GAIN * abox && gain > 0.0f) p = cbrtf(gain * exp2f(-(float)level));
  }
  prio[j] = p;
}

// Pieces per triangle: 1 + min(SPLIT_MAX_PIECES - 1, floor(priority * scale)).
__global__ __launch_bounds__(TB) void k_split_count(const float* __restrict__ prio, uint32_t n, float scale,
                                                    uint32_t* __restrict__ count) {
  const uint32_t j = blockIdx.x * TB + threadIdx.x;
  if (j >= n) return;
  const float s = floorf(prio[j] * scale);
  count[j] = 1u + (s >= (float)(SPLIT_MAX_PIECES - 1) ? SPLIT_MAX_PIECES - 1 : (s > 0.0f ? (uint32_t)s : 0u));
}

// Cuts triangle j into count[j] pieces: a piece with more than one piece
// still to make is cut across the largest axis of its box at the coarsest
// grid plane inside it, and its pieces go to the two halves in proportion
// to their extents on that axis (each half at least one).
__global__ __launch_bounds__(TB) void k_split_emit(const uint32_t* __restrict__ vals, const float4* __restrict__ lo4,
                                                   const float4* __restrict__ hi4, uint32_t n,
                                                   const GeomDesc* __restrict__ geoms, Grid gr, float eps_rel,
                                                   const uint32_t* __restrict__ count,
                                                   const uint32_t* __restrict__ offset, float4* __restrict__ olo,
                                                   float4* __restrict__ ohi, uint32_t* __restrict__ ovals) {
  const uint32_t j = blockIdx.x * TB + threadIdx.x;
  if (j >= n) return;
  const uint32_t o = vals[j];
  const float4 l4 = lo4[o], h4 = hi4[o];
  const uint32_t c = count[j];
  uint32_t out = offset[j];
  if (c <= 1u) {
    olo[out] = l4;
    ohi[out] = h4;
    ovals[out] = out;
    return;
  }
  float v[9];
  load_tri(geoms, l4, h4, v);
  const float tl[3] = {l4.x, l4.y, l4.z}, th[3] = {h4.x, h4.y, h4.z};
  float mx = 0.0f;
  for (int m = 0; m < 3; m++) mx = fmaxf(mx, fmaxf(fabsf(tl[m]), fabsf(th[m])));
  const float eps = mx * eps_rel;
  // pending pieces: region and number of pieces to make of it
  float sl[SPLIT_MAX_PIECES][3], sh[SPLIT_MAX_PIECES][3];
  uint32_t sc[SPLIT_MAX_PIECES];
  int sp = 0;
  for (int m = 0; m < 3; m++) {
    sl[0][m] = tl[m];
    sh[0][m] = th[m];
  }
  sc[0] = c;
  sp = 1;
  while (sp > 0) {
    sp--;
    float rl[3], rh[3];
    for (int m = 0; m < 3; m++) {
      rl[m] = sl[sp][m];
      rh[m] = sh[sp][m];
    }
    const uint32_t k = sc[sp];
    float bl[3], bh[3];
    piece_box(v, rl, rh, tl, th, eps, eps, bl, bh);
    if (k <= 1u || sp + 2 > (int)SPLIT_MAX_PIECES) {
      // one piece (or, never expected, no room to split further: the
      // remaining pieces repeat this box)
      for (uint32_t r = 0; r < k; r++) {
        olo[out] = make_float4(bl[0], bl[1], bl[2], l4.w);
        ohi[out] = make_float4(bh[0], bh[1], bh[2], h4.w);
        ovals[out] = out;
        out++;
      }
      continue;
    }
    int ax = 0;
    for (int m = 1; m < 3; m++)
      if (bh[m] - bl[m] > bh[ax] - bl[ax]) ax = m;
    const float x = plane_at(bl[ax], bh[ax], gr, ax);
    const float ext = bh[ax] - bl[ax];
    uint32_t kl = ext > 0.0f ? (uint32_t)rintf((float)k * (x - bl[ax]) / ext) : k / 2;
    kl = kl < 1u ? 1u : (kl > k - 1u ? k - 1u : kl);
    // right half first on the stack, so the left one is emitted first
    for (int m = 0; m < 3; m++) {
      sl[sp][m] = bl[m];
      sh[sp][m] = bh[m];
      sl[sp + 1][m] = bl[m];
      sh[sp + 1][m] = bh[m];
    }
    sl[sp][ax] = x;
    sc[sp] = k - kl;
    sh[sp + 1][ax] = x;
    sc[sp + 1] = kl;
    sp += 2;
  }
}

__global__ void k_gather_strided(const float* __restrict__ src, uint32_t stride, uint32_t m, float* __restrict__ dst) {
  const uint32_t i = blockIdx.x * TB + threadIdx.x;
  if (i < m) dst[i] = src[(size_t)i * stride];
}

}  // namespace

uint32_t spatial_presplit(uint32_t n, float budget, const GeomDesc* geoms, const float glo[3], const float ghi[3],
                          DevBuf& vals, DevBuf& lo4, DevBuf& hi4, SpatialScratch& S, hipStream_t s) {
  if (n == 0 || !(budget > 0.0f)) return n;
  Grid gr;
  for (int k = 0; k < 3; k++) {
    gr.lo[k] = glo[k];
    const float e = ghi[k] - glo[k];
    gr.scale[k] = e > 0.0f ? 4294967296.0f / e : 0.0f;
  }
  S.prio.reserve((size_t)n * 4);
  S.count.reserve((size_t)n * 4);
  S.offset.reserve(((size_t)n + 1) * 4);
  // cuts on planes finer than the leaves' cells (about n / 4 leaves over
  // the scene's volume) separate nothing the tree would not
  const int max_level = (int)std::floor(std::log2(std::max(2.0, (double)n / 4.0)) / 3.0) - 1;
  if (max_level < 0) return n;
  hipLaunchKernelGGL(k_split_priority, dim3(div_up(n, TB)), dim3(TB), 0, s, vals.as<uint32_t>(), lo4.as<float4>(),
                     hi4.as<float4>(), n, geoms, gr, max_level, 1.0f / 1048576.0f, S.prio.as<float>());
  // the scale that spends the budget, found on a strided sample of the
  // priorities: the largest with (n / m) sum min(15, floor(p scale)) <= extra
  const uint32_t m = n < 65536u ? n : 65536u;
  const uint32_t stride = n / m;
  S.sample.reserve((size_t)m * 4);
  S.sample_host.reserve((size_t)m * 4);
  hipLaunchKernelGGL(k_gather_strided, dim3(div_up(m, TB)), dim3(TB), 0, s, S.prio.as<float>(), stride, m,
                     S.sample.as<float>());
  RTG_HIP(hipMemcpyAsync(S.sample_host.ptr, S.sample.ptr, (size_t)m * 4, hipMemcpyDeviceToHost, s));
  RTG_HIP(hipStreamSynchronize(s));
  const float* smp_p = S.sample_host.as<float>();
  const std::vector<float> smp(smp_p, smp_p + m);
  const double extra = std::floor((double)budget * (double)n);
  float pmin = INFINITY;
  for (float p : smp)
    if (p > 0.0f) pmin = std::min(pmin, p);
  if (!(pmin < INFINITY)) return n;  // no triangle gains from a cut
  auto spent = [&](double sc) {
    double t = 0.0;
    for (float p : smp) t += std::min<double>(SPLIT_MAX_PIECES - 1, std::floor((double)p * sc));
    return t * ((double)n / (double)m);
  };
  double lo = 0.0, hi = (double)SPLIT_MAX_PIECES / pmin;
  if (spent(hi) <= extra) lo = hi;
  for (int it = 0; it < 48 && hi > lo * (1.0 + 1e-6); it++) {
    const double mid = lo > 0.0 ? std::sqrt(lo * hi) : hi * 1e-3;
    if (spent(mid) <= extra) lo = mid; else hi = mid;
  }
  if (!(lo > 0.0)) return n;
  hipLaunchKernelGGL(k_split_count, dim3(div_up(n, TB)), dim3(TB), 0, s, S.prio.as<float>(), n, (float)lo,
                     S.count.as<uint32_t>());
  S.scanner.exclusive(S.count.as<uint32_t>(), S.offset.as<uint32_t>(), n, s);
  uint32_t last[2] = {0, 0};
  RTG_HIP(hipMemcpyAsync(&last[0], S.offset.as<uint32_t>() + (n - 1), 4, hipMemcpyDeviceToHost, s));
  RTG_HIP(hipMemcpyAsync(&last[1], S.count.as<uint32_t>() + (n - 1), 4, hipMemcpyDeviceToHost, s));
  RTG_HIP(hipStreamSynchronize(s));
  const uint64_t total = (uint64_t)last[0] + last[1];
  if (total <= n) return n;  // nothing to split
  if (total >= REF_LEAF) throw hip_error(hipErrorInvalidValue, "too many spatial-split references (2^31 at most)");
  const uint32_t nr = (uint32_t)total;
  S.lo4.reserve((size_t)nr * 16);
  S.hi4.reserve((size_t)nr * 16);
  S.vals.reserve((size_t)nr * 4);
  hipLaunchKernelGGL(k_split_emit, dim3(div_up(n, TB)), dim3(TB), 0, s, vals.as<uint32_t>(), lo4.as<float4>(),
                     hi4.as<float4>(), n, geoms, gr, 1.0f / 2097152.0f, S.count.as<uint32_t>(),
                     S.offset.as<uint32_t>(), S.lo4.as<float4>(), S.hi4.as<float4>(), S.vals.as<uint32_t>());
  RTG_HIP(hipGetLastError());
  std::swap(vals.ptr, S.vals.ptr);
  std::swap(vals.bytes, S.vals.bytes);
  std::swap(lo4.ptr, S.lo4.ptr);
  std::swap(lo4.bytes, S.lo4.bytes);
  std::swap(hi4.ptr, S.hi4.ptr);
  std::swap(hi4.bytes, S.hi4.bytes);
  return nr;
}

}  // namespace rtg
