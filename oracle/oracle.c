/* oracle.c -- CPU restatement of the reference triangle ray-query path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h for the reference anchors and the
 * documented rule differences). Built with -ffp-contract=off so every FMA
 * below is the explicit fmaf() the reference issues, and nothing else is
 * fused. */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* common/math/constants.h: FLT_LARGE bounds valid vertex coordinates
 * (kernels/common/scene_triangle_mesh.h:127-147 via isvalid). */
#define O_FLT_LARGE 1.844E18f
/* common/math/constants.h:31 min_rcp_input, used by zero_fix. */
#define O_MIN_RCP_INPUT 1E-18f
/* Culling margin on ray-interval bounds (documented in oracle.h). */
#define O_MARGIN (1.0f / 4096.0f)
/* Absolute widening of node boxes relative to the largest coordinate. */
#define O_BOX_EPS_REL (1.0f / 1048576.0f)

/* ---------------------------------------------------------------- math -- */

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* prod_diff(a,b,c,d) = msub(a,b,c*d) = fma(a,b,-(c*d)) (common/math/math.h:367-372). */
static inline float prod_diff(float a, float b, float c, float d) { return fmaf(a, b, -(c * d)); }

/* cross(a,b) (common/math/vec3.h:209). */
static inline void cross3(const float a[3], const float b[3], float r[3]) {
  r[0] = prod_diff(a[1], b[2], a[2], b[1]);
  r[1] = prod_diff(a[2], b[0], a[0], b[2]);
  r[2] = prod_diff(a[0], b[1], a[1], b[0]);
}

/* dot(a,b) = madd(a.x,b.x,madd(a.y,b.y,a.z*b.z)) (common/math/vec3.h:204). */
static inline float dot3(const float a[3], const float b[3]) {
  return fmaf(a[0], b[0], fmaf(a[1], b[1], a[2] * b[2]));
}

/* halfArea(d) = madd(d.x,(d.y+d.z),d.y*d.z) (common/math/vec3fa.h:373). */
static inline float half_area_d(float dx, float dy, float dz) { return fmaf(dx, dy + dz, dy * dz); }

float oracle_half_area(const float lo[3], const float hi[3]) {
  return half_area_d(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]);
}

/* ------------------------------------------------------- triangle test -- */

/* MoellerTrumboreIntersector1::intersect (triangle_intersector_moeller.h:62-103)
 * with Ng = cross(e2,e1) (:105-114); hit finalisation per oracle.h. */
int oracle_moeller(const float org[3], const float dir[3], float tnear, float tfar,
                   const float v0[3], const float e1[3], const float e2[3], float out[8]) {
  float C[3] = {v0[0] - org[0], v0[1] - org[1], v0[2] - org[2]};
  float R[3], Ng[3];
  cross3(C, dir, R);
  cross3(e2, e1, Ng);
  const float den = dot3(Ng, dir);
  const float absDen = fabsf(den);
  const uint32_t sgn = f2u(den) & 0x80000000u;
  const float U = u2f(f2u(dot3(R, e2)) ^ sgn);
  const float V = u2f(f2u(dot3(R, e1)) ^ sgn);
  int valid = (den != 0.0f) & (U >= 0.0f) & (V >= 0.0f) & (U + V <= absDen);
  const float T = u2f(f2u(dot3(Ng, C)) ^ sgn);
  valid &= (absDen * tnear < T) & (T <= absDen * tfar);
  out[0] = T; out[1] = absDen; out[2] = U; out[3] = V;
  out[4] = Ng[0]; out[5] = Ng[1]; out[6] = Ng[2];
  out[7] = valid ? T / absDen : INFINITY;
  return valid;
}

int oracle_moeller_verts(const float org[3], const float dir[3], float tnear, float tfar,
                         const float p0[3], const float p1[3], const float p2[3], float out[8]) {
  /* TriangleM(v0,v1,v2): e1 = v0-v1, e2 = v2-v0 (kernels/geometry/triangle.h:40-41). */
  float e1[3] = {p0[0] - p1[0], p0[1] - p1[1], p0[2] - p1[2]};
  float e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
  return oracle_moeller(org, dir, tnear, tfar, p0, e1, e2, out);
}

/* PlueckerIntersector1::intersect (triangle_intersector_pluecker.h:55-110):
 * vertices relative to the origin, edge functions U, V, W, the
 * eps = ulp*|U+V+W| tolerance, stable_triangle_normal (common/math/vec3.h:210-222,
 * msub = fma(a,b,-c)), den and T doubled (math.h:242 twice). min/max of three
 * are min(min(a,b),c) with min(a,b) = a < b ? a : b (SSE semantics).
 * Hit finalisation (PlueckerHitM::finalize, :29-36): u = U/UVW, v = V/UVW,
 * or U*0, V*0 when |UVW| < min_rcp_input; t = T/den. IEEE division replaces
 * the reference's rcp for the same reason as in the Moeller test.
 * out[0..3] = t, U, V, UVW; out[4..6] = Ng. Returns validity for [tnear, tfar]. */
static inline float o_min(float a, float b) { return a < b ? a : b; }
static inline float o_max(float a, float b) { return a > b ? a : b; }

int oracle_pluecker(const float org[3], const float dir[3], float tnear, float tfar,
                    const float p0[3], const float p1[3], const float p2[3], float out[8]) {
  float v0[3], v1[3], v2[3], e0[3], e1[3], e2[3], s[3], c[3];
  for (int k = 0; k < 3; k++) {
    v0[k] = p0[k] - org[k];
    v1[k] = p1[k] - org[k];
    v2[k] = p2[k] - org[k];
  }
  for (int k = 0; k < 3; k++) {
    e0[k] = v2[k] - v0[k];
    e1[k] = v0[k] - v1[k];
    e2[k] = v1[k] - v2[k];
  }
  for (int k = 0; k < 3; k++) s[k] = v2[k] + v0[k];
  cross3(e0, s, c);
  const float U = dot3(c, dir);
  for (int k = 0; k < 3; k++) s[k] = v0[k] + v1[k];
  cross3(e1, s, c);
  const float V = dot3(c, dir);
  for (int k = 0; k < 3; k++) s[k] = v1[k] + v2[k];
  cross3(e2, s, c);
  const float W = dot3(c, dir);
  const float UVW = U + V + W;
  const float eps = FLT_EPSILON * fabsf(UVW);
  int valid = (o_min(o_min(U, V), W) >= -eps) | (o_max(o_max(U, V), W) <= eps);
  /* stable_triangle_normal(a=e0, b=e1, c=e2) */
  const float ab_x = e0[2] * e1[1], ab_y = e0[0] * e1[2], ab_z = e0[1] * e1[0];
  const float bc_x = e1[2] * e2[1], bc_y = e1[0] * e2[2], bc_z = e1[1] * e2[0];
  const float cab[3] = {fmaf(e0[1], e1[2], -ab_x), fmaf(e0[2], e1[0], -ab_y), fmaf(e0[0], e1[1], -ab_z)};
  const float cbc[3] = {fmaf(e1[1], e2[2], -bc_x), fmaf(e1[2], e2[0], -bc_y), fmaf(e1[0], e2[1], -bc_z)};
  float Ng[3];
  Ng[0] = fabsf(ab_x) < fabsf(bc_x) ? cab[0] : cbc[0];
  Ng[1] = fabsf(ab_y) < fabsf(bc_y) ? cab[1] : cbc[1];
  Ng[2] = fabsf(ab_z) < fabsf(bc_z) ? cab[2] : cbc[2];
  const float d = dot3(Ng, dir);
  const float den = d + d;
  const float tt = dot3(v0, Ng);
  const float T = tt + tt;
  const float t = T / den;
  valid &= (tnear <= t) & (t <= tfar);
  valid &= den != 0.0f;
  out[0] = t; out[1] = U; out[2] = V; out[3] = UVW;
  out[4] = Ng[0]; out[5] = Ng[1]; out[6] = Ng[2];
  out[7] = valid ? t : INFINITY;
  return valid;
}

/* ---------------------------------------------------- ray record access -- */

typedef struct {
  float org[3], tnear, dir[3], time, tfar;
  unsigned mask, id, flags;
  float Ng[3], u, v;
  unsigned primID, geomID, instID;
} o_rayhit; /* 80 bytes, RTCRayHit layout */

typedef struct {
  float t, u, v, Ng[3];
  unsigned geomID, primID;
  int found;
} o_best;

/* Lexicographic (t, geomID, primID) improvement test. */
static inline int o_better(float t, unsigned g, unsigned p, const o_best* b) {
  if (!b->found) return 1;
  if (t < b->t) return 1;
  if (t > b->t) return 0;
  if (g != b->geomID) return g < b->geomID;
  return p < b->primID;
}

static inline void o_consider(const float org[3], const float dir[3], float tnear, float tfar,
                              const float p0[3], const float e1[3], const float e2[3],
                              unsigned g, unsigned p, o_best* b) {
  float h[8];
  if (!oracle_moeller(org, dir, tnear, tfar, p0, e1, e2, h)) return;
  const float t = h[7];
  if (!o_better(t, g, p, b)) return;
  b->found = 1;
  b->t = t;
  b->u = h[2] / h[1];
  b->v = h[3] / h[1];
  b->Ng[0] = h[4]; b->Ng[1] = h[5]; b->Ng[2] = h[6];
  b->geomID = g; b->primID = p;
}

static inline void o_consider_pl(const float org[3], const float dir[3], float tnear, float tfar,
                                 const float p0[3], const float p1[3], const float p2[3], unsigned g,
                                 unsigned p, o_best* b) {
  float h[8];
  if (!oracle_pluecker(org, dir, tnear, tfar, p0, p1, p2, h)) return;
  const float t = h[0];
  if (!o_better(t, g, p, b)) return;
  const int tiny = fabsf(h[3]) < O_MIN_RCP_INPUT;
  b->found = 1;
  b->t = t;
  b->u = tiny ? h[1] * 0.0f : h[1] / h[3];
  b->v = tiny ? h[2] * 0.0f : h[2] / h[3];
  b->Ng[0] = h[4]; b->Ng[1] = h[5]; b->Ng[2] = h[6];
  b->geomID = g; b->primID = p;
}

/* Hit write-back (intersector_epilog.h:280-292). */
static inline void o_write_hit(o_rayhit* r, const o_best* b, unsigned inst_id) {
  r->tfar = b->t;
  r->Ng[0] = b->Ng[0]; r->Ng[1] = b->Ng[1]; r->Ng[2] = b->Ng[2];
  r->u = b->u; r->v = b->v;
  r->primID = b->primID; r->geomID = b->geomID; r->instID = inst_id;
}

/* ------------------------------------------------------ triangle access -- */

static inline int o_isvalid_v(const float* v) {
  return (v[0] > -O_FLT_LARGE) & (v[0] < O_FLT_LARGE) & (v[1] > -O_FLT_LARGE) &
         (v[1] < O_FLT_LARGE) & (v[2] > -O_FLT_LARGE) & (v[2] < O_FLT_LARGE);
}

int oracle_triangle_valid(const oracle_mesh* m, unsigned prim) {
  const unsigned* ix = m->indices + 3 * (size_t)prim;
  if (ix[0] >= m->num_vertices || ix[1] >= m->num_vertices || ix[2] >= m->num_vertices) return 0;
  return o_isvalid_v(m->vertices + 3 * (size_t)ix[0]) && o_isvalid_v(m->vertices + 3 * (size_t)ix[1]) &&
         o_isvalid_v(m->vertices + 3 * (size_t)ix[2]);
}

static inline void o_tri_edges(const oracle_mesh* m, unsigned prim, float p0[3], float e1[3], float e2[3]) {
  const unsigned* ix = m->indices + 3 * (size_t)prim;
  const float* a = m->vertices + 3 * (size_t)ix[0];
  const float* b = m->vertices + 3 * (size_t)ix[1];
  const float* c = m->vertices + 3 * (size_t)ix[2];
  for (int k = 0; k < 3; k++) {
    p0[k] = a[k];
    e1[k] = a[k] - b[k];
    e2[k] = c[k] - a[k];
  }
}

static inline void o_tri_verts(const oracle_mesh* m, unsigned prim, const float** a, const float** b,
                               const float** c) {
  const unsigned* ix = m->indices + 3 * (size_t)prim;
  *a = m->vertices + 3 * (size_t)ix[0];
  *b = m->vertices + 3 * (size_t)ix[1];
  *c = m->vertices + 3 * (size_t)ix[2];
}

static int o_threads(int nthreads) {
#ifdef _OPENMP
  return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

/* ---------------------------------------------------------- brute force -- */

static void o_intersect_brute(const oracle_mesh* meshes, unsigned nmeshes, void* rayhits, unsigned nrays,
                              size_t stride, unsigned inst_id, int nthreads, int robust) {
  int nt = o_threads(nthreads);
  (void)nt;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
  for (long i = 0; i < (long)nrays; i++) {
    o_rayhit* r = (o_rayhit*)((char*)rayhits + (size_t)i * stride);
    if (!(r->tnear <= r->tfar)) continue;
    o_best b;
    memset(&b, 0, sizeof(b));
    for (unsigned g = 0; g < nmeshes; g++) {
      const oracle_mesh* m = &meshes[g];
      for (unsigned p = 0; p < m->num_triangles; p++) {
        if (!oracle_triangle_valid(m, p)) continue;
        if (robust) {
          const float *a, *bb, *c;
          o_tri_verts(m, p, &a, &bb, &c);
          o_consider_pl(r->org, r->dir, r->tnear, r->tfar, a, bb, c, g, p, &b);
        } else {
          float p0[3], e1[3], e2[3];
          o_tri_edges(m, p, p0, e1, e2);
          o_consider(r->org, r->dir, r->tnear, r->tfar, p0, e1, e2, g, p, &b);
        }
      }
    }
    if (b.found) o_write_hit(r, &b, inst_id);
  }
}

static void o_occluded_brute(const oracle_mesh* meshes, unsigned nmeshes, void* rays, unsigned nrays,
                             size_t stride, int nthreads, int robust) {
  int nt = o_threads(nthreads);
  (void)nt;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
  for (long i = 0; i < (long)nrays; i++) {
    o_rayhit* r = (o_rayhit*)((char*)rays + (size_t)i * stride);
    if (!(r->tnear <= r->tfar)) continue;
    int hit = 0;
    for (unsigned g = 0; g < nmeshes && !hit; g++) {
      const oracle_mesh* m = &meshes[g];
      for (unsigned p = 0; p < m->num_triangles && !hit; p++) {
        if (!oracle_triangle_valid(m, p)) continue;
        float h[8];
        if (robust) {
          const float *a, *bb, *c;
          o_tri_verts(m, p, &a, &bb, &c);
          hit = oracle_pluecker(r->org, r->dir, r->tnear, r->tfar, a, bb, c, h);
        } else {
          float p0[3], e1[3], e2[3];
          o_tri_edges(m, p, p0, e1, e2);
          hit = oracle_moeller(r->org, r->dir, r->tnear, r->tfar, p0, e1, e2, h);
        }
      }
    }
    if (hit) r->tfar = -INFINITY;
  }
}

void oracle_intersect_brute(const oracle_mesh* meshes, unsigned nmeshes, void* rayhits,
                            unsigned nrays, size_t stride, unsigned inst_id, int nthreads) {
  o_intersect_brute(meshes, nmeshes, rayhits, nrays, stride, inst_id, nthreads, 0);
}

void oracle_occluded_brute(const oracle_mesh* meshes, unsigned nmeshes, void* rays,
                           unsigned nrays, size_t stride, int nthreads) {
  o_occluded_brute(meshes, nmeshes, rays, nrays, stride, nthreads, 0);
}

void oracle_intersect_brute_robust(const oracle_mesh* meshes, unsigned nmeshes, void* rayhits, unsigned nrays,
                                   size_t stride, unsigned inst_id, int nthreads) {
  o_intersect_brute(meshes, nmeshes, rayhits, nrays, stride, inst_id, nthreads, 1);
}

void oracle_occluded_brute_robust(const oracle_mesh* meshes, unsigned nmeshes, void* rays, unsigned nrays,
                                  size_t stride, int nthreads) {
  o_occluded_brute(meshes, nmeshes, rays, nrays, stride, nthreads, 1);
}

/* ------------------------------------------------------------ instances -- */

/* Vec3fa cross: shuffle<1,2,0,3>(prod_diff(a, b.yzx, a.yzx, b)) gives
 * (a.y*b.z - a.z*b.y, ...) with the same fma(a,b,-(c*d)) as cross3. */
static inline float dot_dpps(const float a[3], const float b[3]) {
  const float p0 = a[0] * b[0], p1 = a[1] * b[1], p2 = a[2] * b[2];
  return (p0 + p1) + (p2 + 0.0f);
}

void oracle_affine_inverse(const float l2w[12], float w2l[12]) {
  const float* vx = l2w;
  const float* vy = l2w + 3;
  const float* vz = l2w + 6;
  const float* p = l2w + 9;
  float c0[3], c1[3], c2[3];
  cross3(vy, vz, c0);
  cross3(vz, vx, c1);
  cross3(vx, vy, c2);
  const float det = dot_dpps(vx, c0);
  /* adjoint = LinearSpace3(c0, c1, c2).transposed(): column k = (c0[k], c1[k], c2[k]) */
  float il[9];
  for (int k = 0; k < 3; k++) {
    il[3 * k + 0] = c0[k] / det;
    il[3 * k + 1] = c1[k] / det;
    il[3 * k + 2] = c2[k] / det;
  }
  for (int k = 0; k < 9; k++) w2l[k] = il[k];
  for (int k = 0; k < 3; k++)
    w2l[9 + k] = -fmaf(p[0], il[k], fmaf(p[1], il[3 + k], p[2] * il[6 + k]));
}

void oracle_xfm_point(const float m[12], const float p[3], float out[3]) {
  for (int k = 0; k < 3; k++) out[k] = fmaf(p[0], m[k], fmaf(p[1], m[3 + k], fmaf(p[2], m[6 + k], m[9 + k])));
}

void oracle_xfm_vector(const float m[12], const float v[3], float out[3]) {
  for (int k = 0; k < 3; k++) out[k] = fmaf(v[0], m[k], fmaf(v[1], m[3 + k], v[2] * m[6 + k]));
}

/* Lexicographic (t, instID, geomID, primID) order across instances. */
static inline int o_better_inst(const o_best* a, unsigned ia, const o_best* b, unsigned ib) {
  if (!b->found) return 1;
  if (a->t < b->t) return 1;
  if (a->t > b->t) return 0;
  if (ia != ib) return ia < ib;
  if (a->geomID != b->geomID) return a->geomID < b->geomID;
  return a->primID < b->primID;
}

/* Closest hit in one object for a ray already in its space. */
static void o_object_best(const oracle_object* ob, const float org[3], const float dir[3], float tnear, float tfar,
                          o_best* b) {
  memset(b, 0, sizeof(*b));
  for (unsigned g = 0; g < ob->nmeshes; g++) {
    const oracle_mesh* m = &ob->meshes[g];
    for (unsigned p = 0; p < m->num_triangles; p++) {
      if (!oracle_triangle_valid(m, p)) continue;
      if (ob->robust) {
        const float *a, *bb, *c;
        o_tri_verts(m, p, &a, &bb, &c);
        o_consider_pl(org, dir, tnear, tfar, a, bb, c, g, p, b);
      } else {
        float p0[3], e1[3], e2[3];
        o_tri_edges(m, p, p0, e1, e2);
        o_consider(org, dir, tnear, tfar, p0, e1, e2, g, p, b);
      }
    }
  }
}

static int o_object_any(const oracle_object* ob, const float org[3], const float dir[3], float tnear, float tfar) {
  for (unsigned g = 0; g < ob->nmeshes; g++) {
    const oracle_mesh* m = &ob->meshes[g];
    for (unsigned p = 0; p < m->num_triangles; p++) {
      if (!oracle_triangle_valid(m, p)) continue;
      float h[8];
      if (ob->robust) {
        const float *a, *bb, *c;
        o_tri_verts(m, p, &a, &bb, &c);
        if (oracle_pluecker(org, dir, tnear, tfar, a, bb, c, h)) return 1;
      } else {
        float p0[3], e1[3], e2[3];
        o_tri_edges(m, p, p0, e1, e2);
        if (oracle_moeller(org, dir, tnear, tfar, p0, e1, e2, h)) return 1;
      }
    }
  }
  return 0;
}

/* The ray in an instance's object space (instance_intersector.cpp:59-61). */
static inline void o_inst_ray(const oracle_instance* in, const float org[3], const float dir[3], float o[3],
                              float d[3]) {
  if (in->identity) {
    memcpy(o, org, 12);
    memcpy(d, dir, 12);
  } else {
    oracle_xfm_point(in->w2l, org, o);
    oracle_xfm_vector(in->w2l, dir, d);
  }
}

void oracle_intersect_instances(const oracle_object* objects, const oracle_instance* insts, unsigned ninsts,
                                void* rayhits, unsigned nrays, size_t stride, unsigned ctx_inst, int nthreads) {
  int nt = o_threads(nthreads);
  (void)nt;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
  for (long i = 0; i < (long)nrays; i++) {
    o_rayhit* r = (o_rayhit*)((char*)rayhits + (size_t)i * stride);
    if (!(r->tnear <= r->tfar)) continue;
    o_best best;
    memset(&best, 0, sizeof(best));
    unsigned best_inst = 0xffffffffu;
    for (unsigned k = 0; k < ninsts; k++) {
      const oracle_instance* in = &insts[k];
      if (!in->identity && ctx_inst != 0xffffffffu) continue;
      const unsigned id = in->identity ? ctx_inst : in->instID;
      float o[3], d[3];
      o_inst_ray(in, r->org, r->dir, o, d);
      o_best b;
      o_object_best(&objects[in->object], o, d, r->tnear, r->tfar, &b);
      if (b.found && o_better_inst(&b, id, &best, best_inst)) {
        best = b;
        best_inst = id;
      }
    }
    if (best.found) o_write_hit(r, &best, best_inst);
  }
}

void oracle_occluded_instances(const oracle_object* objects, const oracle_instance* insts, unsigned ninsts,
                               void* rays, unsigned nrays, size_t stride, unsigned ctx_inst, int nthreads) {
  int nt = o_threads(nthreads);
  (void)nt;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
  for (long i = 0; i < (long)nrays; i++) {
    o_rayhit* r = (o_rayhit*)((char*)rays + (size_t)i * stride);
    if (!(r->tnear <= r->tfar)) continue;
    for (unsigned k = 0; k < ninsts; k++) {
      const oracle_instance* in = &insts[k];
      if (!in->identity && ctx_inst != 0xffffffffu) continue;
      float o[3], d[3];
      o_inst_ray(in, r->org, r->dir, o, d);
      if (o_object_any(&objects[in->object], o, d, r->tnear, r->tfar)) {
        r->tfar = -INFINITY;
        break;
      }
    }
  }
}

/* ------------------------------------------------------ binned SAH build -- */

typedef struct {
  float lo[3], hi[3];
  unsigned geomID, primID;
} o_primref; /* kernels/common/primref.h:24-80 */

typedef struct {
  float lo[3], hi[3];
} o_box;

static inline void box_empty(o_box* b) {
  b->lo[0] = b->lo[1] = b->lo[2] = INFINITY;
  b->hi[0] = b->hi[1] = b->hi[2] = -INFINITY;
}
static inline void box_extend_box(o_box* b, const float lo[3], const float hi[3]) {
  for (int k = 0; k < 3; k++) {
    b->lo[k] = fminf(b->lo[k], lo[k]);
    b->hi[k] = fmaxf(b->hi[k], hi[k]);
  }
}
static inline void box_extend_pt(o_box* b, const float p[3]) { box_extend_box(b, p, p); }
static inline float box_half_area(const o_box* b) { return oracle_half_area(b->lo, b->hi); }

/* PrimInfoRange: range plus geometry and centroid(x2) bounds. */
typedef struct {
  size_t begin, end;
  o_box geom, cent;
  unsigned depth;
} o_rec;

static void rec_compute(const o_primref* prims, o_rec* r) {
  box_empty(&r->geom);
  box_empty(&r->cent);
  for (size_t i = r->begin; i < r->end; i++) {
    float c2[3] = {prims[i].lo[0] + prims[i].hi[0], prims[i].lo[1] + prims[i].hi[1],
                   prims[i].lo[2] + prims[i].hi[2]};
    box_extend_box(&r->geom, prims[i].lo, prims[i].hi);
    box_extend_pt(&r->cent, c2);
  }
}

#define O_BINS 32

typedef struct {
  int dim;
  unsigned pos;
  float sah;
  unsigned num;
  float ofs[3], scale[3];
} o_split;

/* BinMapping (heuristic_binning.h:22-30). */
static void o_mapping(size_t N, const o_box* cent, o_split* s) {
  size_t num = (size_t)(4.0f + 0.05f * (float)N);
  s->num = (unsigned)(num < O_BINS ? num : O_BINS);
  const float eps = 1E-34f;
  for (int k = 0; k < 3; k++) {
    float diag = fmaxf(eps, cent->hi[k] - cent->lo[k]);
    s->scale[k] = diag > eps ? (0.99f * (float)s->num) / diag : 0.0f;
    s->ofs[k] = cent->lo[k];
  }
}

static inline int o_bin_of(const o_split* s, float c2, int k) {
  int b = (int)floorf((c2 - s->ofs[k]) * s->scale[k]);
  if (b < 0) b = 0;
  if (b >= (int)s->num) b = (int)s->num - 1;
  return b;
}

/* BinInfoT::bin (heuristic_binning.h:221-258) + best (:340-393), blocks of 4. */
static void o_find(const o_primref* prims, const o_rec* r, o_split* s) {
  const size_t N = r->end - r->begin;
  o_mapping(N, &r->cent, s);
  unsigned counts[O_BINS][3];
  o_box bounds[O_BINS][3];
  for (unsigned i = 0; i < s->num; i++)
    for (int k = 0; k < 3; k++) {
      counts[i][k] = 0;
      box_empty(&bounds[i][k]);
    }
  for (size_t i = r->begin; i < r->end; i++) {
    const o_primref* p = &prims[i];
    for (int k = 0; k < 3; k++) {
      const float c2 = p->lo[k] + p->hi[k];
      const int b = o_bin_of(s, c2, k);
      counts[b][k]++;
      box_extend_box(&bounds[b][k], p->lo, p->hi);
    }
  }
  float rAreas[O_BINS][3];
  unsigned rCounts[O_BINS][3];
  unsigned cnt[3] = {0, 0, 0};
  o_box bb[3];
  for (int k = 0; k < 3; k++) box_empty(&bb[k]);
  for (unsigned i = s->num - 1; i > 0; i--) {
    for (int k = 0; k < 3; k++) {
      cnt[k] += counts[i][k];
      rCounts[i][k] = cnt[k];
      box_extend_box(&bb[k], bounds[i][k].lo, bounds[i][k].hi);
      rAreas[i][k] = box_half_area(&bb[k]);
    }
  }
  float bestSAH[3] = {INFINITY, INFINITY, INFINITY};
  unsigned bestPos[3] = {0, 0, 0};
  cnt[0] = cnt[1] = cnt[2] = 0;
  for (int k = 0; k < 3; k++) box_empty(&bb[k]);
  for (unsigned i = 1; i < s->num; i++) {
    for (int k = 0; k < 3; k++) {
      cnt[k] += counts[i - 1][k];
      box_extend_box(&bb[k], bounds[i - 1][k].lo, bounds[i - 1][k].hi);
      const float lArea = box_half_area(&bb[k]);
      const float rArea = rAreas[i][k];
      const unsigned lC = (cnt[k] + 3) >> 2;
      const unsigned rC = (rCounts[i][k] + 3) >> 2;
      const float sah = fmaf(lArea, (float)lC, rArea * (float)rC);
      if (sah < bestSAH[k]) {
        bestPos[k] = i;
        bestSAH[k] = sah;
      }
    }
  }
  s->dim = -1;
  s->pos = 0;
  s->sah = INFINITY;
  for (int k = 0; k < 3; k++) {
    if (s->scale[k] == 0.0f) continue;
    if (bestSAH[k] < s->sah && bestPos[k] != 0) {
      s->dim = k;
      s->pos = bestPos[k];
      s->sah = bestSAH[k];
    }
  }
}

int oracle_binned_split(const float* lo, const float* hi, unsigned n, unsigned* pos, float* sah,
                        unsigned* num_bins) {
  o_primref* prims = (o_primref*)malloc(sizeof(o_primref) * (n ? n : 1));
  for (unsigned i = 0; i < n; i++) {
    memcpy(prims[i].lo, lo + 3 * (size_t)i, 12);
    memcpy(prims[i].hi, hi + 3 * (size_t)i, 12);
    prims[i].geomID = 0;
    prims[i].primID = i;
  }
  o_rec r;
  r.begin = 0; r.end = n; r.depth = 0;
  rec_compute(prims, &r);
  o_split s;
  o_find(prims, &r, &s);
  free(prims);
  *pos = s.pos;
  *sah = s.sah;
  *num_bins = s.num;
  return s.dim;
}

/* PrimRef ordering used by deterministic_order: by primID then geomID
 * (64-bit ID with primID in the upper half, kernels/common/primref.h). */
static int o_primref_cmp(const void* a, const void* b) {
  const o_primref* x = (const o_primref*)a;
  const o_primref* y = (const o_primref*)b;
  uint64_t ix = ((uint64_t)x->primID << 32) | x->geomID;
  uint64_t iy = ((uint64_t)y->primID << 32) | y->geomID;
  return ix < iy ? -1 : ix > iy ? 1 : 0;
}

static void o_split_set(o_primref* prims, const o_rec* r, const o_split* s, o_rec* L, o_rec* R,
                        o_primref* tmp) {
  size_t mid;
  if (s->dim < 0) {
    /* splitFallback after deterministic_order (heuristic_binning_array_aligned.h:82-85,122-141). */
    qsort(prims + r->begin, r->end - r->begin, sizeof(o_primref), o_primref_cmp);
    mid = (r->begin + r->end) / 2;
  } else {
    /* partition by bin(center2)[dim] < pos; stable (order does not change the sets). */
    size_t nl = 0, nr = 0, n = r->end - r->begin;
    o_primref* t = tmp + r->begin; /* range-local: subtrees may be built concurrently */
    for (size_t i = r->begin; i < r->end; i++) {
      const float c2 = prims[i].lo[s->dim] + prims[i].hi[s->dim];
      if (o_bin_of(s, c2, s->dim) < (int)s->pos) prims[r->begin + nl++] = prims[i];
      else t[nr++] = prims[i];
    }
    memcpy(prims + r->begin + nl, t, nr * sizeof(o_primref));
    (void)n;
    mid = r->begin + nl;
  }
  L->begin = r->begin; L->end = mid; L->depth = r->depth + 1;
  R->begin = mid; R->end = r->end; R->depth = r->depth + 1;
  rec_compute(prims, L);
  rec_compute(prims, R);
}

/* ------------------------------------------------------------ BVH8 tree -- */

typedef struct {
  float lo[8][3], hi[8][3];
  int ref[8]; /* >=0: inner node index; <0: ~leaf index */
  int n;
} o_node;

typedef struct {
  size_t begin, count;
} o_leaf;

struct oracle_bvh {
  o_primref* prims; /* leaf order */
  size_t nprims;
  o_node* nodes;
  size_t nnodes, cap_nodes;
  o_leaf* leaves;
  size_t nleaves, cap_leaves;
  int root; /* node index, or ~leaf, or INT32_MIN when empty */
  o_box bounds;
  float eps;
  unsigned max_depth;
  const oracle_mesh* meshes; /* not owned; vertex data for the triangle test */
  unsigned nmeshes;
  /* leaf triangle data in leaf order: p0, e1, e2 */
  float* tri;
};

/* Nodes and leaves of one subtree under construction. Subtrees of large
 * ranges are built concurrently, each into an arena of its own, and appended
 * to their parent's arena in child order, so the tree (and its numbering in
 * pre-order) does not depend on the thread count or scheduling. */
typedef struct {
  o_node* nodes;
  size_t nnodes, cap_nodes;
  o_leaf* leaves;
  size_t nleaves, cap_leaves;
  unsigned max_depth;
} o_arena;

typedef struct {
  o_primref* prims; /* reordered in place; ranges of concurrent subtrees are disjoint */
  o_primref* tmp;   /* partition scratch, indexed like prims */
} o_build;

static int o_new_node(o_arena* b) {
  if (b->nnodes == b->cap_nodes) {
    b->cap_nodes = b->cap_nodes ? 2 * b->cap_nodes : 1024;
    b->nodes = (o_node*)realloc(b->nodes, b->cap_nodes * sizeof(o_node));
  }
  memset(&b->nodes[b->nnodes], 0, sizeof(o_node));
  return (int)b->nnodes++;
}

static int o_new_leaf(o_arena* b, size_t begin, size_t count) {
  if (b->nleaves == b->cap_leaves) {
    b->cap_leaves = b->cap_leaves ? 2 * b->cap_leaves : 1024;
    b->leaves = (o_leaf*)realloc(b->leaves, b->cap_leaves * sizeof(o_leaf));
  }
  b->leaves[b->nleaves].begin = begin;
  b->leaves[b->nleaves].count = count;
  return ~(int)b->nleaves++;
}

static int o_shift_ref(int ref, size_t node_ofs, size_t leaf_ofs) {
  return ref >= 0 ? ref + (int)node_ofs : ~(~ref + (int)leaf_ofs);
}

/* Appends subtree arena s to a; returns s's root reference renumbered. */
static int o_append(o_arena* a, o_arena* s, int root) {
  const size_t no = a->nnodes, lo = a->nleaves;
  for (size_t i = 0; i < s->nnodes; i++) {
    int id = o_new_node(a);
    a->nodes[id] = s->nodes[i];
    for (int c = 0; c < a->nodes[id].n; c++) a->nodes[id].ref[c] = o_shift_ref(a->nodes[id].ref[c], no, lo);
  }
  for (size_t i = 0; i < s->nleaves; i++) o_new_leaf(a, s->leaves[i].begin, s->leaves[i].count);
  if (s->max_depth > a->max_depth) a->max_depth = s->max_depth;
  free(s->nodes);
  free(s->leaves);
  return o_shift_ref(root, no, lo);
}

/* ranges of at least this many references build their children as tasks */
#define O_TASK_MIN 16384

#define O_MIN_LEAF 4
#define O_MAX_LEAF 28
#define O_MAX_DEPTH 40
#define O_BRANCH 8

static int o_children_cmp(const void* a, const void* b) {
  size_t na = ((const o_rec*)a)->end - ((const o_rec*)a)->begin;
  size_t nb = ((const o_rec*)b)->end - ((const o_rec*)b)->begin;
  return na > nb ? -1 : na < nb ? 1 : 0; /* std::greater: larger first */
}

static int o_make_node(o_arena* b, o_rec* ch, int n) {
  int id = o_new_node(b);
  b->nodes[id].n = n;
  for (int i = 0; i < n; i++) {
    for (int k = 0; k < 3; k++) {
      b->nodes[id].lo[i][k] = ch[i].geom.lo[k];
      b->nodes[id].hi[i][k] = ch[i].geom.hi[k];
    }
  }
  return id;
}

/* createLargeLeaf (kernels/builders/bvh_builder_sah.h:155-220). */
static int o_large_leaf(const o_build* bd, o_arena* b, o_rec* cur) {
  size_t N = cur->end - cur->begin;
  if (cur->depth > b->max_depth) b->max_depth = cur->depth;
  if (N <= O_MAX_LEAF) return o_new_leaf(b, cur->begin, N);
  o_rec ch[O_BRANCH];
  int n = 1;
  ch[0] = *cur;
  do {
    int best = -1;
    size_t bestSize = 0;
    for (int i = 0; i < n; i++) {
      size_t sz = ch[i].end - ch[i].begin;
      if (sz <= O_MAX_LEAF) continue;
      if (sz > bestSize) { bestSize = sz; best = i; }
    }
    if (best < 0) break;
    o_split fb;
    fb.dim = -1;
    o_rec L, R;
    L.depth = R.depth = cur->depth + 1;
    o_split_set(bd->prims, &ch[best], &fb, &L, &R, bd->tmp);
    L.depth = R.depth = cur->depth + 1;
    ch[best] = ch[n - 1];
    ch[n - 1] = L;
    ch[n] = R;
    n++;
  } while (n < O_BRANCH);
  int id = o_make_node(b, ch, n);
  for (int i = 0; i < n; i++) {
    int r = o_large_leaf(bd, b, &ch[i]);
    b->nodes[id].ref[i] = r;
  }
  return id;
}

/* BuilderT::recurse (kernels/builders/bvh_builder_sah.h:222-330); the
 * reference also recurses into the children in parallel above its
 * singleThreadThreshold (:320-327). */
static int o_recurse(const o_build* bd, o_arena* b, o_rec* cur) {
  const size_t N = cur->end - cur->begin;
  if (cur->depth > b->max_depth) b->max_depth = cur->depth;
  o_split s;
  o_find(bd->prims, cur, &s);
  const float ha = box_half_area(&cur->geom);
  const float leafSAH = 1.0f * (ha * (float)((N + 3) >> 2));
  const float splitSAH = 1.0f * ha + 1.0f * s.sah;
  if (N <= O_MIN_LEAF || cur->depth + 8 >= O_MAX_DEPTH || (N <= O_MAX_LEAF && leafSAH <= splitSAH)) {
    qsort(bd->prims + cur->begin, N, sizeof(o_primref), o_primref_cmp);
    return o_large_leaf(bd, b, cur);
  }
  o_rec ch[O_BRANCH];
  o_split_set(bd->prims, cur, &s, &ch[0], &ch[1], bd->tmp);
  int n = 2;
  while (n < O_BRANCH) {
    float bestArea = -INFINITY;
    int best = -1;
    for (int i = 0; i < n; i++) {
      if (ch[i].end - ch[i].begin <= O_MIN_LEAF) continue;
      float a = box_half_area(&ch[i].geom);
      if (a > bestArea) { best = i; bestArea = a; }
    }
    if (best < 0) break;
    o_split cs;
    o_find(bd->prims, &ch[best], &cs);
    o_rec L, R;
    o_split_set(bd->prims, &ch[best], &cs, &L, &R, bd->tmp);
    L.depth = R.depth = cur->depth + 1;
    ch[best] = L;
    ch[n++] = R;
  }
  qsort(ch, n, sizeof(o_rec), o_children_cmp);
  int id = o_make_node(b, ch, n);
  if (N < O_TASK_MIN) {
    for (int i = 0; i < n; i++) {
      int r = o_recurse(bd, b, &ch[i]);
      b->nodes[id].ref[i] = r;
    }
    return id;
  }
  o_arena sub[O_BRANCH];
  int root[O_BRANCH];
  memset(sub, 0, sizeof(sub));
  for (int i = 0; i < n; i++) {
#pragma omp task default(none) firstprivate(i) shared(bd, sub, root, ch)
    root[i] = o_recurse(bd, &sub[i], &ch[i]);
  }
#pragma omp taskwait
  for (int i = 0; i < n; i++) {
    int r = o_append(b, &sub[i], root[i]);
    b->nodes[id].ref[i] = r;
  }
  return id;
}

oracle_bvh* oracle_build_sah(const oracle_mesh* meshes, unsigned nmeshes) {
  return oracle_build_sah_threads(meshes, nmeshes, 0);
}

oracle_bvh* oracle_build_sah_threads(const oracle_mesh* meshes, unsigned nmeshes, int nthreads) {
  oracle_bvh* b = (oracle_bvh*)calloc(1, sizeof(oracle_bvh));
  b->meshes = meshes;
  b->nmeshes = nmeshes;
  size_t total = 0;
  for (unsigned g = 0; g < nmeshes; g++) total += meshes[g].num_triangles;
  b->prims = (o_primref*)malloc(sizeof(o_primref) * (total ? total : 1));
  float maxabs = 0.0f;
  /* createPrimRefArray (primrefgen.cpp:36-56, scene_triangle_mesh.h:220-231). */
  for (unsigned g = 0; g < nmeshes; g++) {
    const oracle_mesh* m = &meshes[g];
    for (unsigned p = 0; p < m->num_triangles; p++) {
      if (!oracle_triangle_valid(m, p)) continue;
      const unsigned* ix = m->indices + 3 * (size_t)p;
      o_primref* r = &b->prims[b->nprims++];
      for (int k = 0; k < 3; k++) {
        float a = m->vertices[3 * (size_t)ix[0] + k], c = m->vertices[3 * (size_t)ix[1] + k],
              d = m->vertices[3 * (size_t)ix[2] + k];
        r->lo[k] = fminf(a, fminf(c, d));
        r->hi[k] = fmaxf(a, fmaxf(c, d));
        maxabs = fmaxf(maxabs, fmaxf(fabsf(r->lo[k]), fabsf(r->hi[k])));
      }
      r->geomID = g;
      r->primID = p;
    }
  }
  b->eps = maxabs * O_BOX_EPS_REL;
  if (b->nprims == 0) {
    b->root = INT32_MIN;
    box_empty(&b->bounds);
    return b;
  }
  o_rec root;
  root.begin = 0; root.end = b->nprims; root.depth = 1;
  rec_compute(b->prims, &root);
  b->bounds = root.geom;
  o_build bd;
  bd.prims = b->prims;
  bd.tmp = (o_primref*)malloc(sizeof(o_primref) * b->nprims);
  o_arena a;
  memset(&a, 0, sizeof(a));
  int top = 0;
#pragma omp parallel num_threads(o_threads(nthreads)) default(none) shared(bd, a, root, top)
#pragma omp single
  top = o_recurse(&bd, &a, &root);
  free(bd.tmp);
  b->root = top;
  b->nodes = a.nodes;
  b->nnodes = a.nnodes;
  b->cap_nodes = a.cap_nodes;
  b->leaves = a.leaves;
  b->nleaves = a.nleaves;
  b->cap_leaves = a.cap_leaves;
  b->max_depth = a.max_depth;
  /* widen every stored child box by eps (conservative culling). */
  for (size_t i = 0; i < b->nnodes; i++)
    for (int c = 0; c < b->nodes[i].n; c++)
      for (int k = 0; k < 3; k++) {
        b->nodes[i].lo[c][k] -= b->eps;
        b->nodes[i].hi[c][k] += b->eps;
      }
  /* leaf triangle data in leaf order */
  b->tri = (float*)malloc(sizeof(float) * 9 * b->nprims);
  for (size_t i = 0; i < b->nprims; i++)
    o_tri_edges(&meshes[b->prims[i].geomID], b->prims[i].primID, b->tri + 9 * i, b->tri + 9 * i + 3,
                b->tri + 9 * i + 6);
  return b;
}

void oracle_bvh_free(oracle_bvh* b) {
  if (!b) return;
  free(b->prims);
  free(b->nodes);
  free(b->leaves);
  free(b->tri);
  free(b);
}

static void o_stats_rec(const oracle_bvh* b, int ref, const float lo[3], const float hi[3], double* nodeSAH,
                        double* leafSAH, double* leafSAHTri) {
  const float ha = oracle_half_area(lo, hi);
  if (ref >= 0) {
    *nodeSAH += ha;
    const o_node* n = &b->nodes[ref];
    for (int i = 0; i < n->n; i++) {
      float l[3], h[3];
      for (int k = 0; k < 3; k++) { l[k] = n->lo[i][k] + b->eps; h[k] = n->hi[i][k] - b->eps; }
      o_stats_rec(b, n->ref[i], l, h, nodeSAH, leafSAH, leafSAHTri);
    }
  } else {
    size_t cnt = b->leaves[~ref].count;
    *leafSAH += (double)ha * (double)((cnt + 3) / 4);
    *leafSAHTri += (double)ha * (double)cnt;
  }
}

void oracle_bvh_stats(const oracle_bvh* b, double out[8]) {
  memset(out, 0, 8 * sizeof(double));
  out[0] = (double)b->nnodes;
  out[1] = (double)b->nleaves;
  out[2] = (double)b->nprims;
  out[3] = (double)b->max_depth;
  if (b->root == INT32_MIN) return;
  double nodeSAH = 0, leafSAH = 0, leafSAHTri = 0;
  o_stats_rec(b, b->root, b->bounds.lo, b->bounds.hi, &nodeSAH, &leafSAH, &leafSAHTri);
  const double root = oracle_half_area(b->bounds.lo, b->bounds.hi);
  out[4] = (nodeSAH + leafSAH) / root;
  out[5] = root;
  out[6] = (nodeSAH + leafSAHTri) / root;
}

static int o_validate_rec(const oracle_bvh* b, int ref, const float* lo, const float* hi, unsigned char* seen) {
  if (ref >= 0) {
    if ((size_t)ref >= b->nnodes) return 10;
    const o_node* n = &b->nodes[ref];
    if (n->n < 1 || n->n > O_BRANCH) return 11;
    for (int i = 0; i < n->n; i++) {
      if (lo && hi)
        for (int k = 0; k < 3; k++)
          if (n->lo[i][k] < lo[k] || n->hi[i][k] > hi[k]) return 12;
      int r = o_validate_rec(b, n->ref[i], n->lo[i], n->hi[i], seen);
      if (r) return r;
    }
    return 0;
  }
  if ((size_t)(~ref) >= b->nleaves) return 20;
  const o_leaf* l = &b->leaves[~ref];
  if (l->count < 1 || l->count > O_MAX_LEAF) return 21;
  for (size_t i = l->begin; i < l->begin + l->count; i++) {
    if (i >= b->nprims) return 22;
    if (seen[i]) return 23;
    seen[i] = 1;
    if (lo && hi)
      for (int k = 0; k < 3; k++)
        if (b->prims[i].lo[k] < lo[k] || b->prims[i].hi[k] > hi[k]) return 24;
  }
  return 0;
}

int oracle_bvh_validate(const oracle_bvh* b) {
  if (b->root == INT32_MIN) return b->nprims == 0 ? 0 : 1;
  unsigned char* seen = (unsigned char*)calloc(b->nprims ? b->nprims : 1, 1);
  int r = o_validate_rec(b, b->root, NULL, NULL, seen);
  for (size_t i = 0; !r && i < b->nprims; i++)
    if (!seen[i]) r = 30;
  free(seen);
  return r;
}

/* ------------------------------------------------------------ traversal -- */

static inline float widen_up(float x) { return isinf(x) ? x : x + fabsf(x) * O_MARGIN; }
static inline float widen_down(float x) { return isinf(x) ? x : x - fabsf(x) * O_MARGIN; }

typedef struct {
  float org[3], dir[3];
  float rdn[3], rdf[3];
  int near_hi[3]; /* 1 if the near plane on this axis is the upper bound */
} o_tray;

/* Robust TravRay (node_intersector1.h:122-190) with zero_fix (vec3fa.h:172-176). */
static void o_tray_init(o_tray* t, const float org[3], const float dir[3]) {
  const float round_down = 1.0f - 3.0f * FLT_EPSILON;
  const float round_up = 1.0f + 3.0f * FLT_EPSILON;
  for (int k = 0; k < 3; k++) {
    t->org[k] = org[k];
    t->dir[k] = dir[k];
    float d = fabsf(dir[k]) < O_MIN_RCP_INPUT ? O_MIN_RCP_INPUT : dir[k];
    float rd = 1.0f / d;
    t->rdn[k] = round_down * rd;
    t->rdf[k] = round_up * rd;
    t->near_hi[k] = !(t->rdn[k] >= 0.0f);
  }
}

/* Slab test of child c; returns 1 and tNear when [tlo, thi] overlaps. */
static inline int o_box_hit(const o_tray* t, const o_node* n, int c, float tlo, float thi, float* dist) {
  float tn = tlo, tf = thi;
  for (int k = 0; k < 3; k++) {
    const float pn = t->near_hi[k] ? n->hi[c][k] : n->lo[c][k];
    const float pf = t->near_hi[k] ? n->lo[c][k] : n->hi[c][k];
    const float a = (pn - t->org[k]) * t->rdn[k];
    const float bb = (pf - t->org[k]) * t->rdf[k];
    tn = fmaxf(tn, a);
    tf = fminf(tf, bb);
  }
  *dist = tn;
  return tn <= tf;
}

typedef struct {
  int ref;
  float dist;
} o_entry;

#define O_STACK 1024

static void o_trace_one(const oracle_bvh* b, o_rayhit* r, int occl, unsigned inst_id, unsigned* cn,
                        unsigned* ct, int robust) {
  if (!(r->tnear <= r->tfar)) return;
  if (b->root == INT32_MIN) return;
  o_tray t;
  o_tray_init(&t, r->org, r->dir);
  const float tnear = r->tnear, tfar = r->tfar;
  const float tlo = widen_down(tnear);
  float thi = widen_up(tfar);
  o_best best;
  memset(&best, 0, sizeof(best));
  o_entry stack[O_STACK];
  int sp = 0;
  stack[sp].ref = b->root;
  stack[sp].dist = -INFINITY;
  sp++;
  unsigned nodes = 0, tris = 0;
  while (sp > 0) {
    o_entry e = stack[--sp];
    if (e.dist > thi) continue;
    if (e.ref >= 0) {
      const o_node* n = &b->nodes[e.ref];
      nodes++;
      o_entry hits[8];
      int nh = 0;
      for (int c = 0; c < n->n; c++) {
        float d;
        if (o_box_hit(&t, n, c, tlo, thi, &d)) {
          hits[nh].ref = n->ref[c];
          hits[nh].dist = d;
          nh++;
        }
      }
      /* push far-to-near so the nearest child is popped first */
      for (int i = 1; i < nh; i++)
        for (int j = i; j > 0 && hits[j].dist > hits[j - 1].dist; j--) {
          o_entry x = hits[j]; hits[j] = hits[j - 1]; hits[j - 1] = x;
        }
      for (int i = 0; i < nh && sp < O_STACK; i++) stack[sp++] = hits[i];
    } else {
      const o_leaf* l = &b->leaves[~e.ref];
      for (size_t i = l->begin; i < l->begin + l->count; i++) {
        const float* tr = b->tri + 9 * i;
        tris++;
        const float *pa = NULL, *pb = NULL, *pc = NULL;
        if (robust) o_tri_verts(&b->meshes[b->prims[i].geomID], b->prims[i].primID, &pa, &pb, &pc);
        if (occl) {
          float h[8];
          if (robust ? oracle_pluecker(r->org, r->dir, tnear, tfar, pa, pb, pc, h)
                     : oracle_moeller(r->org, r->dir, tnear, tfar, tr, tr + 3, tr + 6, h)) {
            r->tfar = -INFINITY;
            if (cn) { *cn = nodes; *ct = tris; }
            return;
          }
        } else if (robust) {
          o_consider_pl(r->org, r->dir, tnear, tfar, pa, pb, pc, b->prims[i].geomID, b->prims[i].primID, &best);
        } else {
          o_consider(r->org, r->dir, tnear, tfar, tr, tr + 3, tr + 6, b->prims[i].geomID, b->prims[i].primID,
                     &best);
        }
      }
      if (!occl && best.found) thi = widen_up(best.t < tfar ? best.t : tfar);
    }
  }
  if (!occl && best.found) o_write_hit(r, &best, inst_id);
  if (cn) { *cn = nodes; *ct = tris; }
}

void oracle_intersect_bvh(const oracle_bvh* b, void* rayhits, unsigned nrays, size_t stride,
                          unsigned inst_id, int nthreads) {
  int nt = o_threads(nthreads);
  (void)nt;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
  for (long i = 0; i < (long)nrays; i++)
    o_trace_one(b, (o_rayhit*)((char*)rayhits + (size_t)i * stride), 0, inst_id, NULL, NULL, 0);
}

void oracle_intersect_bvh_robust(const oracle_bvh* b, void* rayhits, unsigned nrays, size_t stride,
                                 unsigned inst_id, int nthreads) {
  int nt = o_threads(nthreads);
  (void)nt;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
  for (long i = 0; i < (long)nrays; i++)
    o_trace_one(b, (o_rayhit*)((char*)rayhits + (size_t)i * stride), 0, inst_id, NULL, NULL, 1);
}

void oracle_occluded_bvh_robust(const oracle_bvh* b, void* rays, unsigned nrays, size_t stride, int nthreads) {
  int nt = o_threads(nthreads);
  (void)nt;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
  for (long i = 0; i < (long)nrays; i++)
    o_trace_one(b, (o_rayhit*)((char*)rays + (size_t)i * stride), 1, 0, NULL, NULL, 1);
}

void oracle_occluded_bvh(const oracle_bvh* b, void* rays, unsigned nrays, size_t stride, int nthreads) {
  int nt = o_threads(nthreads);
  (void)nt;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
  for (long i = 0; i < (long)nrays; i++)
    o_trace_one(b, (o_rayhit*)((char*)rays + (size_t)i * stride), 1, 0, NULL, NULL, 0);
}

void oracle_intersect_bvh_counted(const oracle_bvh* b, void* rayhits, unsigned nrays, size_t stride,
                                  unsigned* nodes, unsigned* tris, int nthreads) {
  int nt = o_threads(nthreads);
  (void)nt;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
  for (long i = 0; i < (long)nrays; i++) {
    nodes[i] = tris[i] = 0;
    o_trace_one(b, (o_rayhit*)((char*)rayhits + (size_t)i * stride), 0, 0xFFFFFFFFu, &nodes[i],
                &tris[i], 0);
  }
}

/* --------------------------------------------------------------- Morton -- */

/* bitInterleave (common/math/math.h:390-409). */
unsigned oracle_bit_interleave(unsigned x, unsigned y, unsigned z) {
  x = (x | (x << 16)) & 0x030000FF;
  x = (x | (x << 8)) & 0x0300F00F;
  x = (x | (x << 4)) & 0x030C30C3;
  x = (x | (x << 2)) & 0x09249249;
  y = (y | (y << 16)) & 0x030000FF;
  y = (y | (y << 8)) & 0x0300F00F;
  y = (y | (y << 4)) & 0x030C30C3;
  y = (y | (y << 2)) & 0x09249249;
  z = (z | (z << 16)) & 0x030000FF;
  z = (z | (z << 8)) & 0x0300F00F;
  z = (z | (z << 4)) & 0x030C30C3;
  z = (z | (z << 2)) & 0x09249249;
  return x | (y << 1) | (z << 2);
}

/* MortonCodeMapping over centroid-x2 bounds (bvh_builder_morton.h:70-101);
 * rcp(diag) taken as IEEE 1/diag. */
void oracle_morton_codes(const float* lo, const float* hi, unsigned n, unsigned* codes) {
  o_box cb;
  box_empty(&cb);
  for (unsigned i = 0; i < n; i++) {
    float c2[3] = {lo[3 * i] + hi[3 * i], lo[3 * i + 1] + hi[3 * i + 1], lo[3 * i + 2] + hi[3 * i + 2]};
    box_extend_pt(&cb, c2);
  }
  float scale[3];
  for (int k = 0; k < 3; k++) {
    float diag = cb.hi[k] - cb.lo[k];
    scale[k] = diag > 1E-19f ? (1.0f / diag) * (1024 * 0.99f) : 0.0f;
  }
  for (unsigned i = 0; i < n; i++) {
    unsigned q[3];
    for (int k = 0; k < 3; k++) {
      float c2 = lo[3 * i + k] + hi[3 * i + k];
      q[k] = (unsigned)(int)((c2 - cb.lo[k]) * scale[k]);
    }
    codes[i] = oracle_bit_interleave(q[0], q[1], q[2]);
  }
}

/* Instances over per-object BVHs: same result as the brute-force instance
 * functions (each object BVH's result equals its brute force), for scenes too
 * large to brute-force. objects[k] is the BVH of object k and robust[k]
 * selects its triangle test. */
void oracle_intersect_instances_bvh(const oracle_bvh* const* objects, const int* robust,
                                    const oracle_instance* insts, unsigned ninsts, void* rayhits, unsigned nrays,
                                    size_t stride, unsigned ctx_inst, int nthreads) {
  int nt = o_threads(nthreads);
  (void)nt;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
  for (long i = 0; i < (long)nrays; i++) {
    o_rayhit* r = (o_rayhit*)((char*)rayhits + (size_t)i * stride);
    if (!(r->tnear <= r->tfar)) continue;
    o_best best;
    memset(&best, 0, sizeof(best));
    unsigned best_inst = 0xffffffffu;
    for (unsigned k = 0; k < ninsts; k++) {
      const oracle_instance* in = &insts[k];
      if (!in->identity && ctx_inst != 0xffffffffu) continue;
      const unsigned id = in->identity ? ctx_inst : in->instID;
      o_rayhit q = *r;
      o_inst_ray(in, r->org, r->dir, q.org, q.dir);
      q.geomID = 0xffffffffu;
      o_trace_one(objects[in->object], &q, 0, id, NULL, NULL, robust[in->object]);
      if (q.geomID == 0xffffffffu) continue;
      o_best b;
      b.found = 1;
      b.t = q.tfar;
      b.u = q.u;
      b.v = q.v;
      b.Ng[0] = q.Ng[0]; b.Ng[1] = q.Ng[1]; b.Ng[2] = q.Ng[2];
      b.geomID = q.geomID;
      b.primID = q.primID;
      if (o_better_inst(&b, id, &best, best_inst)) {
        best = b;
        best_inst = id;
      }
    }
    if (best.found) o_write_hit(r, &best, best_inst);
  }
}

void oracle_occluded_instances_bvh(const oracle_bvh* const* objects, const int* robust,
                                   const oracle_instance* insts, unsigned ninsts, void* rays, unsigned nrays,
                                   size_t stride, unsigned ctx_inst, int nthreads) {
  int nt = o_threads(nthreads);
  (void)nt;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
  for (long i = 0; i < (long)nrays; i++) {
    o_rayhit* r = (o_rayhit*)((char*)rays + (size_t)i * stride);
    if (!(r->tnear <= r->tfar)) continue;
    for (unsigned k = 0; k < ninsts; k++) {
      const oracle_instance* in = &insts[k];
      if (!in->identity && ctx_inst != 0xffffffffu) continue;
      o_rayhit q = *r;
      o_inst_ray(in, r->org, r->dir, q.org, q.dir);
      o_trace_one(objects[in->object], &q, 1, 0, NULL, NULL, robust[in->object]);
      if (q.tfar == -INFINITY) {
        r->tfar = -INFINITY;
        break;
      }
    }
  }
}
