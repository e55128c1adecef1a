/* oracle.h -- CPU restatement of the reference triangle ray-query path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the parity tests,
 * the package smoke check and the CPU-baseline leg of the benchmark compare
 * the GPU library against. Nothing in the product links, loads or calls it.
 *
 * What it restates (reference files, relative to the reference root):
 *   - Pluecker triangle test of robust scenes, same operation sequence:
 *       kernels/geometry/triangle_intersector_pluecker.h:23-110,
 *       stable_triangle_normal common/math/vec3.h:210-222, ulp = FLT_EPSILON
 *       (common/math/constants.h:130-136).
 *   - Moeller-Trumbore triangle test, exact FP32 operation sequence:
 *       kernels/geometry/triangle_intersector_moeller.h:62-103 (1-ray form),
 *       :288-330 (k-th ray of a packet, same arithmetic), cross/dot as
 *       common/math/vec3.h:204,209 with prod_diff common/math/math.h:367-372;
 *       triangle edges e1 = v0-v1, e2 = v2-v0 as kernels/geometry/triangle.h:40-41.
 *   - Closest-hit epilog fields: kernels/geometry/intersector_epilog.h:215-295.
 *   - Stream entry rules (rays with !(tnear <= tfar) untouched; occluded ->
 *     tfar = -inf): kernels/common/rtcore.cpp:595-623, :848-874,
 *     kernels/bvh/bvh_intersector_stream_filters.cpp:135-151,
 *     kernels/bvh/bvh_intersector1.cpp:191-193.
 *   - Triangle validity (index range, finite vertices with |x| < 1.844e18):
 *     kernels/common/scene_triangle_mesh.h:127-147, PrimRef generation
 *     kernels/common/scene_triangle_mesh.h:220-231, kernels/common/primref.h:24-80.
 *   - Binned SAH builder, BVH8 with largest-area child expansion:
 *     kernels/builders/bvh_builder_sah.h:155-330, heuristic_binning.h:16-112,
 *     :211-258, :340-393, heuristic_binning_array_aligned.h:37-143,
 *     priminfo.h:139-147, bbox.h/vec3fa.h halfArea, kernels/bvh/bvh_builder_sah.cpp:560-565
 *     (sahBlockSize 4, intCost 1, minLeaf 4, maxLeaf 28), bvh_builder.cpp:11-21
 *     (branching 8, maxDepth 40).
 *   - SAH statistic: kernels/bvh/bvh_statistics.cpp:52-56,139-146,
 *     bvh_statistics.h:36-38,99-101.
 *   - Robust slab test (rdir scaled by 1 -/+ 3 ulp): kernels/bvh/node_intersector1.h:122-190,
 *     zero_fix / min_rcp_input: common/math/vec3fa.h:172-176, common/math/constants.h:31.
 *   - Morton codes: kernels/builders/bvh_builder_morton.h:70-101,
 *     bitInterleave common/math/math.h:390-409.
 *
 * Deliberate, documented differences from the reference (the GPU library
 * follows the same rules, so both are order-independent):
 *   - t, u, v are T/absDen, U/absDen, V/absDen with IEEE division; the
 *     reference multiplies by rcp(absDen) (estimate + one Newton step on x86,
 *     kernels/geometry/triangle_intersector_moeller.h:29-35), which is up to
 *     about 1 ulp lower (single-triangle known answer t = 0.99999994 vs 1).
 *   - Closest hit is the lexicographic minimum of (t, geomID, primID) over
 *     all valid Moeller hits; the reference keeps the last visited of equal
 *     t, which depends on its tree.
 *   - The triangle test uses the ray's original [tnear, tfar]; the
 *     reference shrinks tfar to the current hit, so its choice among hits
 *     within a few ulp depends on visit order.
 *   - BVH culling is conservative (robust slab test, boxes widened by an
 *     absolute epsilon, a relative margin of 2^-12 on the current hit
 *     distance), so the BVH result equals the brute-force result. The
 *     reference's non-robust slab test (node_intersector1.h:526-575) can cull
 *     a box holding a hit its own triangle test accepts.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One indexed triangle mesh: tightly packed FLOAT3 vertices, UINT3 indices. */
typedef struct {
  const float* vertices;
  unsigned num_vertices;
  const unsigned* indices;
  unsigned num_triangles;
} oracle_mesh;

/* Moeller test of one ray against one triangle given v0 and edges e1, e2.
 * out[0..3] = T, absDen, U, V (sign-corrected); out[4..6] = Ng;
 * out[7] = t = T/absDen. Returns 1 when the hit is valid for [tnear, tfar]. */
int oracle_moeller(const float org[3], const float dir[3], float tnear, float tfar,
                   const float v0[3], const float e1[3], const float e2[3], float out[8]);

/* Same, from the three vertices (edges formed as e1=v0-v1, e2=v2-v0). */
int oracle_moeller_verts(const float org[3], const float dir[3], float tnear, float tfar,
                         const float p0[3], const float p1[3], const float p2[3], float out[8]);

/* Pluecker (robust) test of one ray against the triangle p0, p1, p2
 * (kernels/geometry/triangle_intersector_pluecker.h:55-110, finalize :29-36).
 * out[0..3] = t, U, V, U+V+W; out[4..6] = Ng (stable_triangle_normal,
 * common/math/vec3.h:210-222). t = T/den, u = U/UVW, v = V/UVW with IEEE
 * division (the reference uses rcp). Returns 1 when tnear <= t <= tfar. */
int oracle_pluecker(const float org[3], const float dir[3], float tnear, float tfar,
                    const float p0[3], const float p1[3], const float p2[3], float out[8]);

/* 1 if triangle prim of mesh is valid (index range and finite vertices). */
int oracle_triangle_valid(const oracle_mesh* m, unsigned prim);

/* Brute-force closest hit over every triangle of every mesh; rays are
 * RTCRayHit records (80 B) spaced `stride` bytes apart, updated in place.
 * inst_id is written to instID[0] on a hit. nthreads<=0 uses all cores. */
void oracle_intersect_brute(const oracle_mesh* meshes, unsigned nmeshes, void* rayhits,
                            unsigned nrays, size_t stride, unsigned inst_id, int nthreads);

/* Brute-force occlusion; rays are RTCRay records (48 B) spaced `stride`. */
void oracle_occluded_brute(const oracle_mesh* meshes, unsigned nmeshes, void* rays,
                           unsigned nrays, size_t stride, int nthreads);

/* Same with the Pluecker test (RTC_SCENE_FLAG_ROBUST scenes). */
void oracle_intersect_brute_robust(const oracle_mesh* meshes, unsigned nmeshes, void* rayhits,
                                   unsigned nrays, size_t stride, unsigned inst_id, int nthreads);
void oracle_occluded_brute_robust(const oracle_mesh* meshes, unsigned nmeshes, void* rays,
                                  unsigned nrays, size_t stride, int nthreads);

/* ---- instances (two-level scenes) ----
 * world2local = rcp(local2world) as Instance::commit computes it
 * (kernels/common/scene_instance.cpp commit, common/math/affinespace.h:83):
 * il = adjoint/det (common/math/linearspace3.h:44-53 with the column-vector
 * cross of common/math/vec3fa.h:358-365 and the SSE4.1 dpps dot of
 * vec3fa.h:349-351, i.e. (x0*y0 + x1*y1) + x2*y2), p' = -(il * p)
 * (linearspace3.h:146). Matrices are 12 floats, columns vx | vy | vz | p. */
void oracle_affine_inverse(const float l2w[12], float w2l[12]);

/* xfmPoint / xfmVector (common/math/affinespace.h:102-103,
 * common/math/linearspace3.h:156): fma(x, vx, fma(y, vy, fma(z, vz, p))) and
 * fma(x, vx, fma(y, vy, z * vz)) per component. */
void oracle_xfm_point(const float m[12], const float p[3], float out[3]);
void oracle_xfm_vector(const float m[12], const float v[3], float out[3]);

/* An instanced object: the triangle meshes of a committed scene (geomID =
 * index in meshes) and whether that scene is robust (Pluecker test). */
typedef struct {
  const oracle_mesh* meshes;
  unsigned nmeshes;
  int robust;
} oracle_object;

/* One instance: its world2local transform, object and instID. identity = 1
 * marks the parent scene's own triangles (no transform; hits report the
 * context's instID[0]). */
typedef struct {
  float w2l[12];
  unsigned object;
  unsigned instID;
  int identity;
} oracle_instance;

/* Brute-force closest hit over every instance: the ray moves into each
 * instance's object space (kernels/geometry/instance_intersector.cpp:52-71),
 * keeps [tnear, tfar], and the hit is the lexicographic minimum of
 * (t, instID, geomID, primID); Ng stays in object space. When ctx_inst is not
 * RTC_INVALID_GEOMETRY_ID (0xffffffff) the one-level instance stack is full
 * and only identity instances are tested (kernels/common/instance_stack.h:21-38). */
void oracle_intersect_instances(const oracle_object* objects, const oracle_instance* insts, unsigned ninsts,
                                void* rayhits, unsigned nrays, size_t stride, unsigned ctx_inst, int nthreads);
void oracle_occluded_instances(const oracle_object* objects, const oracle_instance* insts, unsigned ninsts,
                               void* rays, unsigned nrays, size_t stride, unsigned ctx_inst, int nthreads);

typedef struct oracle_bvh oracle_bvh;

/* Builds the reference-style binned-SAH BVH8 over all valid triangles. The
 * meshes array and the data it points to must outlive the BVH. */
oracle_bvh* oracle_build_sah(const oracle_mesh* meshes, unsigned nmeshes);
/* The same tree built on nthreads threads (0: all); subtrees are numbered in
 * pre-order whatever the thread count. */
oracle_bvh* oracle_build_sah_threads(const oracle_mesh* meshes, unsigned nmeshes, int nthreads);
void oracle_bvh_free(oracle_bvh* bvh);

/* Statistics: [0] inner nodes, [1] leaves, [2] primitives, [3] max depth,
 * [4] SAH cost (reference statistic, leaf cost in blocks of 4),
 * [5] root half area, [6] SAH cost with leaf cost per triangle. */
void oracle_bvh_stats(const oracle_bvh* bvh, double out[8]);

/* Structural check: every valid primitive in exactly one leaf, leaves of 1..28
 * primitives, 1..8 children per inner node, stored child boxes containing
 * their subtree's primitive boxes. Returns 0 when all hold, else a code. */
int oracle_bvh_validate(const oracle_bvh* bvh);

/* BVH closest hit / occlusion with the conservative rules above. */
void oracle_intersect_bvh(const oracle_bvh* bvh, void* rayhits, unsigned nrays, size_t stride,
                          unsigned inst_id, int nthreads);
void oracle_occluded_bvh(const oracle_bvh* bvh, void* rays, unsigned nrays, size_t stride,
                         int nthreads);

/* Same with the Pluecker test. */
void oracle_intersect_bvh_robust(const oracle_bvh* bvh, void* rayhits, unsigned nrays, size_t stride,
                                 unsigned inst_id, int nthreads);
void oracle_occluded_bvh_robust(const oracle_bvh* bvh, void* rays, unsigned nrays, size_t stride,
                                int nthreads);

/* Per-ray traversal counters of the last oracle_intersect_bvh call pattern:
 * fills nodes[i] and tris[i] (node visits and triangle tests per ray). */
void oracle_intersect_bvh_counted(const oracle_bvh* bvh, void* rayhits, unsigned nrays,
                                  size_t stride, unsigned* nodes, unsigned* tris, int nthreads);

/* Morton code pieces (reference bvh_builder_morton.h:70-101). */
unsigned oracle_bit_interleave(unsigned x, unsigned y, unsigned z);
/* codes[i] for PrimRef bounds lo/hi (n x 3 floats each), mapping over the
 * centroid-times-two bounds of all n boxes. */
void oracle_morton_codes(const float* lo, const float* hi, unsigned n, unsigned* codes);

/* Reference halfArea of a box: fma(dx, dy+dz, dy*dz). */
float oracle_half_area(const float lo[3], const float hi[3]);

/* Binned SAH split search over PrimRef boxes (reference BinInfoT::bin/best):
 * returns dim (-1 if none), writes pos and sah. Used by the learned-cost
 * tests as the plain-SAH judge on single nodes. */
int oracle_binned_split(const float* lo, const float* hi, unsigned n, unsigned* pos, float* sah,
                        unsigned* num_bins);

/* Instances over per-object oracle BVHs (objects[k] built over object k's
 * meshes): the same results as oracle_intersect_instances /
 * oracle_occluded_instances, fast enough for large scenes. */
void oracle_intersect_instances_bvh(const oracle_bvh* const* objects, const int* robust,
                                    const oracle_instance* insts, unsigned ninsts, void* rayhits, unsigned nrays,
                                    size_t stride, unsigned ctx_inst, int nthreads);
void oracle_occluded_instances_bvh(const oracle_bvh* const* objects, const int* robust,
                                   const oracle_instance* insts, unsigned ninsts, void* rays, unsigned nrays,
                                   size_t stride, unsigned ctx_inst, int nthreads);

#ifdef __cplusplus
}
#endif
