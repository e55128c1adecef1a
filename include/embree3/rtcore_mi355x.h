/* Extensions of the MI355X triangle ray-query library.
 *
 * None of these has a counterpart in the reference API. They exist for
 * callers that keep rays resident in GPU memory (a renderer's wavefront
 * queue, a benchmark, a training loop) and for the workload generators the
 * tests and benchmarks use. Every function reports failures through the
 * device error slot like the standard API. */
#pragma once

#include "rtcore.h"

RTC_NAMESPACE_BEGIN

/* Device memory on the device's GPU (hipMalloc / hipFree). */
RTC_API void* rtcgDeviceMalloc(RTCDevice device, size_t bytes);
RTC_API void rtcgDeviceFree(RTCDevice device, void* ptr);

/* Synchronous copy between any two host or device pointers (hipMemcpyDefault). */
RTC_API void rtcgMemcpy(RTCDevice device, void* dst, const void* src, size_t bytes);

/* Same contract as rtcIntersect1M / rtcOccluded1M, but the records live in
 * device memory and the call is asynchronous on `hipStream` (a hipStream_t;
 * NULL selects the calling thread's stream of this device). */
RTC_API void rtcgIntersect1MAsync(RTCScene scene, const struct RTCIntersectContext* context, struct RTCRayHit* d_rayhit,
                                  unsigned int M, size_t byteStride, void* hipStream);
RTC_API void rtcgOccluded1MAsync(RTCScene scene, const struct RTCIntersectContext* context, struct RTCRay* d_ray,
                                 unsigned int M, size_t byteStride, void* hipStream);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the query of rtcgIntersect1MAsync (occluded = 0,
 * RTCRayHit records) or rtcgOccluded1MAsync (occluded = 1, RTCRay records)
 * over records in this device's memory, with the same results, allocating
 * nothing. Its kernels keep their ray counter and error bits in a fixed
 * 256-byte array per GPU that the library holds until the process ends.
 * At most 2^31 rays per call (more is RTC_ERROR_INVALID_ARGUMENT); host
 * records are RTC_ERROR_INVALID_ARGUMENT; the device's ray-reordering setting
 * is not applied. Every call on one GPU shares that GPU's array, so calls that
 * overlap on one GPU (other streams, other threads) race: the caller orders
 * them. hipStream NULL: synchronous on the null stream, and reports traversal
 * errors, including those left by earlier asynchronous calls on this GPU;
 * a caller's hipStream_t: asynchronous. */
RTC_API void rtcgQuery1MNoAlloc(RTCScene scene, const struct RTCIntersectContext* context, void* d_rays,
                                unsigned int M, size_t byteStride, int occluded, void* hipStream);

/* Message of the calling thread's most recent error (any device); the
 * pointer stays valid until the thread's next error. */
RTC_API const char* rtcgGetErrorString(void);

/* Waits for the calling thread's stream of this device. */
RTC_API void rtcgSynchronize(RTCDevice device);

/* Timing harness: `steps` times, copies M records from d_src to d_rays
 * (untimed) and traces d_rays with HIP events recorded around the trace
 * kernel on the stream it runs on; per-step milliseconds go to ms_out. */
RTC_API void rtcgTimeTrace(RTCScene scene, void* d_rays, const void* d_src, unsigned int M, size_t byteStride,
                           int occluded, int steps, float* ms_out);

/* Timing harness for the scene build: rebuilds the committed scene `steps`
 * times and reports per-build milliseconds (wall clock, build fully
 * synchronised). */
RTC_API void rtcgTimeBuild(RTCScene scene, int steps, float* ms_out);

/* Counting trace: like rtcgIntersect1MAsync/rtcgOccluded1MAsync (synchronous)
 * but also writes per-ray node visits and triangle tests to the device
 * arrays d_nodes[M] and d_tris[M]. */
RTC_API void rtcgTraceCounted(RTCScene scene, void* d_rays, unsigned int M, size_t byteStride, int occluded,
                              unsigned int* d_nodes, unsigned int* d_tris);

/* Built acceleration structure summary. */
struct RTCGSceneStats
{
  unsigned int num_prims;      /* valid triangles in the BVH */
  unsigned int num_nodes;      /* 8-wide nodes */
  unsigned int max_depth;      /* node levels */
  unsigned int node_bytes;     /* bytes a visit reads of a node record (80; a node sits in a 128 B slot) */
  unsigned int tri_bytes;      /* bytes per triangle record */
  unsigned int max_stack;      /* traversal stack entries the deepest path can need */
  double sah;                  /* SAH statistic, leaf cost per block of 4 triangles */
  double sah_tri;              /* SAH statistic, leaf cost per triangle */
  double build_ms;             /* last commit's build time */
};
RTC_API void rtcgGetSceneStats(RTCScene scene, struct RTCGSceneStats* stats);

/* Copies the device BVH to host memory: nodes packed at node_bytes (80 B)
 * each and triangle records (48 B each); sizes from rtcgGetSceneStats.
 * Either pointer may be NULL. */
RTC_API void rtcgReadBVH(RTCScene scene, void* nodes, void* tris);

/* ---- workload generators (host memory, deterministic) ---------------- */

/* UV sphere of the reference tutorial generator with numPhi rings:
 * numTheta = 2*numPhi, vertices numTheta*(numPhi+1), triangles
 * 2*numTheta*(numPhi-1). Returns the triangle count; pass NULL buffers to
 * query sizes (vertex count written to *numVertices). */
RTC_API unsigned int rtcgGenSphere(unsigned int numPhi, float cx, float cy, float cz, float radius,
                                   float* vertices, unsigned int* indices, unsigned int* numVertices);

/* Pinhole camera rays (closest-hit records, tnear 0, tfar +inf) for a
 * width x height image from `org` looking down -z with the given vertical
 * field of view in degrees. */
RTC_API void rtcgGenPinholeRays(unsigned int width, unsigned int height, float ox, float oy, float oz,
                                float fovDeg, struct RTCRayHit* rays);

/* Random triangle soup: centroids uniform in [0,1]^3, vertex offsets
 * uniform in [-h,h]^3; 3 unique vertices per triangle (indexed). */
RTC_API void rtcgGenSoup(unsigned int numTris, unsigned long long seed, float h, float* vertices,
                         unsigned int* indices);

/* Incoherent rays: origins uniform in [lo,hi]^3, directions uniform on the
 * sphere; tnear 0 and the given tfar. */
RTC_API void rtcgGenRandomRays(unsigned int numRays, unsigned long long seed, float lo, float hi, float tfar,
                               struct RTCRayHit* rays);

/* Heightfield: n x n vertices over [0,1]^2, z = amplitude * Perlin
 * gradient-noise fBm (improved noise with a quintic fade; octaves at
 * frequency 4*2^o, weight 2^-o); returns 2*(n-1)^2 triangles
 * (NULL buffers: size query). Vertices n*n*3 floats, indices tris*3. */
RTC_API unsigned int rtcgGenHeightfield(unsigned int n, unsigned long long seed, float amplitude,
                                        unsigned int octaves, float* vertices, unsigned int* indices);

/* Look-at pinhole camera rays (closest-hit records) for width x height pixels. */
RTC_API void rtcgGenCameraRays(unsigned int width, unsigned int height, const float* org, const float* target,
                               const float* up, float fovDeg, struct RTCRayHit* rays);

/* Cosine-weighted next-bounce rays from the hits among M traced records;
 * returns how many were written (misses spawn none). */
RTC_API unsigned int rtcgGenBounceRays(const struct RTCRayHit* hits, unsigned int M, unsigned long long seed,
                                       float epsilon, struct RTCRayHit* out);

/* Unit UV sphere (numPhi rings) displaced radially by amplitude times a
 * 3-octave value noise; sizes as rtcgGenSphere. */
RTC_API unsigned int rtcgGenDisplacedSphere(unsigned int numPhi, unsigned long long seed, float amplitude,
                                            float* vertices, unsigned int* indices, unsigned int* numVertices);

/* Rays from uniform points on a sphere of the given radius towards uniform
 * points inside the unit ball. */
RTC_API void rtcgGenSphereRays(unsigned int numRays, unsigned long long seed, float radius, struct RTCRayHit* rays);

/* ---- autotuner ------------------------------------------------------- */
/* No counterpart in the reference (its BVH and traversal settings are fixed
 * per ISA). Times traversal configurations on M device-resident sample rays
 * (closest hit, or occlusion records) of a committed scene and keeps the
 * fastest as the device's setting: LDS stack entries (0, 8, 12, 16), refill
 * threshold (8..32 idle lanes) and resident blocks per CU (4, 6, 8); with
 * rebuildTree != 0 first the tree shape, rebuilding the scene for each node
 * width (4, 8), leaf size (4..12) and collapse triangle cost, ending with the
 * best. The sample rays are not modified. Writes the chosen settings as an
 * rtcNewDevice config string into out (outSize bytes) and returns the best
 * Mrays/s. */
RTC_API float rtcgAutotune(RTCScene scene, const void* rays, unsigned int M, size_t byteStride, int occluded,
                           int rebuildTree, char* out, size_t outSize);

/* ---- batched scenes ------------------------------------------------- */
/* No counterpart in the reference (each RTCScene commits on its own,
 * kernels/common/scene.cpp:655-750); for callers with hundreds of scenes.
 * Builds every scene in one pass (one segmented sort and one binned-SAH
 * forest over all their triangles, sharing one node and triangle array).
 * Each scene then answers rtcIntersect1M/rtcOccluded1M as if committed on
 * its own; a later rtcCommitScene gives it arrays of its own again. All
 * scenes must belong to one device and agree on RTC_SCENE_FLAG_ROBUST. */
RTC_API void rtcgCommitScenes(RTCScene* scenes, unsigned int numScenes);

/* One ray stream of a batched traversal: device records (RTCRayHit for
 * closest hit, RTCRay for occlusion) of a scene of one rtcgCommitScenes. */
struct RTCGBatchItem
{
  RTCScene scene;
  void* rays;
  unsigned int numRays;
  unsigned int byteStride;
};

/* Traces every stream in one persistent launch (lanes pull rays from one
 * atomic counter over the concatenated streams); records updated in place
 * as rtcIntersect1M/rtcOccluded1M would. Synchronous. */
RTC_API void rtcgTraceBatch(const struct RTCGBatchItem* items, unsigned int numItems, int occluded);

/* Launch settings for the batched traversal, tuned on the batch itself:
 * times rtcgTraceBatch's launch over the streams (on a scratch copy: the
 * records are only read) for each LDS stack ring (8 or 12 entries per
 * lane), refill threshold (8..24 idle lanes) and resident blocks per CU (4,
 * 6, 8), keeps the fastest as the setting of the scenes' device (batch
 * graphs made afterwards capture it), writes it as an rtcNewDevice config
 * string into out (outSize bytes) and returns the best Mrays/s. Needs twice
 * the streams' bytes of scratch device memory. No counterpart in the
 * reference. */
RTC_API float rtcgAutotuneBatch(const struct RTCGBatchItem* items, unsigned int numItems, int occluded, char* out,
                                size_t outSize);

/* The same traversal captured once in a hipGraph (counter clear + launch;
 * the stream table stays on the device), for callers that replay it. The
 * graph keeps the scenes and their device arrays alive; a launch after any
 * of its scenes was committed again (rtcCommitScene, another
 * rtcgCommitScenes, rtcgAutotune's rebuilds) is RTC_ERROR_INVALID_OPERATION:
 * make a new graph. The ray buffers must stay valid. A launch on
 * a NULL stream is synchronous and reports traversal errors; on a caller's
 * hipStream_t it is asynchronous. Replays of one graph must not overlap. */
typedef struct RTCGBatchGraphTy* RTCGBatchGraph;
RTC_API RTCGBatchGraph rtcgNewBatchGraph(const struct RTCGBatchItem* items, unsigned int numItems, int occluded);
RTC_API void rtcgLaunchBatchGraph(RTCGBatchGraph graph, void* hipStream);
RTC_API void rtcgReleaseBatchGraph(RTCGBatchGraph graph);

/* ---- learned split cost --------------------------------------------- */
/* No counterpart in the reference: its binned-SAH split score
 * (kernels/builders/heuristic_binning.h:340-393, BinInfoT::best) is the plain
 * cost these functions correct. A split model is a small FP16 MLP (12
 * features -> 32 -> 32 -> 1, FP32 accumulate) on the per-bin statistics of a
 * split candidate: for a child C (one side's bins: box and primitive count)
 * of a segment S (box and count) it predicts y = log2 P(C | S) -
 * log2(A_C / A_S), how far the fraction of the target distribution's rays
 * that enter S and then C departs from the SAH's area ratio. The binned-SAH
 * builder then prices a split A_L 2^y_L n_L + A_R 2^y_R n_R (y held within
 * [-3, 3] and to P <= 1) for segments with more than min_count (default 64)
 * primitives. The model is trained on the GPU from the node-entry counts of
 * the library's own traversal of a binary (width-2) tree of the scene. Hits
 * do not depend on the tree, so scenes built with a model give the same
 * records.
 * Features of C in S, scene box (slo, sext): log2(A_C/A_S + 2^-30)/16,
 * n_C/n_S, (centre_C - centre_S)/extent_S (3), extent_C/extent_S (3),
 * (centre_S - slo)/sext (3), log2(A_S/A_scene + 2^-30)/16. */
typedef struct RTCGSplitModelTy* RTCGSplitModel;

RTC_API RTCGSplitModel rtcgNewSplitModel(RTCDevice device, unsigned int seed);
RTC_API void rtcgRetainSplitModel(RTCGSplitModel model);
RTC_API void rtcgReleaseSplitModel(RTCGSplitModel model);

/* Parameter vector (FP32 master copy): W1[32][12] b1[32] W2[32][32] b2[32] w3[32] b3. */
RTC_API unsigned int rtcgSplitModelParamCount(void);
RTC_API void rtcgSplitModelSetParams(RTCGSplitModel model, const float* params);
RTC_API void rtcgSplitModelGetParams(RTCGSplitModel model, float* params);

/* Inference on n feature vectors (host arrays, n x 12 in, n out). */
RTC_API void rtcgSplitModelPredict(RTCGSplitModel model, const float* features, unsigned int n, float* out);

/* Loss sum_i w_i (y_i - t_i)^2 / n of the FP16 forward pass and its FP32
 * gradient with respect to the parameter vector (host arrays). */
RTC_API float rtcgSplitModelGradient(RTCGSplitModel model, const float* features, const float* targets,
                                     const float* weights, unsigned int n, float* grad);

/* Builds a binary (width-2) binned-SAH tree of the committed scene's
 * triangles, traces M closest-hit records (device memory, not modified)
 * through it counting the rays that enter each node, and appends one
 * training sample per (parent, internal child) pair whose parent at least 8
 * rays entered: the child's features in the parent and the target
 * log2((v_C + 0.5) / (v_S + 0.5)) - log2(A_C / A_S). */
RTC_API void rtcgSplitModelCollect(RTCGSplitModel model, RTCScene scene, const void* d_rays, unsigned int M,
                                   size_t byteStride);
/* The same for a committed two-level scene (at least one instance): with
 * object NULL, samples for the scene's top tree over its instances (the
 * tree the scene's own split model prices); else for the tree of `object`,
 * a scene the top tree instances, from the visits of all its instances in
 * object space. The records are traced through the two-level scene with a
 * width-2 build of that one tree in its place. Neither tree is changed. */
RTC_API void rtcgSplitModelCollectInstanced(RTCGSplitModel model, RTCScene scene, RTCScene object,
                                            const void* d_rays, unsigned int M, size_t byteStride);
RTC_API unsigned int rtcgSplitModelSampleCount(RTCGSplitModel model);
RTC_API void rtcgSplitModelGetSamples(RTCGSplitModel model, float* features, float* targets);
RTC_API void rtcgSplitModelClearSamples(RTCGSplitModel model);

/* Adam on random minibatches of the collected samples, on the GPU; returns
 * the last minibatch loss. */
RTC_API float rtcgSplitModelTrain(RTCGSplitModel model, unsigned int steps, unsigned int batch,
                                  float learningRate, unsigned int seed);
RTC_API void rtcgSetSplitModelMinCount(RTCGSplitModel model, unsigned int minCount);
/* The model only chooses among split candidates whose plain binned-SAH cost
 * is within (1 + slack) of the segment's best (default 0.05; INFINITY lets it
 * choose freely). Negative or NaN slack is RTC_ERROR_INVALID_ARGUMENT. */
RTC_API void rtcgSetSplitModelSahSlack(RTCGSplitModel model, float slack);

/* Later commits of the scene use the model's split cost (binned-SAH builder);
 * NULL restores the plain cost. */
RTC_API void rtcgSetSceneSplitModel(RTCScene scene, RTCGSplitModel model);

RTC_NAMESPACE_END
