// C API of the MI355X triangle ray-query library.
//
// Object model and error semantics follow the reference's API layer
// (kernels/common/rtcore.cpp, rtcore.h:28-122): opaque reference-counted
// handles, every entry point converts failures into an error code stored in
// a per-device, per-thread first-error slot (kernels/common/device.cpp:258-285,
// state.h:162-175) and reported to the optional error callback. The compute
// behind rtcCommitScene and the ray queries runs on the GPU (builder.hip,
// trace.hip); there is no CPU execution path.
#define RTC_EXPORT_API
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "builder.h"
#include "embree3/rtcore.h"
#include "embree3/rtcore_mi355x.h"
#include "trace.h"

namespace {

struct rtc_error {
  RTCError code;
  std::string msg;
};
#define RTC_THROW(c, m) throw rtc_error{c, m}

struct RefObj {
  std::atomic<long> refs{1};
  virtual ~RefObj() {}
};
inline void ref_inc(RefObj* o) { o->refs.fetch_add(1); }
inline void ref_dec(RefObj* o) {
  if (o->refs.fetch_sub(1) == 1) delete o;
}

// ---------------------------------------------------------------- device --

struct Device;

struct ThreadDev {
  hipStream_t stream = nullptr;
  rtg::DevBuf staging;
  rtg::DevBuf errword;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // host records: two pinned chunks the queries of host memory move through
  rtg::PinnedBuf stage[2];
  hipEvent_t ev_stage[2] = {nullptr, nullptr};
  rtg::RaySorter sorter;
  rtg::DevBuf batch, batch_counter;  // batched traversal: stream table, work counter
};

struct Device : RefObj {
  int ordinal = 0;
  int verbose = 0;
  rtg::BuildParams build;
  rtg::TraceParams trace;
  RTCErrorFunction errfn = nullptr;
  void* errptr = nullptr;
  RTCMemoryMonitorFunction memfn = nullptr;
  void* memptr = nullptr;
  std::mutex build_mtx;
  rtg::BuildScratch scratch;
  // tri_builder / builder setting: "default" lets the build quality choose
  // (reference bvh8_factory.cpp:865-870); an unknown name fails the commit.
  enum { TB_DEFAULT, TB_SAH, TB_MORTON, TB_UNKNOWN } tri_builder = TB_DEFAULT;
  std::string tri_builder_name = "default";
  // HIGH quality: spatial splits add up to this fraction of extra triangle
  // references (the reference's spatial-split SAH at HIGH, bvh8_factory.cpp:867-870)
  float spatial_budget = 0.3f;
  std::mutex tls_mtx;
  std::unordered_map<std::thread::id, ThreadDev*> threads;
  ~Device() override {
    (void)hipSetDevice(ordinal);
    for (auto& kv : threads) {
      ThreadDev* t = kv.second;
      if (t->stream) (void)hipStreamDestroy(t->stream);
      if (t->ev0) (void)hipEventDestroy(t->ev0);
      if (t->ev1) (void)hipEventDestroy(t->ev1);
      for (hipEvent_t e : t->ev_stage)
        if (e) (void)hipEventDestroy(e);
      delete t;
    }
  }
  void activate() { RTG_HIP(hipSetDevice(ordinal)); }
  ThreadDev& td() {
    std::lock_guard<std::mutex> g(tls_mtx);
    auto it = threads.find(std::this_thread::get_id());
    if (it != threads.end()) return *it->second;
    ThreadDev* t = new ThreadDev();
    RTG_HIP(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
    RTG_HIP(hipEventCreate(&t->ev0));
    RTG_HIP(hipEventCreate(&t->ev1));
    for (hipEvent_t& e : t->ev_stage) RTG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    t->errword.reserve(256);
    RTG_HIP(hipMemsetAsync(t->errword.ptr, 0, 4, t->stream));
    threads[std::this_thread::get_id()] = t;
    return *t;
  }
  // build and trace settings: written by rtcNewDevice and rtcgAutotune,
  // read by commits and queries on any thread (copies taken under the lock)
  mutable std::mutex settings_mtx;
  rtg::TraceParams trace_params() const {
    std::lock_guard<std::mutex> g(settings_mtx);
    return trace;
  }
  rtg::BuildParams build_settings() const {
    std::lock_guard<std::mutex> g(settings_mtx);
    return build;
  }
  void monitor(ssize_t bytes) {
    if (memfn && !memfn(memptr, bytes, false)) RTC_THROW(RTC_ERROR_OUT_OF_MEMORY, "memory monitor forced termination");
  }
};

// per-thread first-error slots (reference: device.cpp:258-285)
thread_local std::unordered_map<const Device*, RTCError> tls_errors;
thread_local RTCError tls_null_error = RTC_ERROR_NONE;
thread_local std::string tls_last_message;

void process_error(Device* d, RTCError code, const char* msg) {
  tls_last_message = msg ? msg : "";
  if (!d) {
    if (tls_null_error == RTC_ERROR_NONE) tls_null_error = code;
    return;
  }
  RTCError& e = tls_errors[d];
  if (e == RTC_ERROR_NONE) e = code;
  if (d->verbose) fprintf(stderr, "Embree-MI355X: error %d: %s\n", (int)code, msg);
  if (d->errfn) d->errfn(d->errptr, code, msg);
}

// ---------------------------------------------------------------- buffer --

struct Buffer : RefObj {
  Device* dev;
  char* data;
  size_t size;
  bool shared;
  Buffer(Device* d, void* p, size_t n, bool sh) : dev(d), data((char*)p), size(n), shared(sh) { ref_inc(d); }
  ~Buffer() override {
    if (!shared) free(data);
    ref_dec(dev);
  }
};

// -------------------------------------------------------------- geometry --

struct Slot {
  Buffer* buf = nullptr;  // retained when set through a buffer object
  const char* ptr = nullptr;
  size_t stride = 0, count = 0;
  RTCFormat format = RTC_FORMAT_UNDEFINED;
  bool set = false;
  void clear() {
    if (buf) ref_dec(buf);
    buf = nullptr;
    ptr = nullptr;
    set = false;
  }
};

struct Geometry : RefObj {
  Device* dev;
  RTCGeometryType type;
  Slot vertex, index;
  bool enabled = true;
  bool committed = false;
  unsigned mask = 0xffffffffu;
  RTCBuildQuality quality = RTC_BUILD_QUALITY_MEDIUM;
  void* user = nullptr;
  // RTC_GEOMETRY_TYPE_INSTANCE: the instanced scene (a Scene, retained) and
  // its local-to-world transform, columns vx | vy | vz | p
  RefObj* inst_scene = nullptr;
  float l2w[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
  Geometry(Device* d, RTCGeometryType t) : dev(d), type(t) { ref_inc(d); }
  ~Geometry() override {
    vertex.clear();
    index.clear();
    if (inst_scene) ref_dec(inst_scene);
    ref_dec(dev);
  }
};

// ------------------------------------------------------------ user BVH --

// Memory handed to the caller's node and leaf callbacks: bump allocation in
// blocks owned by the RTCBVH, released when it is rebuilt or released.
struct Arena {
  std::vector<std::unique_ptr<char[]>> blocks;
  size_t used = 0, cap = 0;
  void reset() {
    blocks.clear();
    used = cap = 0;
  }
  void* alloc(size_t bytes, size_t align) {
    if (align == 0 || (align & (align - 1))) align = 16;
    size_t off = (used + align - 1) & ~(align - 1);
    if (blocks.empty() || off + bytes > cap) {
      const size_t want = std::max<size_t>(bytes + align, 1 << 20);
      blocks.emplace_back(new char[want]);
      cap = want;
      used = 0;
      const uintptr_t base = (uintptr_t)blocks.back().get();
      off = (size_t)(((base + align - 1) & ~(uintptr_t)(align - 1)) - base);
    }
    char* p = blocks.back().get() + off;
    used = off + bytes;
    memset(p, 0, bytes);
    return p;
  }
};

struct BVHUser : RefObj {
  Device* dev;
  Arena arena;
  std::mutex mtx;
  explicit BVHUser(Device* d) : dev(d) { ref_inc(d); }
  ~BVHUser() override { ref_dec(dev); }
};

// ----------------------------------------------------------- split model --

struct SplitModelH : RefObj {
  Device* dev;
  rtg::SplitModel m;
  uint32_t min_count = 64;
  float sah_slack = 0.05f;
  std::mutex mtx;
  explicit SplitModelH(Device* d) : dev(d) { ref_inc(d); }
  ~SplitModelH() override { ref_dec(dev); }
};

// ----------------------------------------------------------------- scene --

struct Scene : RefObj {
  Device* dev;
  std::vector<Geometry*> geoms;
  RTCSceneFlags flags = RTC_SCENE_FLAG_NONE;
  RTCBuildQuality quality = RTC_BUILD_QUALITY_MEDIUM;
  RTCProgressMonitorFunction progress = nullptr;
  void* progress_ptr = nullptr;
  bool committed = false;
  std::mutex mtx;
  std::shared_ptr<rtg::DeviceBVH> bvh = std::make_shared<rtg::DeviceBVH>();  // shared by the scenes of a batch
  uint32_t slot = 0;                                                          // this scene's slot in bvh
  std::vector<rtg::DevBuf*> gdata;
  std::vector<rtg::GeomDesc> descs;
  SplitModelH* model = nullptr;  // learned split cost for the next commits (retained)
  // Two-level scenes (at least one instance at the last commit): the tree
  // over instance records, the object table it indexes, and the object
  // trees it was built against with their build generations. Holding the
  // trees keeps them alive: a later commit of an instanced scene builds into
  // new arrays, so this scene keeps tracing the trees of its own commit; a
  // rebuild in place (generation change) is refused at query time.
  bool two_level = false;
  rtg::DeviceBVH top;
  rtg::DevBuf top_objects;
  rtg::DevTop top_view{};
  std::vector<std::pair<std::shared_ptr<rtg::DeviceBVH>, uint64_t>> top_deps;
  std::vector<RefObj*> top_scenes;  // instanced scenes, retained while the top tree refers to them
  std::vector<rtg::InstRec> top_recs;  // the top tree's records and their world boxes (lo xyz, hi xyz), as built
  std::vector<float> top_boxes;
  uint32_t top_obj_stack = 0;          // deepest object tree's traversal stack
  std::vector<rtg::DevScene> top_object_views;  // the object table as uploaded
  std::vector<Scene*> top_object_scenes;        // the scene of each object table entry
  float bounds_lo[3] = {INFINITY, INFINITY, INFINITY};
  float bounds_hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  explicit Scene(Device* d) : dev(d) { ref_inc(d); }
  ~Scene() override {
    for (RefObj* o : top_scenes) ref_dec(o);
    if (model) ref_dec(model);
    for (Geometry* g : geoms)
      if (g) ref_dec(g);
    for (auto* b : gdata) delete b;
    ref_dec(dev);
  }
};

Device* dev_of(Device* d) { return d; }
Device* dev_of(Buffer* b) { return b ? b->dev : nullptr; }
Device* dev_of(Geometry* g) { return g ? g->dev : nullptr; }
Device* dev_of(Scene* s) { return s ? s->dev : nullptr; }
Device* dev_of(SplitModelH* m) { return m ? m->dev : nullptr; }
Device* dev_of(BVHUser* b) { return b ? b->dev : nullptr; }

#define API_BEGIN try {
#define API_END(obj)                                                                                 \
  }                                                                                                  \
  catch (const rtc_error& e) {                                                                       \
    process_error(dev_of(obj), e.code, e.msg.c_str());                                               \
  }                                                                                                  \
  catch (const rtg::hip_error& e) {                                                                  \
    process_error(dev_of(obj), e.code == hipErrorOutOfMemory ? RTC_ERROR_OUT_OF_MEMORY : RTC_ERROR_UNKNOWN, \
                  e.what());                                                                         \
  }                                                                                                  \
  catch (const std::bad_alloc&) {                                                                    \
    process_error(dev_of(obj), RTC_ERROR_OUT_OF_MEMORY, "out of memory");                            \
  }                                                                                                  \
  catch (const std::exception& e) {                                                                  \
    process_error(dev_of(obj), RTC_ERROR_UNKNOWN, e.what());                                         \
  }

#define VERIFY_HANDLE(h) \
  if ((h) == nullptr) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid argument");

std::mutex g_api_mutex;  // device creation / retain / release (reference rtcore.cpp:20-38)

void parse_config(Device* d, const char* cfg) {
  if (!cfg) return;
  std::string s(cfg);
  size_t pos = 0;
  while (pos < s.size()) {
    size_t end = s.find(',', pos);
    if (end == std::string::npos) end = s.size();
    std::string kv = s.substr(pos, end - pos);
    pos = end + 1;
    size_t eq = kv.find('=');
    std::string k = kv.substr(0, eq), v = eq == std::string::npos ? "" : kv.substr(eq + 1);
    while (!k.empty() && k.front() == ' ') k.erase(0, 1);
    while (!k.empty() && k.back() == ' ') k.pop_back();
    if (k == "verbose") d->verbose = atoi(v.c_str());
    else if (k == "gpu") d->ordinal = atoi(v.c_str());
    else if (k == "bvh_leaf" || k == "max_leaf") d->build.max_leaf = (uint32_t)atoi(v.c_str());
    else if (k == "cost_node") d->build.cost_node = (float)atof(v.c_str());
    else if (k == "cost_tri") d->build.cost_tri = (float)atof(v.c_str());
    else if (k == "sah_block") d->build.block = (uint32_t)std::max(1, atoi(v.c_str()));
    else if (k == "spatial_budget") d->spatial_budget = std::max(0.0f, (float)atof(v.c_str()));
    else if (k == "builder" || k == "tri_builder") {
      // reference state.cpp:309-310 (aliases) and bvh8_factory.cpp:871-878
      // (names): sah and the two-level "dynamic" SAH build the binned-SAH
      // tree here, morton (and lbvh) the Morton radix tree
      d->tri_builder_name = v;
      if (v == "default") d->tri_builder = Device::TB_DEFAULT;
      else if (v == "sah" || v == "dynamic") d->tri_builder = Device::TB_SAH;
      else if (v == "morton" || v == "lbvh") d->tri_builder = Device::TB_MORTON;
      else d->tri_builder = Device::TB_UNKNOWN;
    }
    else if (k == "bvh_width") d->build.width = (uint32_t)atoi(v.c_str());
    else if (k == "sah_median_below") d->build.median_below = (uint32_t)atoi(v.c_str());
    else if (k == "trace_stack") d->trace.lds_slots = atoi(v.c_str());
    else if (k == "ray_sort") d->trace.sort_rays = atoi(v.c_str());
    else if (k == "ray_sort_min") d->trace.sort_min = (uint32_t)atol(v.c_str());
    else if (k == "persistent") d->trace.persistent = atoi(v.c_str());
    else if (k == "refill_min") d->trace.refill_min = atoi(v.c_str());
    else if (k == "refill_min_occ") d->trace.refill_min_occ = atoi(v.c_str());
    else if (k == "refill_min_occ_inst") d->trace.refill_min_occ_inst = atoi(v.c_str());
    else if (k == "blocks_per_cu") d->trace.blocks_per_cu = atoi(v.c_str());
    else if (k == "dual") d->trace.dual = atoi(v.c_str());
    // threads, isa, tri_accel, set_affinity, ... are CPU settings: accepted, ignored
  }
}

// Build settings for a scene. With the default builder the quality chooses
// (reference scene.cpp:93-116, bvh8_factory.cpp:865-870): LOW the Morton
// radix tree, MEDIUM and HIGH the binned-SAH tree; a builder named in the
// device config applies at every quality; an unknown name is an error here,
// at commit, as in the reference's factory (bvh8_factory.cpp:878).
rtg::BuildParams build_params(const Device* d, RTCBuildQuality q, RTCSceneFlags flags) {
  rtg::BuildParams p = d->build_settings();
  p.quality = q;
  p.robust = (flags & RTC_SCENE_FLAG_ROBUST) ? 1 : 0;
  p.split_budget = q == RTC_BUILD_QUALITY_HIGH ? d->spatial_budget : 0.0f;
  switch (d->tri_builder) {
    case Device::TB_DEFAULT: p.builder = q == RTC_BUILD_QUALITY_LOW ? rtg::BUILDER_MORTON : rtg::BUILDER_SAH; break;
    case Device::TB_SAH: p.builder = rtg::BUILDER_SAH; break;
    case Device::TB_MORTON: p.builder = rtg::BUILDER_MORTON; break;
    default:
      RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "unknown builder " + d->tri_builder_name + " for BVH8<Triangle4>");
  }
  return p;
}

Scene* SC(RTCScene s) { return (Scene*)s; }
Geometry* GE(RTCGeometry g) { return (Geometry*)g; }
Device* DE(RTCDevice d) { return (Device*)d; }
Buffer* BU(RTCBuffer b) { return (Buffer*)b; }

size_t format_size(RTCFormat f) {
  switch (f) {
    case RTC_FORMAT_FLOAT3: return 12;
    case RTC_FORMAT_UINT3: return 12;
    case RTC_FORMAT_FLOAT: case RTC_FORMAT_UINT: case RTC_FORMAT_INT: return 4;
    case RTC_FORMAT_FLOAT2: case RTC_FORMAT_UINT2: case RTC_FORMAT_INT2: return 8;
    case RTC_FORMAT_FLOAT4: case RTC_FORMAT_UINT4: case RTC_FORMAT_INT4: return 16;
    default: return 0;
  }
}

// Validates and binds a triangle-mesh buffer (reference scene_triangle_mesh.cpp:35-80).
void bind_slot(Geometry* g, RTCBufferType type, unsigned slot, RTCFormat format, Buffer* buf, const void* ptr,
               size_t offset, size_t stride, size_t count) {
  if (g->type != RTC_GEOMETRY_TYPE_TRIANGLE) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "unsupported geometry type");
  if (count > 0xffffffffull) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "buffer too large");
  const char* p = (const char*)ptr + offset;
  if (((size_t)p & 3) || (stride & 3)) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "data must be 4 bytes aligned");
  Slot* s = nullptr;
  if (type == RTC_BUFFER_TYPE_INDEX) {
    if (slot != 0) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid index buffer slot");
    if (format != RTC_FORMAT_UINT3) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "invalid index buffer format");
    s = &g->index;
  } else if (type == RTC_BUFFER_TYPE_VERTEX) {
    if (slot != 0) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid vertex buffer slot");
    if (format != RTC_FORMAT_FLOAT3) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "invalid vertex buffer format");
    s = &g->vertex;
  } else if (type == RTC_BUFFER_TYPE_VERTEX_ATTRIBUTE) {
    return;  // accepted; interpolation is outside the ray-query path
  } else {
    RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "unknown buffer type");
  }
  if (stride < format_size(format)) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "stride smaller than element");
  if (buf) ref_inc(buf);
  s->clear();
  s->buf = buf;
  s->ptr = p;
  s->stride = stride;
  s->count = count;
  s->format = format;
  s->set = true;
  g->committed = false;
}

// Build settings of a scene, including its split model (lc receives the
// model's device view and must outlive the build).
rtg::BuildParams scene_params(Scene* sc, rtg::LearnedCost& lc, const rtg::BuildParams* shape = nullptr) {
  rtg::BuildParams p = build_params(sc->dev, sc->quality, sc->flags);
  if (shape) {  // a tree shape under trial (autotuner)
    p.width = shape->width;
    p.max_leaf = shape->max_leaf;
    p.cost_tri = shape->cost_tri;
  }
  if (sc->model) {
    std::lock_guard<std::mutex> ml(sc->model->mtx);
    if (!sc->model->m.has_norm) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "split model has no training data");
    lc.P = sc->model->m.packed.as<rtg::MlpPacked>();
    lc.sn = sc->model->m.norm;
    lc.min_count = sc->model->min_count;
    lc.sah_slack = sc->model->sah_slack;
    p.learned = &lc;
    p.builder = rtg::BUILDER_SAH;
  }
  return p;
}

// Builds the scene's own tree into sc->bvh. Where spatial splits cut
// triangles (HIGH quality), the plain binned-SAH tree over the same
// triangles is built as well and the tree with the lower SAH statistic is
// kept: the splits are chosen by a per-triangle estimate before the tree
// exists, and this keeps them from making a tree worse than it would be
// without them (the reference's spatial-split builder compares both kinds
// of split at every node, kernels/builders/heuristic_spatial_array.h).
void build_scene_tree(Scene* sc, Device* d, const std::vector<rtg::GeomDesc>& descs, const rtg::BuildParams& p,
                      hipStream_t st) {
  rtg::build_bvh(descs, p, d->scratch, *sc->bvh, st);
  if (sc->bvh->stats.num_split == 0) return;
  rtg::BuildParams q = p;
  q.split_budget = 0.0f;
  rtg::DeviceBVH plain;
  rtg::build_bvh(descs, q, d->scratch, plain, st);
  RTG_HIP(hipStreamSynchronize(st));
  if (plain.stats.sah <= sc->bvh->stats.sah) sc->bvh->swap_tree(plain);
}

// Uploads the enabled triangle geometries of sc (scene index `scene` of a
// build) into device buffers owned by sc, appending their descriptors;
// prim_offset advances over their triangles.
void upload_geometry(Scene* sc, Device* d, ThreadDev& t, uint32_t scene, uint32_t& prim_offset,
                     std::vector<rtg::GeomDesc>& out) {
  for (auto* b : sc->gdata) delete b;
  sc->gdata.clear();
  sc->descs.clear();
  for (size_t id = 0; id < sc->geoms.size(); id++) {
    Geometry* g = sc->geoms[id];
    if (!g || !g->enabled || g->type != RTC_GEOMETRY_TYPE_TRIANGLE) continue;
    if (!g->vertex.set || !g->index.set || g->index.count == 0) continue;
    const size_t vbytes = g->vertex.count ? (g->vertex.count - 1) * g->vertex.stride + 12 : 0;
    const size_t ibytes = (g->index.count - 1) * g->index.stride + 12;
    if ((uint64_t)prim_offset + g->index.count > 0xffffffffull) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "too many primitives");
    d->monitor((ssize_t)(vbytes + ibytes));
    auto* vb = new rtg::DevBuf();
    sc->gdata.push_back(vb);
    auto* ib = new rtg::DevBuf();
    sc->gdata.push_back(ib);
    vb->reserve(vbytes ? vbytes : 16);
    ib->reserve(ibytes);
    if (vbytes) RTG_HIP(hipMemcpyAsync(vb->ptr, g->vertex.ptr, vbytes, hipMemcpyHostToDevice, t.stream));
    RTG_HIP(hipMemcpyAsync(ib->ptr, g->index.ptr, ibytes, hipMemcpyHostToDevice, t.stream));
    rtg::GeomDesc gd;
    gd.vertices = (const char*)vb->ptr;
    gd.indices = (const char*)ib->ptr;
    gd.vstride = g->vertex.stride;
    gd.istride = g->index.stride;
    gd.num_vertices = (uint32_t)g->vertex.count;
    gd.num_tris = (uint32_t)g->index.count;
    gd.prim_offset = prim_offset;
    gd.geomID = (uint32_t)id;
    gd.scene = scene;
    gd.kind = rtg::GEOM_TRIANGLES;
    prim_offset += gd.num_tris;
    sc->descs.push_back(gd);
    out.push_back(gd);
  }
}

// ------------------------------------------------------------ instances --

// world2local = rcp(local2world) as the reference's Instance::commit forms it
// (kernels/common/scene_instance.cpp, common/math/affinespace.h:83): the
// inverse of the linear part is adjoint / det (common/math/linearspace3.h:44-53)
// with the Vec3fa cross (common/math/vec3fa.h:358-365: fma(a, b, -(c * d)))
// and the SSE4.1 dpps dot (vec3fa.h:349-351: (x0*y0 + x1*y1) + x2*y2), each
// element divided by det; the translation is -(il * p) (linearspace3.h:146).
// Matrices are columns vx | vy | vz | p.
void affine_inverse(const float* l2w, float* w2l) {
  auto pd = [](float a, float b, float c, float d) { return fmaf(a, b, -(c * d)); };
  auto cross = [&](const float* a, const float* b, float* r) {
    r[0] = pd(a[1], b[2], a[2], b[1]);
    r[1] = pd(a[2], b[0], a[0], b[2]);
    r[2] = pd(a[0], b[1], a[1], b[0]);
  };
  const float *vx = l2w, *vy = l2w + 3, *vz = l2w + 6, *p = l2w + 9;
  float c0[3], c1[3], c2[3];
  cross(vy, vz, c0);
  cross(vz, vx, c1);
  cross(vx, vy, c2);
  const float d01 = vx[0] * c0[0] + vx[1] * c0[1];
  const float det = d01 + (vx[2] * c0[2] + 0.0f);
  for (int k = 0; k < 3; k++) {  // column k of the adjoint is (c0[k], c1[k], c2[k])
    w2l[3 * k + 0] = c0[k] / det;
    w2l[3 * k + 1] = c1[k] / det;
    w2l[3 * k + 2] = c2[k] / det;
  }
  for (int k = 0; k < 3; k++) w2l[9 + k] = -fmaf(p[0], w2l[k], fmaf(p[1], w2l[3 + k], p[2] * w2l[6 + k]));
}

// xfmBounds (common/math/affinespace.h:106-118): the box of the 8
// transformed corners, each transformed as xfmPoint.
void xfm_bounds(const float* m, const float* lo, const float* hi, float* olo, float* ohi) {
  for (int k = 0; k < 3; k++) {
    olo[k] = INFINITY;
    ohi[k] = -INFINITY;
  }
  for (int c = 0; c < 8; c++) {
    const float x = (c & 4) ? hi[0] : lo[0], y = (c & 2) ? hi[1] : lo[1], z = (c & 1) ? hi[2] : lo[2];
    for (int k = 0; k < 3; k++) {
      const float w = fmaf(x, m[k], fmaf(y, m[3 + k], fmaf(z, m[6 + k], m[9 + k])));
      olo[k] = fminf(olo[k], w);
      ohi[k] = fmaxf(ohi[k], w);
    }
  }
}

// The box an instance occupies in the top tree: the transformed corners in
// double precision, widened by 2^-12 of (largest coordinate + extent) and
// rounded outwards, so that hits found in object space with the rounded
// inverse transform stay inside it.
void instance_box(const float* m, const float* lo, const float* hi, float* olo, float* ohi) {
  double dl[3] = {INFINITY, INFINITY, INFINITY}, dh[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int c = 0; c < 8; c++) {
    const double x = (c & 4) ? hi[0] : lo[0], y = (c & 2) ? hi[1] : lo[1], z = (c & 1) ? hi[2] : lo[2];
    for (int k = 0; k < 3; k++) {
      const double w = x * m[k] + y * m[3 + k] + z * m[6 + k] + m[9 + k];
      dl[k] = std::min(dl[k], w);
      dh[k] = std::max(dh[k], w);
    }
  }
  double big = 0.0, ext = 0.0;
  for (int k = 0; k < 3; k++) {
    big = std::max(big, std::max(std::fabs(dl[k]), std::fabs(dh[k])));
    ext = std::max(ext, dh[k] - dl[k]);
  }
  const double margin = (big + ext) * (1.0 / 4096.0);
  for (int k = 0; k < 3; k++) {
    olo[k] = std::nextafter((float)(dl[k] - margin), -INFINITY);
    ohi[k] = std::nextafter((float)(dh[k] + margin), INFINITY);
  }
}

// Builds the top tree of a scene with instances over its instance records
// (one per enabled instance of a committed, non-empty scene, plus one
// identity record for the scene's own triangles). Leaves hold one record,
// as the reference's instance BVH does (kernels/bvh/bvh_builder_sah.cpp,
// BVH8InstanceSceneBuilderSAH: leaf size 1). Without instances the scene
// stays single-level.
void build_top(Scene* sc, Device* d, ThreadDev& t) {
  for (RefObj* o : sc->top_scenes) ref_dec(o);
  sc->top_scenes.clear();
  sc->top_deps.clear();
  sc->two_level = false;
  const rtg::SceneSlot own = sc->slot < sc->bvh->slots.size() ? sc->bvh->slots[sc->slot] : rtg::SceneSlot();
  for (int k = 0; k < 3; k++) {
    sc->bounds_lo[k] = own.num_prims ? own.bounds_lo[k] : INFINITY;
    sc->bounds_hi[k] = own.num_prims ? own.bounds_hi[k] : -INFINITY;
  }
  std::vector<std::pair<uint32_t, Geometry*>> ginst;
  for (size_t id = 0; id < sc->geoms.size(); id++) {
    Geometry* g = sc->geoms[id];
    if (g && g->enabled && g->type == RTC_GEOMETRY_TYPE_INSTANCE && g->inst_scene) ginst.push_back({(uint32_t)id, g});
  }
  if (ginst.empty()) return;

  std::vector<rtg::DevScene> objects;
  std::vector<Scene*> object_scene;
  std::vector<rtg::InstRec> recs;
  std::vector<float> boxes;
  uint32_t obj_stack = 0;
  uint64_t obj_size = 0;
  auto object_of = [&](Scene* o) -> uint32_t {
    for (size_t k = 0; k < object_scene.size(); k++)
      if (object_scene[k] == o) return (uint32_t)k;
    objects.push_back(o->bvh->view(o->slot));
    object_scene.push_back(o);
    obj_stack = std::max(obj_stack, o->bvh->stats.max_stack);
    obj_size = std::max(obj_size, (uint64_t)o->bvh->stats.num_nodes + o->bvh->stats.num_prims);
    sc->top_deps.push_back({o->bvh, o->bvh->generation});
    return (uint32_t)(objects.size() - 1);
  };
  auto add_record = [&](const float* w2l, uint32_t inst, uint32_t obj, uint32_t flags, const float* lo, const float* hi) {
    rtg::InstRec r{};
    for (int k = 0; k < 12; k++) r.m[k] = w2l[k];
    r.instID = inst;
    r.object = obj;
    r.flags = flags;
    recs.push_back(r);
    for (int k = 0; k < 3; k++) boxes.push_back(lo[k]);
    for (int k = 0; k < 3; k++) boxes.push_back(hi[k]);
  };
  if (own.num_prims) {
    const float id12[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
    add_record(id12, RTC_INVALID_GEOMETRY_ID, object_of(sc), rtg::INST_IDENTITY, own.bounds_lo, own.bounds_hi);
  }
  for (auto& gi : ginst) {
    Scene* o = (Scene*)gi.second->inst_scene;
    if (o == sc) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "a scene cannot instance itself");
    if (!o->committed) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "instanced scene not committed");
    const rtg::SceneSlot os = o->slot < o->bvh->slots.size() ? o->bvh->slots[o->slot] : rtg::SceneSlot();
    if (os.num_prims == 0) continue;  // nothing to hit (instances inside it lie beyond the one instance level)
    float w2l[12], lo[3], hi[3], blo[3], bhi[3];
    affine_inverse(gi.second->l2w, w2l);
    instance_box(gi.second->l2w, os.bounds_lo, os.bounds_hi, lo, hi);
    xfm_bounds(gi.second->l2w, os.bounds_lo, os.bounds_hi, blo, bhi);
    for (int k = 0; k < 3; k++) {
      sc->bounds_lo[k] = std::min(sc->bounds_lo[k], blo[k]);
      sc->bounds_hi[k] = std::max(sc->bounds_hi[k], bhi[k]);
    }
    add_record(w2l, gi.first, object_of(o), 0u, lo, hi);
  }
  if (recs.empty() || (recs.size() == 1 && own.num_prims)) {
    sc->top_deps.clear();
    return;  // no instance to enter: the scene's own tree serves
  }
  const uint32_t n = (uint32_t)recs.size();
  rtg::DevBuf dboxes, drecs;
  dboxes.reserve((size_t)n * 24);
  drecs.reserve((size_t)n * sizeof(rtg::InstRec));
  RTG_HIP(hipMemcpyAsync(dboxes.ptr, boxes.data(), (size_t)n * 24, hipMemcpyHostToDevice, t.stream));
  RTG_HIP(hipMemcpyAsync(drecs.ptr, recs.data(), (size_t)n * sizeof(rtg::InstRec), hipMemcpyHostToDevice, t.stream));
  rtg::GeomDesc gd{};
  gd.vertices = (const char*)dboxes.ptr;
  gd.indices = (const char*)drecs.ptr;
  gd.vstride = 24;
  gd.istride = sizeof(rtg::InstRec);
  gd.num_vertices = n;
  gd.num_tris = n;
  gd.prim_offset = 0;
  gd.geomID = 0;
  gd.scene = 0;
  gd.kind = rtg::GEOM_INSTANCES;
  // the scene's split model, if any, prices the top tree's splits too
  rtg::LearnedCost lc{};
  rtg::BuildParams p = scene_params(sc, lc);
  p.builder = rtg::BUILDER_SAH;
  p.max_leaf = 1;
  p.split_budget = 0.0f;
  d->monitor((ssize_t)n * (80 + 64));
  {
    std::lock_guard<std::mutex> bl(d->build_mtx);
    rtg::build_bvh({gd}, p, d->scratch, sc->top, t.stream);
  }
  sc->top_objects.reserve(objects.size() * sizeof(rtg::DevScene));
  RTG_HIP(hipMemcpyAsync(sc->top_objects.ptr, objects.data(), objects.size() * sizeof(rtg::DevScene),
                         hipMemcpyHostToDevice, t.stream));
  RTG_HIP(hipStreamSynchronize(t.stream));  // the uploads read host vectors and free dboxes/drecs below
  if (sc->top.stats.num_prims != n)
    RTC_THROW(RTC_ERROR_INVALID_OPERATION, "invalid instance transform or bounds");
  if (sc->top.stats.max_stack + obj_stack + 2 > rtg::TRACE_STACK)
    RTC_THROW(RTC_ERROR_UNKNOWN, "two-level tree too deep for the traversal stack");
  rtg::DevTop& v = sc->top_view;
  v.nodes = sc->top.nodes.as<rtg::QNode8>();
  v.insts = sc->top.insts.as<rtg::InstRec>();
  v.objects = sc->top_objects.as<rtg::DevScene>();
  v.num_nodes = sc->top.stats.num_nodes;
  v.num_insts = n;
  v.num_objects = (uint32_t)objects.size();
  v.root = sc->top.slots.empty() ? 0u : sc->top.slots[0].root;
  v.max_iter = (uint64_t)v.num_nodes + n + (uint64_t)n * (obj_size + 2) + 16;
  sc->top_recs = std::move(recs);
  sc->top_boxes = std::move(boxes);
  sc->top_obj_stack = obj_stack;
  sc->top_object_views = objects;
  sc->top_object_scenes = object_scene;
  for (Scene* o : object_scene)
    if (o != sc) {
      ref_inc(o);
      sc->top_scenes.push_back(o);
    }
  sc->two_level = true;
}

// Throws when an object tree the top tree refers to was rebuilt since.
void check_top(Scene* sc) {
  for (auto& dp : sc->top_deps)
    if (dp.first->generation != dp.second)
      RTC_THROW(RTC_ERROR_INVALID_OPERATION, "an instanced scene was committed again: commit this scene again");
}

bool scene_empty(Scene* sc) { return sc->two_level ? false : sc->bvh->empty(sc->slot); }

// Traverses the scene over device records on `stream`.
void launch_scene(Scene* sc, ThreadDev& t, void* rays, unsigned M, size_t stride, bool occluded, uint32_t inst,
                  hipStream_t stream, const rtg::TraceParams& tp) {
  if (sc->two_level) {
    check_top(sc);
    rtg::launch_trace_inst(sc->top_view, rays, M, stride, occluded, inst, t.errword.as<uint32_t>(), stream, tp,
                           &t.sorter);
  } else if (!sc->bvh->empty(sc->slot)) {
    rtg::launch_trace(sc->bvh->view(sc->slot), rays, M, stride, occluded, inst, nullptr, nullptr,
                      t.errword.as<uint32_t>(), stream, tp, &t.sorter);
  }
}

void single_level_only(Scene* sc, const char* what) {
  if (sc->two_level) RTC_THROW(RTC_ERROR_INVALID_OPERATION, std::string(what) + " is not supported for scenes with instances");
}

// Uploads geometry buffers and builds the device BVH.
void commit_scene(Scene* sc) {
  Device* d = sc->dev;
  d->activate();
  ThreadDev& t = d->td();
  std::lock_guard<std::mutex> lk(sc->mtx);
  if (sc->progress && !sc->progress(sc->progress_ptr, 0.0)) RTC_THROW(RTC_ERROR_CANCELLED, "progress monitor forced termination");
  uint32_t prim_offset = 0;
  std::vector<rtg::GeomDesc> descs;
  upload_geometry(sc, d, t, 0, prim_offset, descs);
  {
    std::lock_guard<std::mutex> bl(d->build_mtx);
    rtg::LearnedCost lc{};
    const rtg::BuildParams p = scene_params(sc, lc);
    d->monitor((ssize_t)prim_offset * (80 + 48));
    if (sc->bvh.use_count() > 1 || sc->slot != 0) {  // leave a batch's shared arrays to the other scenes
      sc->bvh = std::make_shared<rtg::DeviceBVH>();
      sc->slot = 0;
    }
    build_scene_tree(sc, d, descs, p, t.stream);
  }
  RTG_HIP(hipStreamSynchronize(t.stream));
  build_top(sc, d, t);
  sc->committed = true;
  if (sc->progress && !sc->progress(sc->progress_ptr, 1.0)) RTC_THROW(RTC_ERROR_CANCELLED, "progress monitor forced termination");
}

bool is_device_ptr(Device* d, const void* p) {
  hipPointerAttribute_t attr;
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice && attr.device == d->ordinal;
}

void check_trace_error(ThreadDev& t) {
  uint32_t err = 0;
  RTG_HIP(hipMemcpyAsync(&err, t.errword.ptr, 4, hipMemcpyDeviceToHost, t.stream));
  RTG_HIP(hipStreamSynchronize(t.stream));
  if (err) {
    RTG_HIP(hipMemsetAsync(t.errword.ptr, 0, 4, t.stream));
    RTG_HIP(hipStreamSynchronize(t.stream));
    RTC_THROW(RTC_ERROR_UNKNOWN, (err & 1u) ? "traversal stack overflow" : "corrupt acceleration structure");
  }
}

// Stream query over host or device records (reference rtcore.cpp:595-623, 848-874).
// Host records per chunk of a query in host memory (80 MB of RTCRayHit).
constexpr size_t HOST_CHUNK_RAYS = size_t(1) << 20;

// Copies n records of rec bytes from src (stride sstride) to dst (stride
// dstride), on several host threads when there are many.
void copy_records(void* dst, size_t dstride, const void* src, size_t sstride, size_t n, size_t rec) {
  auto part = [&](size_t a, size_t b) {
    if (dstride == rec && sstride == rec) {
      memcpy((char*)dst + a * rec, (const char*)src + a * rec, (b - a) * rec);
    } else {
      for (size_t i = a; i < b; i++) memcpy((char*)dst + i * dstride, (const char*)src + i * sstride, rec);
    }
  };
  const unsigned hw = std::thread::hardware_concurrency();
  const size_t nt = std::min<size_t>({8, hw > 1 ? hw / 2 : 1, std::max<size_t>(1, n * rec / (4u << 20))});
  if (nt <= 1) {
    part(0, n);
    return;
  }
  std::vector<std::thread> th;
  const size_t per = (n + nt - 1) / nt;
  size_t j = 1;
  try {
    for (; j < nt; j++) {
      const size_t a = std::min(n, j * per), b = std::min(n, (j + 1) * per);
      if (a < b) th.emplace_back(part, a, b);
    }
  } catch (...) {  // no more threads: this one copies the parts not started
  }
  part(0, std::min(n, per));
  for (; j < nt; j++) part(std::min(n, j * per), std::min(n, (j + 1) * per));
  for (auto& x : th) x.join();
}

void trace_stream(Scene* sc, const RTCIntersectContext* ctx, void* rays, unsigned M, size_t stride, bool occluded) {
  if (!sc->committed) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "scene not committed");
  if (ctx && ctx->filter) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "filter functions are not supported");
  if (M == 0) return;
  if (!rays) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid argument");
  const size_t rec = occluded ? sizeof(RTCRay) : sizeof(RTCRayHit);
  if (stride < rec && M > 1) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "stride smaller than the ray record");
  Device* d = sc->dev;
  d->activate();
  ThreadDev& t = d->td();
  const uint32_t inst = ctx ? ctx->instID[0] : RTC_INVALID_GEOMETRY_ID;
  if (scene_empty(sc)) return;  // empty scene: every ray misses
  if (is_device_ptr(d, rays)) {
    launch_scene(sc, t, rays, M, stride, occluded, inst, t.stream, d->trace_params());
    check_trace_error(t);
    return;
  }
  // Host records move in chunks through two pinned buffers: host threads
  // pack chunk k into one while the GPU uploads, traces and downloads chunk
  // k - 1 through the other, and unpack a chunk once its download is done.
  // (Pageable memory cannot be registered for DMA here, so every byte goes
  // through a host copy; several threads copy faster than one.)
  const size_t chunk = std::min<size_t>(M, HOST_CHUNK_RAYS);
  const size_t nchunks = (M + chunk - 1) / chunk;
  t.staging.reserve((size_t)M * rec);
  for (auto& b : t.stage) b.reserve(chunk * rec);
  const rtg::TraceParams tp = d->trace_params();
  auto unpack = [&](size_t k) {
    const size_t first = k * chunk, n = std::min<size_t>(chunk, M - first);
    RTG_HIP(hipEventSynchronize(t.ev_stage[k & 1]));
    copy_records((char*)rays + first * stride, stride, t.stage[k & 1].ptr, rec, n, rec);
  };
  for (size_t k = 0; k < nchunks; k++) {
    if (k >= 2) unpack(k - 2);  // frees this chunk's pinned buffer
    const size_t first = k * chunk, n = std::min<size_t>(chunk, M - first);
    char* dev = t.staging.as<char>() + first * rec;
    copy_records(t.stage[k & 1].ptr, rec, (const char*)rays + first * stride, stride, n, rec);
    RTG_HIP(hipMemcpyAsync(dev, t.stage[k & 1].ptr, n * rec, hipMemcpyHostToDevice, t.stream));
    launch_scene(sc, t, dev, (unsigned)n, rec, occluded, inst, t.stream, tp);
    RTG_HIP(hipMemcpyAsync(t.stage[k & 1].ptr, dev, n * rec, hipMemcpyDeviceToHost, t.stream));
    RTG_HIP(hipEventRecord(t.ev_stage[k & 1], t.stream));
  }
  for (size_t k = nchunks >= 2 ? nchunks - 2 : 0; k < nchunks; k++) unpack(k);
  check_trace_error(t);
}

}  // namespace

extern "C" {

// ------------------------------------------------------------------ device --

RTC_API RTCDevice rtcNewDevice(const char* config) {
  Device* d = nullptr;
  API_BEGIN
  std::lock_guard<std::mutex> g(g_api_mutex);
  d = new Device();
  parse_config(d, config);
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= d->ordinal) {
    (void)hipGetLastError();
    delete d;
    d = nullptr;
    RTC_THROW(RTC_ERROR_UNSUPPORTED_CPU, "no usable GPU");
  }
  d->activate();
  rtg::noalloc_setup(d->ordinal);
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, d->ordinal) == hipSuccess && prop.multiProcessorCount > 0)
      d->trace.cus = prop.multiProcessorCount;
  }
  if (d->verbose) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, d->ordinal) == hipSuccess)
      fprintf(stderr, "Embree-MI355X %s on %s (%s), %d CUs\n", RTC_VERSION_STRING, prop.name, prop.gcnArchName,
              prop.multiProcessorCount);
  }
  return (RTCDevice)d;
  API_END((Device*)nullptr)
  return nullptr;
}

RTC_API void rtcRetainDevice(RTCDevice device) {
  API_BEGIN
  VERIFY_HANDLE(device);
  std::lock_guard<std::mutex> g(g_api_mutex);
  ref_inc(DE(device));
  API_END((Device*)nullptr)
}

RTC_API void rtcReleaseDevice(RTCDevice device) {
  API_BEGIN
  VERIFY_HANDLE(device);
  std::lock_guard<std::mutex> g(g_api_mutex);
  ref_dec(DE(device));
  API_END((Device*)nullptr)
}

RTC_API ssize_t rtcGetDeviceProperty(RTCDevice device, RTCDeviceProperty prop) {
  Device* d = DE(device);
  API_BEGIN
  VERIFY_HANDLE(device);
  switch (prop) {
    case RTC_DEVICE_PROPERTY_VERSION: return RTC_VERSION;
    case RTC_DEVICE_PROPERTY_VERSION_MAJOR: return RTC_VERSION_MAJOR;
    case RTC_DEVICE_PROPERTY_VERSION_MINOR: return RTC_VERSION_MINOR;
    case RTC_DEVICE_PROPERTY_VERSION_PATCH: return RTC_VERSION_PATCH;
    case RTC_DEVICE_PROPERTY_NATIVE_RAY4_SUPPORTED:
    case RTC_DEVICE_PROPERTY_NATIVE_RAY8_SUPPORTED:
    case RTC_DEVICE_PROPERTY_NATIVE_RAY16_SUPPORTED: return 0;
    case RTC_DEVICE_PROPERTY_RAY_STREAM_SUPPORTED: return 1;
    case RTC_DEVICE_PROPERTY_BACKFACE_CULLING_CURVES_ENABLED:
    case RTC_DEVICE_PROPERTY_RAY_MASK_SUPPORTED:
    case RTC_DEVICE_PROPERTY_BACKFACE_CULLING_ENABLED:
    case RTC_DEVICE_PROPERTY_FILTER_FUNCTION_SUPPORTED:
    case RTC_DEVICE_PROPERTY_IGNORE_INVALID_RAYS_ENABLED:
    case RTC_DEVICE_PROPERTY_COMPACT_POLYS_ENABLED: return 0;
    case RTC_DEVICE_PROPERTY_TRIANGLE_GEOMETRY_SUPPORTED: return 1;
    case RTC_DEVICE_PROPERTY_QUAD_GEOMETRY_SUPPORTED:
    case RTC_DEVICE_PROPERTY_SUBDIVISION_GEOMETRY_SUPPORTED:
    case RTC_DEVICE_PROPERTY_CURVE_GEOMETRY_SUPPORTED:
    case RTC_DEVICE_PROPERTY_USER_GEOMETRY_SUPPORTED:
    case RTC_DEVICE_PROPERTY_POINT_GEOMETRY_SUPPORTED: return 0;
    case RTC_DEVICE_PROPERTY_TASKING_SYSTEM: return 0;
    case RTC_DEVICE_PROPERTY_JOIN_COMMIT_SUPPORTED: return 1;
    case RTC_DEVICE_PROPERTY_PARALLEL_COMMIT_SUPPORTED: return 1;
    default: RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "unknown readable property");
  }
  API_END(d)
  return 0;
}

RTC_API void rtcSetDeviceProperty(RTCDevice device, const RTCDeviceProperty prop, ssize_t value) {
  Device* d = DE(device);
  API_BEGIN
  (void)prop;
  (void)value;
  VERIFY_HANDLE(device);
  RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "unknown writable property");
  API_END(d)
}

RTC_API RTCError rtcGetDeviceError(RTCDevice device) {
  if (!device) {
    RTCError e = tls_null_error;
    tls_null_error = RTC_ERROR_NONE;
    return e;
  }
  auto it = tls_errors.find(DE(device));
  if (it == tls_errors.end()) return RTC_ERROR_NONE;
  RTCError e = it->second;
  it->second = RTC_ERROR_NONE;
  return e;
}

RTC_API void rtcSetDeviceErrorFunction(RTCDevice device, RTCErrorFunction error, void* userPtr) {
  Device* d = DE(device);
  API_BEGIN
  VERIFY_HANDLE(device);
  d->errfn = error;
  d->errptr = userPtr;
  API_END(d)
}

RTC_API void rtcSetDeviceMemoryMonitorFunction(RTCDevice device, RTCMemoryMonitorFunction fn, void* userPtr) {
  Device* d = DE(device);
  API_BEGIN
  VERIFY_HANDLE(device);
  d->memfn = fn;
  d->memptr = userPtr;
  API_END(d)
}

// ------------------------------------------------------------------ buffer --

RTC_API RTCBuffer rtcNewBuffer(RTCDevice device, size_t byteSize) {
  Device* d = DE(device);
  API_BEGIN
  VERIFY_HANDLE(device);
  void* p = malloc(byteSize + 16);
  if (!p) RTC_THROW(RTC_ERROR_OUT_OF_MEMORY, "out of memory");
  return (RTCBuffer) new Buffer(d, p, byteSize, false);
  API_END(d)
  return nullptr;
}

RTC_API RTCBuffer rtcNewSharedBuffer(RTCDevice device, void* ptr, size_t byteSize) {
  Device* d = DE(device);
  API_BEGIN
  VERIFY_HANDLE(device);
  return (RTCBuffer) new Buffer(d, ptr, byteSize, true);
  API_END(d)
  return nullptr;
}

RTC_API void* rtcGetBufferData(RTCBuffer buffer) {
  Buffer* b = BU(buffer);
  API_BEGIN
  VERIFY_HANDLE(buffer);
  return b->data;
  API_END(b)
  return nullptr;
}

RTC_API void rtcRetainBuffer(RTCBuffer buffer) {
  Buffer* b = BU(buffer);
  API_BEGIN
  VERIFY_HANDLE(buffer);
  ref_inc(b);
  API_END(b)
}

RTC_API void rtcReleaseBuffer(RTCBuffer buffer) {
  Buffer* b = BU(buffer);
  API_BEGIN
  VERIFY_HANDLE(buffer);
  ref_dec(b);
  API_END((Buffer*)nullptr)
}

// ---------------------------------------------------------------- geometry --

RTC_API RTCGeometry rtcNewGeometry(RTCDevice device, RTCGeometryType type) {
  Device* d = DE(device);
  API_BEGIN
  VERIFY_HANDLE(device);
  if (type != RTC_GEOMETRY_TYPE_TRIANGLE && type != RTC_GEOMETRY_TYPE_INSTANCE)
    RTC_THROW(RTC_ERROR_INVALID_OPERATION, "geometry type not supported");
  return (RTCGeometry) new Geometry(d, type);
  API_END(d)
  return nullptr;
}

RTC_API void rtcRetainGeometry(RTCGeometry geometry) {
  Geometry* g = GE(geometry);
  API_BEGIN
  VERIFY_HANDLE(geometry);
  ref_inc(g);
  API_END(g)
}

RTC_API void rtcReleaseGeometry(RTCGeometry geometry) {
  Geometry* g = GE(geometry);
  API_BEGIN
  VERIFY_HANDLE(geometry);
  ref_dec(g);
  API_END((Geometry*)nullptr)
}

RTC_API void rtcCommitGeometry(RTCGeometry geometry) {
  Geometry* g = GE(geometry);
  API_BEGIN
  VERIFY_HANDLE(geometry);
  g->committed = true;
  API_END(g)
}

RTC_API void rtcEnableGeometry(RTCGeometry geometry) {
  Geometry* g = GE(geometry);
  API_BEGIN
  VERIFY_HANDLE(geometry);
  g->enabled = true;
  API_END(g)
}

RTC_API void rtcDisableGeometry(RTCGeometry geometry) {
  Geometry* g = GE(geometry);
  API_BEGIN
  VERIFY_HANDLE(geometry);
  g->enabled = false;
  API_END(g)
}

RTC_API void rtcSetGeometryTimeStepCount(RTCGeometry geometry, unsigned int timeStepCount) {
  Geometry* g = GE(geometry);
  API_BEGIN
  VERIFY_HANDLE(geometry);
  if (timeStepCount != 1) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "motion blur is not supported");
  API_END(g)
}

RTC_API void rtcSetGeometryMask(RTCGeometry geometry, unsigned int mask) {
  Geometry* g = GE(geometry);
  API_BEGIN
  VERIFY_HANDLE(geometry);
  g->mask = mask;
  API_END(g)
}

RTC_API void rtcSetGeometryBuildQuality(RTCGeometry geometry, RTCBuildQuality quality) {
  Geometry* g = GE(geometry);
  API_BEGIN
  VERIFY_HANDLE(geometry);
  if (quality > RTC_BUILD_QUALITY_REFIT) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid build quality");
  g->quality = quality;
  API_END(g)
}

RTC_API void rtcSetGeometryBuffer(RTCGeometry geometry, RTCBufferType type, unsigned int slot, RTCFormat format,
                                  RTCBuffer buffer, size_t byteOffset, size_t byteStride, size_t itemCount) {
  Geometry* g = GE(geometry);
  API_BEGIN
  VERIFY_HANDLE(geometry);
  VERIFY_HANDLE(buffer);
  Buffer* b = BU(buffer);
  if (b->dev != g->dev) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "inputs are from different devices");
  if (itemCount > 0 && byteOffset + (itemCount - 1) * byteStride + format_size(format) > b->size)
    RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "buffer region out of range");
  bind_slot(g, type, slot, format, b, b->data, byteOffset, byteStride, itemCount);
  API_END(g)
}

RTC_API void rtcSetSharedGeometryBuffer(RTCGeometry geometry, RTCBufferType type, unsigned int slot, RTCFormat format,
                                        const void* ptr, size_t byteOffset, size_t byteStride, size_t itemCount) {
  Geometry* g = GE(geometry);
  API_BEGIN
  VERIFY_HANDLE(geometry);
  bind_slot(g, type, slot, format, nullptr, ptr, byteOffset, byteStride, itemCount);
  API_END(g)
}

RTC_API void* rtcSetNewGeometryBuffer(RTCGeometry geometry, RTCBufferType type, unsigned int slot, RTCFormat format,
                                      size_t byteStride, size_t itemCount) {
  Geometry* g = GE(geometry);
  API_BEGIN
  VERIFY_HANDLE(geometry);
  size_t bytes = byteStride * itemCount;
  if (type == RTC_BUFFER_TYPE_VERTEX || type == RTC_BUFFER_TYPE_VERTEX_ATTRIBUTE) bytes += 16;
  void* p = calloc(1, bytes + 16);
  if (!p) RTC_THROW(RTC_ERROR_OUT_OF_MEMORY, "out of memory");
  Buffer* b = new Buffer(g->dev, p, bytes, false);
  try {
    bind_slot(g, type, slot, format, b, p, 0, byteStride, itemCount);
  } catch (...) {
    ref_dec(b);
    throw;
  }
  ref_dec(b);  // the slot holds the only reference
  return p;
  API_END(g)
  return nullptr;
}

RTC_API void* rtcGetGeometryBufferData(RTCGeometry geometry, RTCBufferType type, unsigned int slot) {
  Geometry* g = GE(geometry);
  API_BEGIN
  VERIFY_HANDLE(geometry);
  if (slot != 0) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid buffer slot");
  const Slot* s = type == RTC_BUFFER_TYPE_INDEX ? &g->index : type == RTC_BUFFER_TYPE_VERTEX ? &g->vertex : nullptr;
  if (!s || !s->set) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "buffer not set");
  return (void*)s->ptr;
  API_END(g)
  return nullptr;
}

RTC_API void rtcUpdateGeometryBuffer(RTCGeometry geometry, RTCBufferType type, unsigned int slot) {
  Geometry* g = GE(geometry);
  API_BEGIN
  VERIFY_HANDLE(geometry);
  (void)type;
  (void)slot;
  g->committed = false;
  API_END(g)
}

RTC_API void rtcSetGeometryIntersectFilterFunction(RTCGeometry geometry, RTCFilterFunctionN filter) {
  Geometry* g = GE(geometry);
  API_BEGIN
  VERIFY_HANDLE(geometry);
  if (filter) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "filter functions are not supported");
  API_END(g)
}

RTC_API void rtcSetGeometryOccludedFilterFunction(RTCGeometry geometry, RTCFilterFunctionN filter) {
  Geometry* g = GE(geometry);
  API_BEGIN
  VERIFY_HANDLE(geometry);
  if (filter) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "filter functions are not supported");
  API_END(g)
}

RTC_API void rtcSetGeometryUserData(RTCGeometry geometry, void* ptr) {
  Geometry* g = GE(geometry);
  API_BEGIN
  VERIFY_HANDLE(geometry);
  g->user = ptr;
  API_END(g)
}

RTC_API void* rtcGetGeometryUserData(RTCGeometry geometry) {
  Geometry* g = GE(geometry);
  API_BEGIN
  VERIFY_HANDLE(geometry);
  return g->user;
  API_END(g)
  return nullptr;
}

// Instance geometry (reference rtcore.cpp:1000-1120, kernels/common/scene_instance.cpp).
RTC_API void rtcSetGeometryInstancedScene(RTCGeometry geometry, RTCScene scene) {
  Geometry* g = GE(geometry);
  API_BEGIN
  VERIFY_HANDLE(geometry);
  VERIFY_HANDLE(scene);
  if (g->type != RTC_GEOMETRY_TYPE_INSTANCE) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "operation not supported for this geometry");
  Scene* s = SC(scene);
  ref_inc(s);
  if (g->inst_scene) ref_dec(g->inst_scene);
  g->inst_scene = s;
  g->committed = false;
  API_END(g)
}

// loadTransform (reference rtcore.cpp:1008-1039): FLOAT3X4 row or column
// major and FLOAT4X4 column major; other formats are invalid.
RTC_API void rtcSetGeometryTransform(RTCGeometry geometry, unsigned int timeStep, RTCFormat format, const void* xfm) {
  Geometry* g = GE(geometry);
  API_BEGIN
  VERIFY_HANDLE(geometry);
  VERIFY_HANDLE(xfm);
  if (g->type != RTC_GEOMETRY_TYPE_INSTANCE) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "operation not supported for this geometry");
  if (timeStep != 0) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "invalid timestep");
  const float* x = (const float*)xfm;
  float m[12];
  switch (format) {
    case RTC_FORMAT_FLOAT3X4_ROW_MAJOR:
      for (int c = 0; c < 4; c++)
        for (int r = 0; r < 3; r++) m[3 * c + r] = x[4 * r + c];
      break;
    case RTC_FORMAT_FLOAT3X4_COLUMN_MAJOR:
      for (int k = 0; k < 12; k++) m[k] = x[k];
      break;
    case RTC_FORMAT_FLOAT4X4_COLUMN_MAJOR:
      for (int c = 0; c < 4; c++)
        for (int r = 0; r < 3; r++) m[3 * c + r] = x[4 * c + r];
      break;
    default:
      RTC_THROW(RTC_ERROR_INVALID_OPERATION, "invalid matrix format");
  }
  memcpy(g->l2w, m, sizeof(m));
  g->committed = false;
  API_END(g)
}

// storeTransform (reference rtcore.cpp:1041-1069); one time step, so `time`
// does not matter.
RTC_API void rtcGetGeometryTransform(RTCGeometry geometry, float time, RTCFormat format, void* xfm) {
  Geometry* g = GE(geometry);
  API_BEGIN
  VERIFY_HANDLE(geometry);
  VERIFY_HANDLE(xfm);
  (void)time;
  if (g->type != RTC_GEOMETRY_TYPE_INSTANCE) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "operation not supported for this geometry");
  float* x = (float*)xfm;
  const float* m = g->l2w;
  switch (format) {
    case RTC_FORMAT_FLOAT3X4_ROW_MAJOR:
      for (int c = 0; c < 4; c++)
        for (int r = 0; r < 3; r++) x[4 * r + c] = m[3 * c + r];
      break;
    case RTC_FORMAT_FLOAT3X4_COLUMN_MAJOR:
      for (int k = 0; k < 12; k++) x[k] = m[k];
      break;
    case RTC_FORMAT_FLOAT4X4_COLUMN_MAJOR:
      for (int c = 0; c < 4; c++) {
        for (int r = 0; r < 3; r++) x[4 * c + r] = m[3 * c + r];
        x[4 * c + 3] = c == 3 ? 1.0f : 0.0f;
      }
      break;
    default:
      RTC_THROW(RTC_ERROR_INVALID_OPERATION, "invalid matrix format");
  }
  API_END(g)
}

// ------------------------------------------------------------------- scene --

RTC_API RTCScene rtcNewScene(RTCDevice device) {
  Device* d = DE(device);
  API_BEGIN
  VERIFY_HANDLE(device);
  return (RTCScene) new Scene(d);
  API_END(d)
  return nullptr;
}

RTC_API RTCDevice rtcGetSceneDevice(RTCScene scene) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  ref_inc(s->dev);
  return (RTCDevice)s->dev;
  API_END(s)
  return nullptr;
}

RTC_API void rtcRetainScene(RTCScene scene) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  ref_inc(s);
  API_END(s)
}

RTC_API void rtcReleaseScene(RTCScene scene) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  ref_dec(s);
  API_END((Scene*)nullptr)
}

RTC_API unsigned int rtcAttachGeometry(RTCScene scene, RTCGeometry geometry) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  VERIFY_HANDLE(geometry);
  Geometry* g = GE(geometry);
  if (g->dev != s->dev) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "inputs are from different devices");
  std::lock_guard<std::mutex> lk(s->mtx);
  size_t id = 0;
  while (id < s->geoms.size() && s->geoms[id]) id++;  // smallest free ID
  if (id == s->geoms.size()) s->geoms.push_back(nullptr);
  ref_inc(g);
  s->geoms[id] = g;
  s->committed = false;
  return (unsigned)id;
  API_END(s)
  return RTC_INVALID_GEOMETRY_ID;
}

RTC_API void rtcAttachGeometryByID(RTCScene scene, RTCGeometry geometry, unsigned int geomID) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  VERIFY_HANDLE(geometry);
  if (geomID == RTC_INVALID_GEOMETRY_ID) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "invalid geometry identifier");
  Geometry* g = GE(geometry);
  if (g->dev != s->dev) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "inputs are from different devices");
  std::lock_guard<std::mutex> lk(s->mtx);
  if (geomID >= s->geoms.size()) s->geoms.resize((size_t)geomID + 1, nullptr);
  if (s->geoms[geomID]) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "geometry identifier already used");
  ref_inc(g);
  s->geoms[geomID] = g;
  s->committed = false;
  API_END(s)
}

RTC_API void rtcDetachGeometry(RTCScene scene, unsigned int geomID) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  std::lock_guard<std::mutex> lk(s->mtx);
  if (geomID >= s->geoms.size() || !s->geoms[geomID]) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid geometry identifier");
  ref_dec(s->geoms[geomID]);
  s->geoms[geomID] = nullptr;
  s->committed = false;
  API_END(s)
}

RTC_API RTCGeometry rtcGetGeometry(RTCScene scene, unsigned int geomID) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  if (geomID >= s->geoms.size() || !s->geoms[geomID]) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid geometry identifier");
  return (RTCGeometry)s->geoms[geomID];
  API_END(s)
  return nullptr;
}

RTC_API void rtcCommitScene(RTCScene scene) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  commit_scene(s);
  API_END(s)
}

RTC_API void rtcJoinCommitScene(RTCScene scene) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  commit_scene(s);
  API_END(s)
}

RTC_API void rtcSetSceneProgressMonitorFunction(RTCScene scene, RTCProgressMonitorFunction progress, void* ptr) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  s->progress = progress;
  s->progress_ptr = ptr;
  API_END(s)
}

RTC_API void rtcSetSceneBuildQuality(RTCScene scene, RTCBuildQuality quality) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  if (quality != RTC_BUILD_QUALITY_LOW && quality != RTC_BUILD_QUALITY_MEDIUM && quality != RTC_BUILD_QUALITY_HIGH)
    RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid build quality");
  s->quality = quality;
  API_END(s)
}

RTC_API void rtcSetSceneFlags(RTCScene scene, RTCSceneFlags flags) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  if (flags & RTC_SCENE_FLAG_CONTEXT_FILTER_FUNCTION)
    RTC_THROW(RTC_ERROR_INVALID_OPERATION, "filter functions are not supported");
  s->flags = flags;
  API_END(s)
}

RTC_API RTCSceneFlags rtcGetSceneFlags(RTCScene scene) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  return s->flags;
  API_END(s)
  return RTC_SCENE_FLAG_NONE;
}

RTC_API void rtcGetSceneBounds(RTCScene scene, RTCBounds* b) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  if (!s->committed) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "scene not committed");
  const rtg::SceneSlot st = s->slot < s->bvh->slots.size() ? s->bvh->slots[s->slot] : rtg::SceneSlot();
  b->lower_x = st.bounds_lo[0]; b->lower_y = st.bounds_lo[1]; b->lower_z = st.bounds_lo[2];
  b->upper_x = st.bounds_hi[0]; b->upper_y = st.bounds_hi[1]; b->upper_z = st.bounds_hi[2];
  b->align0 = b->align1 = 0.0f;
  if (s->two_level) {  // own triangles and transformed instance bounds
    b->lower_x = s->bounds_lo[0]; b->lower_y = s->bounds_lo[1]; b->lower_z = s->bounds_lo[2];
    b->upper_x = s->bounds_hi[0]; b->upper_y = s->bounds_hi[1]; b->upper_z = s->bounds_hi[2];
  } else if (st.num_prims == 0) {
    b->lower_x = b->lower_y = b->lower_z = INFINITY;
    b->upper_x = b->upper_y = b->upper_z = -INFINITY;
  }
  API_END(s)
}

RTC_API void rtcIntersect1(RTCScene scene, RTCIntersectContext* context, RTCRayHit* rayhit) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  VERIFY_HANDLE(rayhit);
  if (!(rayhit->ray.tnear <= rayhit->ray.tfar)) return;
  trace_stream(s, context, rayhit, 1, sizeof(RTCRayHit), false);
  API_END(s)
}

RTC_API void rtcIntersect1M(RTCScene scene, RTCIntersectContext* context, RTCRayHit* rayhit, unsigned int M,
                            size_t byteStride) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  trace_stream(s, context, rayhit, M, byteStride, false);
  API_END(s)
}

RTC_API void rtcOccluded1(RTCScene scene, RTCIntersectContext* context, RTCRay* ray) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  VERIFY_HANDLE(ray);
  if (!(ray->tnear <= ray->tfar)) return;
  trace_stream(s, context, ray, 1, sizeof(RTCRay), true);
  API_END(s)
}

RTC_API void rtcOccluded1M(RTCScene scene, RTCIntersectContext* context, RTCRay* ray, unsigned int M,
                           size_t byteStride) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  trace_stream(s, context, ray, M, byteStride, true);
  API_END(s)
}

// -------------------------------------------------------------- extensions --

RTC_API void* rtcgDeviceMalloc(RTCDevice device, size_t bytes) {
  Device* d = DE(device);
  API_BEGIN
  VERIFY_HANDLE(device);
  d->activate();
  void* p = nullptr;
  RTG_HIP(hipMalloc(&p, bytes ? bytes : 16));
  return p;
  API_END(d)
  return nullptr;
}

RTC_API void rtcgDeviceFree(RTCDevice device, void* ptr) {
  Device* d = DE(device);
  API_BEGIN
  VERIFY_HANDLE(device);
  d->activate();
  if (ptr) RTG_HIP(hipFree(ptr));
  API_END(d)
}

RTC_API void rtcgMemcpy(RTCDevice device, void* dst, const void* src, size_t bytes) {
  Device* d = DE(device);
  API_BEGIN
  VERIFY_HANDLE(device);
  d->activate();
  ThreadDev& t = d->td();
  RTG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, t.stream));
  RTG_HIP(hipStreamSynchronize(t.stream));
  API_END(d)
}

static void async_trace(Scene* s, const RTCIntersectContext* ctx, void* rays, unsigned M, size_t stride,
                        bool occluded, void* stream) {
  if (!s->committed) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "scene not committed");
  if (ctx && ctx->filter) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "filter functions are not supported");
  Device* d = s->dev;
  d->activate();
  ThreadDev& t = d->td();
  const uint32_t inst = ctx ? ctx->instID[0] : RTC_INVALID_GEOMETRY_ID;
  launch_scene(s, t, rays, M, stride, occluded, inst, stream ? (hipStream_t)stream : t.stream, d->trace_params());
}

RTC_API void rtcgIntersect1MAsync(RTCScene scene, const RTCIntersectContext* context, RTCRayHit* d_rayhit,
                                  unsigned int M, size_t byteStride, void* hipStream) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  async_trace(s, context, d_rayhit, M, byteStride, false, hipStream);
  API_END(s)
}

RTC_API void rtcgOccluded1MAsync(RTCScene scene, const RTCIntersectContext* context, RTCRay* d_ray, unsigned int M,
                                 size_t byteStride, void* hipStream) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  async_trace(s, context, d_ray, M, byteStride, true, hipStream);
  API_END(s)
}

// No reference counterpart: rtcgIntersect1MAsync / rtcgOccluded1MAsync for
// callers that cannot allocate. The kernels' ray counter and error bits are
// words 0 and 1 of the GPU's fixed scratch array; nothing here allocates
// device memory or creates a stream.
RTC_API void rtcgQuery1MNoAlloc(RTCScene scene, const RTCIntersectContext* context, void* d_rays, unsigned int M,
                                size_t byteStride, int occluded, void* hipStream) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  if (!s->committed) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "scene not committed");
  if (context && context->filter) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "filter functions are not supported");
  if (M == 0) return;
  if (M > rtg::NOALLOC_MAX_RAYS)
    RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "rtcgQuery1MNoAlloc: at most 2^31 rays per call");
  const size_t rec = occluded ? sizeof(RTCRay) : sizeof(RTCRayHit);
  if (!d_rays) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid argument");
  if (byteStride < rec && M > 1) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "stride smaller than the ray record");
  Device* d = s->dev;
  d->activate();
  if (!is_device_ptr(d, d_rays))
    RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "rtcgQuery1MNoAlloc: the records must be in this device's memory");
  uint32_t* scr = rtg::noalloc_scratch(d->ordinal);
  if (!scr) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "rtcgQuery1MNoAlloc: no scratch on this GPU");
  const hipStream_t st = (hipStream_t)hipStream;  // NULL: the null stream, synchronous
  const uint32_t inst = context ? context->instID[0] : RTC_INVALID_GEOMETRY_ID;
  const rtg::TraceParams tp = d->trace_params();
  if (s->two_level) {
    check_top(s);
    rtg::launch_trace_inst(s->top_view, d_rays, M, byteStride, occluded != 0, inst, scr + 1, st, tp, nullptr, scr);
  } else if (!s->bvh->empty(s->slot)) {
    rtg::launch_trace(s->bvh->view(s->slot), d_rays, M, byteStride, occluded != 0, inst, nullptr, nullptr, scr + 1,
                      st, tp, nullptr, nullptr, scr);
  }
  if (hipStream) return;
  RTG_HIP(hipStreamSynchronize(st));
  uint32_t err = 0;
  RTG_HIP(hipMemcpy(&err, scr + 1, 4, hipMemcpyDeviceToHost));
  if (err) {
    rtg::fill_words(scr + 1, 1, 0u, st);
    RTG_HIP(hipStreamSynchronize(st));
    RTC_THROW(RTC_ERROR_UNKNOWN, (err & 1u) ? "traversal stack overflow" : "corrupt acceleration structure");
  }
  API_END(s)
}

RTC_API const char* rtcgGetErrorString(void) { return tls_last_message.c_str(); }

RTC_API void rtcgSynchronize(RTCDevice device) {
  Device* d = DE(device);
  API_BEGIN
  VERIFY_HANDLE(device);
  d->activate();
  ThreadDev& t = d->td();
  RTG_HIP(hipStreamSynchronize(t.stream));
  check_trace_error(t);
  API_END(d)
}

RTC_API void rtcgTimeTrace(RTCScene scene, void* d_rays, const void* d_src, unsigned int M, size_t byteStride,
                           int occluded, int steps, float* ms_out) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  if (!s->committed) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "scene not committed");
  Device* d = s->dev;
  d->activate();
  ThreadDev& t = d->td();
  for (int i = 0; i < steps; i++) {
    if (d_src) RTG_HIP(hipMemcpyAsync(d_rays, d_src, (size_t)M * byteStride, hipMemcpyDeviceToDevice, t.stream));
    RTG_HIP(hipEventRecord(t.ev0, t.stream));
    launch_scene(s, t, d_rays, M, byteStride, occluded != 0, RTC_INVALID_GEOMETRY_ID, t.stream, d->trace_params());
    RTG_HIP(hipEventRecord(t.ev1, t.stream));
    RTG_HIP(hipEventSynchronize(t.ev1));
    float ms = 0.0f;
    RTG_HIP(hipEventElapsedTime(&ms, t.ev0, t.ev1));
    if (ms_out) ms_out[i] = ms;
  }
  check_trace_error(t);
  API_END(s)
}

RTC_API void rtcgTimeBuild(RTCScene scene, int steps, float* ms_out) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  if (!s->committed) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "scene not committed");
  single_level_only(s, "rebuild timing");
  Device* d = s->dev;
  d->activate();
  ThreadDev& t = d->td();
  std::lock_guard<std::mutex> lk(s->mtx);
  for (int i = 0; i < steps; i++) {
    std::lock_guard<std::mutex> bl(d->build_mtx);
    RTG_HIP(hipStreamSynchronize(t.stream));
    auto t0 = std::chrono::steady_clock::now();
    if (s->bvh.use_count() > 1 || s->slot != 0)
      RTC_THROW(RTC_ERROR_INVALID_OPERATION, "scene shares its BVH with a batch: commit it on its own first");
    rtg::LearnedCost lc{};
    const rtg::BuildParams p = scene_params(s, lc);
    build_scene_tree(s, d, s->descs, p, t.stream);
    RTG_HIP(hipStreamSynchronize(t.stream));
    if (ms_out) ms_out[i] = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  API_END(s)
}

RTC_API void rtcgTraceCounted(RTCScene scene, void* d_rays, unsigned int M, size_t byteStride, int occluded,
                              unsigned int* d_nodes, unsigned int* d_tris) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  if (!s->committed) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "scene not committed");
  single_level_only(s, "counted traversal");
  if (!d_nodes || !d_tris) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid argument");
  Device* d = s->dev;
  d->activate();
  ThreadDev& t = d->td();
  if (s->bvh->empty(s->slot)) {
    RTG_HIP(hipMemsetAsync(d_nodes, 0, (size_t)M * 4, t.stream));
    RTG_HIP(hipMemsetAsync(d_tris, 0, (size_t)M * 4, t.stream));
  } else {
    rtg::launch_trace(s->bvh->view(s->slot), d_rays, M, byteStride, occluded != 0, RTC_INVALID_GEOMETRY_ID, d_nodes, d_tris,
                      t.errword.as<uint32_t>(), t.stream, d->trace_params(), &t.sorter);
  }
  RTG_HIP(hipStreamSynchronize(t.stream));
  check_trace_error(t);
  API_END(s)
}

RTC_API void rtcgGetSceneStats(RTCScene scene, RTCGSceneStats* out) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  VERIFY_HANDLE(out);
  const rtg::BuildStats& st = s->bvh->stats;
  memset(out, 0, sizeof(*out));
  out->num_prims = st.num_prims;
  out->num_nodes = st.num_nodes;
  out->max_depth = st.max_depth;
  out->node_bytes = (unsigned)rtg::QNODE_BYTES;
  out->tri_bytes = sizeof(rtg::TriRec);
  out->max_stack = st.max_stack;
  out->sah = st.sah;
  out->sah_tri = st.sah_tri;
  out->build_ms = st.ms_total;
  API_END(s)
}

RTC_API void rtcgReadBVH(RTCScene scene, void* nodes, void* tris) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  if (!s->committed) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "scene not committed");
  Device* d = s->dev;
  d->activate();
  ThreadDev& t = d->td();
  const rtg::BuildStats& st = s->bvh->stats;
  // nodes are handed out packed, QNODE_BYTES each (the device keeps each in
  // a 128 B slot)
  std::vector<rtg::QNode8> slots(nodes ? st.num_nodes : 0);
  if (nodes && st.num_nodes)
    RTG_HIP(hipMemcpyAsync(slots.data(), s->bvh->nodes.ptr, (size_t)st.num_nodes * sizeof(rtg::QNode8),
                           hipMemcpyDeviceToHost, t.stream));
  if (tris && st.num_prims)
    RTG_HIP(hipMemcpyAsync(tris, s->bvh->tris.ptr, (size_t)st.num_prims * sizeof(rtg::TriRec), hipMemcpyDeviceToHost,
                           t.stream));
  RTG_HIP(hipStreamSynchronize(t.stream));
  for (size_t i = 0; i < slots.size(); i++)
    memcpy((char*)nodes + i * rtg::QNODE_BYTES, &slots[i], rtg::QNODE_BYTES);
  API_END(s)
}

// ------------------------------------------------------ learned split cost --

}  // extern "C"

namespace {
SplitModelH* SM(RTCGSplitModel m) { return (SplitModelH*)m; }

rtg::SceneNorm scene_norm(const rtg::SceneSlot& st) {
  rtg::SceneNorm sn;
  float e[3];
  for (int k = 0; k < 3; k++) {
    sn.lo[k] = st.bounds_lo[k];
    e[k] = st.bounds_hi[k] - st.bounds_lo[k];
    sn.inv_ext[k] = e[k] > 0.0f ? 1.0f / e[k] : 0.0f;
  }
  const float a = rtg::half_area(e[0], e[1], e[2]);
  sn.inv_area = a > 0.0f ? 1.0f / a : 0.0f;
  return sn;
}

// Spatial splits of a user build at HIGH quality (the reference passes
// splitPrimitive to its spatial-split builder there,
// kernels/common/rtcore_builder.cpp:379-380): while the caller's array has
// room (primitiveArrayCapacity), the primitive whose cut most reduces the
// sum of box half areas is cut by splitPrimitive across the largest axis of
// its box, at the coarsest plane of the dyadic grid over all primitives'
// bounds that lies inside it (the midpoint when rounding puts that plane
// outside); the two halves replace it and may be cut again. Cuts that do
// not reduce the half-area sum are not made.
void user_spatial_splits(const RTCBuildArguments& a, std::vector<RTCBuildPrimitive>& prims) {
  const size_t room = a.primitiveArrayCapacity > prims.size() ? a.primitiveArrayCapacity - prims.size() : 0;
  if (room == 0 || prims.empty()) return;
  float glo[3] = {INFINITY, INFINITY, INFINITY}, ghi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (const RTCBuildPrimitive& p : prims) {
    glo[0] = std::min(glo[0], p.lower_x); glo[1] = std::min(glo[1], p.lower_y); glo[2] = std::min(glo[2], p.lower_z);
    ghi[0] = std::max(ghi[0], p.upper_x); ghi[1] = std::max(ghi[1], p.upper_y); ghi[2] = std::max(ghi[2], p.upper_z);
  }
  auto lo = [](const RTCBuildPrimitive& p, int k) { return k == 0 ? p.lower_x : (k == 1 ? p.lower_y : p.lower_z); };
  auto hi = [](const RTCBuildPrimitive& p, int k) { return k == 0 ? p.upper_x : (k == 1 ? p.upper_y : p.upper_z); };
  auto area = [&](const RTCBuildPrimitive& p) {
    return rtg::half_area(std::max(0.0f, p.upper_x - p.lower_x), std::max(0.0f, p.upper_y - p.lower_y),
                          std::max(0.0f, p.upper_z - p.lower_z));
  };
  struct Cut {
    float gain;
    uint32_t i;
    RTCBuildPrimitive l, r;
  };
  auto cut = [&](uint32_t i, Cut& c) {
    const RTCBuildPrimitive& p = prims[i];
    int ax = 0;
    for (int k = 1; k < 3; k++)
      if (hi(p, k) - lo(p, k) > hi(p, ax) - lo(p, ax)) ax = k;
    const float l = lo(p, ax), h = hi(p, ax), ext = ghi[ax] - glo[ax];
    if (!(h > l) || !(ext > 0.0f)) return false;
    // coarsest dyadic plane in (l, h) of the grid over [glo, ghi]
    const double ul = (l - glo[ax]) / (double)ext, uh = (h - glo[ax]) / (double)ext;
    float pos = 0.5f * (l + h);
    for (int level = 1; level <= 24; level++) {
      const double step = std::ldexp(1.0, -level), m = std::floor(uh / step) * step;
      if (m > ul && m < uh) {
        const float x = (float)(glo[ax] + m * ext);
        if (x > l && x < h) pos = x;
        break;
      }
    }
    RTCBounds lb{}, rb{};
    a.splitPrimitive(&p, (unsigned)ax, pos, &lb, &rb, a.userPtr);
    c.i = i;
    c.l = p;
    c.r = p;
    c.l.lower_x = lb.lower_x; c.l.lower_y = lb.lower_y; c.l.lower_z = lb.lower_z;
    c.l.upper_x = lb.upper_x; c.l.upper_y = lb.upper_y; c.l.upper_z = lb.upper_z;
    c.r.lower_x = rb.lower_x; c.r.lower_y = rb.lower_y; c.r.lower_z = rb.lower_z;
    c.r.upper_x = rb.upper_x; c.r.upper_y = rb.upper_y; c.r.upper_z = rb.upper_z;
    c.gain = area(p) - area(c.l) - area(c.r);
    return c.gain > 0.0f;
  };
  auto cmp = [](const Cut& x, const Cut& y) { return x.gain < y.gain; };
  std::vector<Cut> heap;
  for (uint32_t i = 0; i < prims.size(); i++) {
    Cut c;
    if (cut(i, c)) heap.push_back(c);
  }
  std::make_heap(heap.begin(), heap.end(), cmp);
  size_t made = 0;
  while (made < room && !heap.empty()) {
    std::pop_heap(heap.begin(), heap.end(), cmp);
    const Cut c = heap.back();
    heap.pop_back();
    prims[c.i] = c.l;
    prims.push_back(c.r);
    made++;
    Cut x;
    if (cut(c.i, x)) {
      heap.push_back(x);
      std::push_heap(heap.begin(), heap.end(), cmp);
    }
    if (cut((uint32_t)prims.size() - 1, x)) {
      heap.push_back(x);
      std::push_heap(heap.begin(), heap.end(), cmp);
    }
  }
}

// Host array -> device scratch.
template <class T>
T* upload(rtg::DevBuf& b, const T* host, size_t n, hipStream_t s) {
  b.reserve(n * sizeof(T) + 16);
  if (n) RTG_HIP(hipMemcpyAsync(b.ptr, host, n * sizeof(T), hipMemcpyHostToDevice, s));
  return b.as<T>();
}
}  // namespace

extern "C" {

RTC_API RTCGSplitModel rtcgNewSplitModel(RTCDevice device, unsigned int seed) {
  Device* d = DE(device);
  API_BEGIN
  VERIFY_HANDLE(device);
  d->activate();
  ThreadDev& t = d->td();
  auto* h = new SplitModelH(d);
  try {
    rtg::mlp_init(h->m, seed, t.stream);
    RTG_HIP(hipStreamSynchronize(t.stream));
  } catch (...) {
    ref_dec(h);
    throw;
  }
  return (RTCGSplitModel)h;
  API_END(d)
  return nullptr;
}

RTC_API void rtcgRetainSplitModel(RTCGSplitModel model) {
  SplitModelH* m = SM(model);
  API_BEGIN
  VERIFY_HANDLE(model);
  ref_inc(m);
  API_END(m)
}

RTC_API void rtcgReleaseSplitModel(RTCGSplitModel model) {
  SplitModelH* m = SM(model);
  API_BEGIN
  VERIFY_HANDLE(model);
  ref_dec(m);
  API_END((SplitModelH*)nullptr)
}

RTC_API unsigned int rtcgSplitModelParamCount(void) { return (unsigned)rtg::MLP_NPARAM; }

RTC_API void rtcgSplitModelSetParams(RTCGSplitModel model, const float* params) {
  SplitModelH* m = SM(model);
  API_BEGIN
  VERIFY_HANDLE(model);
  VERIFY_HANDLE(params);
  std::lock_guard<std::mutex> lk(m->mtx);
  m->dev->activate();
  ThreadDev& t = m->dev->td();
  rtg::mlp_set_params(m->m, params, t.stream);
  RTG_HIP(hipStreamSynchronize(t.stream));
  API_END(m)
}

RTC_API void rtcgSplitModelGetParams(RTCGSplitModel model, float* params) {
  SplitModelH* m = SM(model);
  API_BEGIN
  VERIFY_HANDLE(model);
  VERIFY_HANDLE(params);
  std::lock_guard<std::mutex> lk(m->mtx);
  m->dev->activate();
  rtg::mlp_get_params(m->m, params, m->dev->td().stream);
  API_END(m)
}

RTC_API void rtcgSplitModelPredict(RTCGSplitModel model, const float* features, unsigned int n, float* out) {
  SplitModelH* m = SM(model);
  API_BEGIN
  VERIFY_HANDLE(model);
  if (n && (!features || !out)) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid argument");
  std::lock_guard<std::mutex> lk(m->mtx);
  m->dev->activate();
  ThreadDev& t = m->dev->td();
  rtg::DevBuf dx, dy;
  const float* x = upload(dx, features, (size_t)n * rtg::MLP_IN, t.stream);
  dy.reserve((size_t)n * 4 + 16);
  rtg::mlp_predict(m->m, x, n, dy.as<float>(), t.stream);
  if (n) RTG_HIP(hipMemcpyAsync(out, dy.ptr, (size_t)n * 4, hipMemcpyDeviceToHost, t.stream));
  RTG_HIP(hipStreamSynchronize(t.stream));
  API_END(m)
}

RTC_API float rtcgSplitModelGradient(RTCGSplitModel model, const float* features, const float* targets,
                                     const float* weights, unsigned int n, float* grad) {
  SplitModelH* m = SM(model);
  API_BEGIN
  VERIFY_HANDLE(model);
  VERIFY_HANDLE(grad);
  if (n && (!features || !targets || !weights)) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid argument");
  std::lock_guard<std::mutex> lk(m->mtx);
  m->dev->activate();
  ThreadDev& t = m->dev->td();
  rtg::DevBuf dx, dt, dw;
  const float* x = upload(dx, features, (size_t)n * rtg::MLP_IN, t.stream);
  const float* tt = upload(dt, targets, n, t.stream);
  const float* w = upload(dw, weights, n, t.stream);
  const float loss = rtg::mlp_gradient(m->m, x, tt, w, nullptr, n, t.stream);
  RTG_HIP(hipMemcpyAsync(grad, m->m.grad.ptr, rtg::MLP_NPARAM * 4, hipMemcpyDeviceToHost, t.stream));
  RTG_HIP(hipMemsetAsync(m->m.grad.ptr, 0, rtg::MLP_NPARAM * 4, t.stream));
  RTG_HIP(hipStreamSynchronize(t.stream));
  return loss;
  API_END(m)
  return 0.0f;
}

RTC_API void rtcgSplitModelCollect(RTCGSplitModel model, RTCScene scene, const void* d_rays, unsigned int M,
                                   size_t byteStride) {
  SplitModelH* m = SM(model);
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(model);
  VERIFY_HANDLE(scene);
  if (M && !d_rays) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid argument");
  if (!s->committed) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "scene not committed");
  single_level_only(s, "visit collection");
  if (byteStride < sizeof(RTCRayHit) || byteStride % 4) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid stride");
  if (M == 0 || s->bvh->empty(s->slot)) return;
  std::lock_guard<std::mutex> lk(m->mtx);
  std::lock_guard<std::mutex> slk(s->mtx);
  Device* d = s->dev;
  d->activate();
  ThreadDev& t = d->td();
  // a binary tree of the scene's triangles: its parent/child pairs are the
  // binary splits the builder prices
  std::vector<rtg::GeomDesc> descs = s->descs;
  uint32_t off = 0;
  for (rtg::GeomDesc& g : descs) {
    g.scene = 0;
    g.prim_offset = off;
    off += g.num_tris;
  }
  rtg::LearnedCost none{};
  rtg::BuildParams p = scene_params(s, none);
  p.learned = nullptr;
  p.builder = rtg::BUILDER_SAH;
  p.width = 2;
  p.split_budget = 0.0f;
  rtg::DeviceBVH bin;
  {
    std::lock_guard<std::mutex> bl(d->build_mtx);
    rtg::build_bvh(descs, p, d->scratch, bin, t.stream);
  }
  const uint32_t nn = bin.stats.num_nodes;
  if (nn == 0) return;
  rtg::DevBuf copy;
  copy.reserve((size_t)M * byteStride);
  RTG_HIP(hipMemcpyAsync(copy.ptr, d_rays, (size_t)M * byteStride, hipMemcpyDeviceToDevice, t.stream));
  m->m.visits.reserve((size_t)nn * 4);
  m->m.counts.reserve((size_t)M * 8);
  RTG_HIP(hipMemsetAsync(m->m.visits.ptr, 0, (size_t)nn * 4, t.stream));
  uint32_t* cn = m->m.counts.as<uint32_t>();
  rtg::launch_trace(bin.view(0), copy.ptr, M, byteStride, false, RTC_INVALID_GEOMETRY_ID, cn, cn + M,
                    t.errword.as<uint32_t>(), t.stream, d->trace_params(), &t.sorter, m->m.visits.as<uint32_t>());
  const rtg::SceneNorm sn = m->m.has_norm ? m->m.norm : scene_norm(s->bvh->slots[s->slot]);
  rtg::mlp_append_samples(m->m, bin.nodes.as<rtg::QNode8>(), bin.info.as<rtg::NodeInfo>(), m->m.visits.as<uint32_t>(),
                          nn, sn, t.stream);
  m->m.norm = sn;
  m->m.has_norm = true;
  RTG_HIP(hipStreamSynchronize(t.stream));
  check_trace_error(t);
  API_END(m)
}

// Two-level scenes: samples for the top tree (object == NULL) or for one
// instanced scene's own tree, from the visits of the two-level traversal.
// The tree collected on is a width-2 build of the same records or triangles
// (its parent/child pairs are the binary splits the builder prices), swapped
// into the two-level view for the one traversal.
RTC_API void rtcgSplitModelCollectInstanced(RTCGSplitModel model, RTCScene scene, RTCScene object, const void* d_rays,
                                            unsigned int M, size_t byteStride) {
  SplitModelH* m = SM(model);
  Scene* s = SC(scene);
  Scene* o = object ? SC(object) : nullptr;
  API_BEGIN
  VERIFY_HANDLE(model);
  VERIFY_HANDLE(scene);
  if (M && !d_rays) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid argument");
  if (!s->committed) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "scene not committed");
  if (!s->two_level) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "scene has no instances: use rtcgSplitModelCollect");
  if (byteStride < sizeof(RTCRayHit) || byteStride % 4) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid stride");
  check_top(s);
  uint32_t target = 0xffffffffu;
  if (o) {
    for (size_t k = 0; k < s->top_object_scenes.size(); k++)
      if (s->top_object_scenes[k] == o) target = (uint32_t)k;
    if (target == 0xffffffffu) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "object is not instanced by the scene");
  }
  if (M == 0) return;
  std::lock_guard<std::mutex> lk(m->mtx);
  std::lock_guard<std::mutex> slk(s->mtx);
  Device* d = s->dev;
  d->activate();
  ThreadDev& t = d->td();
  rtg::LearnedCost none{};
  rtg::DeviceBVH bin;
  rtg::DevBuf dboxes, drecs, dobjects;
  rtg::DevTop v = s->top_view;
  rtg::SceneNorm sn;
  if (!o) {
    // width-2 top tree over the scene's instance records
    const uint32_t n = (uint32_t)s->top_recs.size();
    dboxes.reserve((size_t)n * 24);
    drecs.reserve((size_t)n * sizeof(rtg::InstRec));
    RTG_HIP(hipMemcpyAsync(dboxes.ptr, s->top_boxes.data(), (size_t)n * 24, hipMemcpyHostToDevice, t.stream));
    RTG_HIP(hipMemcpyAsync(drecs.ptr, s->top_recs.data(), (size_t)n * sizeof(rtg::InstRec), hipMemcpyHostToDevice,
                           t.stream));
    rtg::GeomDesc gd{};
    gd.vertices = (const char*)dboxes.ptr;
    gd.indices = (const char*)drecs.ptr;
    gd.vstride = 24;
    gd.istride = sizeof(rtg::InstRec);
    gd.num_vertices = n;
    gd.num_tris = n;
    gd.kind = rtg::GEOM_INSTANCES;
    rtg::BuildParams p = scene_params(s, none);
    p.learned = nullptr;
    p.builder = rtg::BUILDER_SAH;
    p.max_leaf = 1;
    p.width = 2;
    p.split_budget = 0.0f;
    {
      std::lock_guard<std::mutex> bl(d->build_mtx);
      rtg::build_bvh({gd}, p, d->scratch, bin, t.stream);
    }
    v.nodes = bin.nodes.as<rtg::QNode8>();
    v.insts = bin.insts.as<rtg::InstRec>();
    v.num_nodes = bin.stats.num_nodes;
    v.root = bin.slots.empty() ? 0u : bin.slots[0].root;
    if (bin.stats.max_stack + s->top_obj_stack + 2 > rtg::TRACE_STACK)
      RTC_THROW(RTC_ERROR_UNKNOWN, "two-level tree too deep for the traversal stack");
    rtg::SceneSlot world;
    for (int k = 0; k < 3; k++) {
      world.bounds_lo[k] = s->bounds_lo[k];
      world.bounds_hi[k] = s->bounds_hi[k];
    }
    sn = scene_norm(world);
  } else {
    // width-2 tree of the object's triangles in its slot of the object table
    std::vector<rtg::GeomDesc> descs = o->descs;
    uint32_t off = 0;
    for (rtg::GeomDesc& g : descs) {
      g.scene = 0;
      g.prim_offset = off;
      off += g.num_tris;
    }
    rtg::BuildParams p = scene_params(o, none);
    p.learned = nullptr;
    p.builder = rtg::BUILDER_SAH;
    p.width = 2;
    p.split_budget = 0.0f;
    {
      std::lock_guard<std::mutex> bl(d->build_mtx);
      rtg::build_bvh(descs, p, d->scratch, bin, t.stream);
    }
    std::vector<rtg::DevScene> views = s->top_object_views;
    views[target] = bin.view(0);
    dobjects.reserve(views.size() * sizeof(rtg::DevScene));
    RTG_HIP(hipMemcpyAsync(dobjects.ptr, views.data(), views.size() * sizeof(rtg::DevScene), hipMemcpyHostToDevice,
                           t.stream));
    v.objects = dobjects.as<rtg::DevScene>();
    if (s->top.stats.max_stack + bin.stats.max_stack + 2 > rtg::TRACE_STACK)
      RTC_THROW(RTC_ERROR_UNKNOWN, "two-level tree too deep for the traversal stack");
    sn = scene_norm(o->bvh->slots[o->slot]);
  }
  const uint32_t nn = bin.stats.num_nodes;
  if (nn == 0) return;
  uint64_t obj_size = 0;
  for (const rtg::DevScene& ov : s->top_object_views)
    obj_size = std::max<uint64_t>(obj_size, (uint64_t)ov.num_nodes + ov.num_tris);
  if (o) obj_size = std::max<uint64_t>(obj_size, (uint64_t)bin.stats.num_nodes + bin.stats.num_prims);
  v.max_iter = (uint64_t)v.num_nodes + v.num_insts + (uint64_t)v.num_insts * (obj_size + 2) + 16;
  rtg::DevBuf copy;
  copy.reserve((size_t)M * byteStride);
  RTG_HIP(hipMemcpyAsync(copy.ptr, d_rays, (size_t)M * byteStride, hipMemcpyDeviceToDevice, t.stream));
  m->m.visits.reserve((size_t)nn * 4);
  m->m.counts.reserve(256);
  RTG_HIP(hipMemsetAsync(m->m.visits.ptr, 0, (size_t)nn * 4, t.stream));
  uint32_t* vis = m->m.visits.as<uint32_t>();
  rtg::launch_trace_inst_counted(v, copy.ptr, M, byteStride, t.errword.as<uint32_t>(), m->m.counts.as<uint32_t>(),
                                 o ? nullptr : vis, o ? vis : nullptr, target, t.stream, d->trace_params());
  if (m->m.has_norm) sn = m->m.norm;
  rtg::mlp_append_samples(m->m, bin.nodes.as<rtg::QNode8>(), bin.info.as<rtg::NodeInfo>(), vis, nn, sn, t.stream);
  m->m.norm = sn;
  m->m.has_norm = true;
  RTG_HIP(hipStreamSynchronize(t.stream));
  check_trace_error(t);
  API_END(m)
}

RTC_API unsigned int rtcgSplitModelSampleCount(RTCGSplitModel model) {
  SplitModelH* m = SM(model);
  API_BEGIN
  VERIFY_HANDLE(model);
  return m->m.nsamples;
  API_END(m)
  return 0;
}

RTC_API void rtcgSplitModelGetSamples(RTCGSplitModel model, float* features, float* targets) {
  SplitModelH* m = SM(model);
  API_BEGIN
  VERIFY_HANDLE(model);
  std::lock_guard<std::mutex> lk(m->mtx);
  m->dev->activate();
  ThreadDev& t = m->dev->td();
  const size_t n = m->m.nsamples;
  if (n && features) RTG_HIP(hipMemcpyAsync(features, m->m.X.ptr, n * rtg::MLP_IN * 4, hipMemcpyDeviceToHost, t.stream));
  if (n && targets) RTG_HIP(hipMemcpyAsync(targets, m->m.T.ptr, n * 4, hipMemcpyDeviceToHost, t.stream));
  RTG_HIP(hipStreamSynchronize(t.stream));
  API_END(m)
}

RTC_API void rtcgSplitModelClearSamples(RTCGSplitModel model) {
  SplitModelH* m = SM(model);
  API_BEGIN
  VERIFY_HANDLE(model);
  std::lock_guard<std::mutex> lk(m->mtx);
  m->m.nsamples = 0;
  m->m.has_norm = false;
  API_END(m)
}

RTC_API float rtcgSplitModelTrain(RTCGSplitModel model, unsigned int steps, unsigned int batch, float learning_rate,
                                  unsigned int seed) {
  SplitModelH* m = SM(model);
  API_BEGIN
  VERIFY_HANDLE(model);
  if (!(learning_rate > 0.0f)) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "learning rate must be positive");
  std::lock_guard<std::mutex> lk(m->mtx);
  if (m->m.nsamples == 0) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "split model has no training data");
  m->dev->activate();
  return rtg::mlp_train(m->m, steps, batch, learning_rate, seed, m->dev->td().stream);
  API_END(m)
  return 0.0f;
}

RTC_API void rtcgSetSplitModelMinCount(RTCGSplitModel model, unsigned int min_count) {
  SplitModelH* m = SM(model);
  API_BEGIN
  VERIFY_HANDLE(model);
  m->min_count = min_count < 64 ? 64 : min_count;
  API_END(m)
}

RTC_API void rtcgSetSplitModelSahSlack(RTCGSplitModel model, float slack) {
  SplitModelH* m = SM(model);
  API_BEGIN
  VERIFY_HANDLE(model);
  if (!(slack >= 0.0f)) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "SAH slack must be >= 0");
  m->sah_slack = slack;
  API_END(m)
}

RTC_API void rtcgSetSceneSplitModel(RTCScene scene, RTCGSplitModel model) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  SplitModelH* m = SM(model);
  if (m && m->dev != s->dev) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "split model belongs to another device");
  std::lock_guard<std::mutex> lk(s->mtx);
  if (m) ref_inc(m);
  if (s->model) ref_dec(s->model);
  s->model = m;
  API_END(s)
}

// ------------------------------------------------------------ batches --

RTC_API void rtcgCommitScenes(RTCScene* scenes, unsigned int numScenes) {
  Scene* s0 = (numScenes && scenes) ? SC(scenes[0]) : nullptr;
  API_BEGIN
  if (!scenes || numScenes == 0 || !s0) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid argument");
  if (numScenes > 65535) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "at most 65535 scenes per batch");
  Device* d = s0->dev;
  std::vector<Scene*> sc(numScenes);
  for (unsigned i = 0; i < numScenes; i++) {
    sc[i] = SC(scenes[i]);
    if (!sc[i]) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid argument");
    if (sc[i]->dev != d) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "scenes of a batch must share a device");
    if ((sc[i]->flags & RTC_SCENE_FLAG_ROBUST) != (s0->flags & RTC_SCENE_FLAG_ROBUST))
      RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "scenes of a batch must agree on RTC_SCENE_FLAG_ROBUST");
    if (sc[i]->model) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "batched commits use the plain split cost");
    for (Geometry* g : sc[i]->geoms)
      if (g && g->enabled && g->type == RTC_GEOMETRY_TYPE_INSTANCE)
        RTC_THROW(RTC_ERROR_INVALID_OPERATION, "batched commits hold triangle scenes only");
    for (unsigned j = 0; j < i; j++)
      if (sc[j] == sc[i]) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "a scene appears twice in the batch");
  }
  d->activate();
  ThreadDev& t = d->td();
  // lock in address order (no deadlock with other batch commits)
  std::vector<Scene*> order(sc);
  std::sort(order.begin(), order.end());
  std::vector<std::unique_lock<std::mutex>> locks;
  for (Scene* x : order) locks.emplace_back(x->mtx);
  uint32_t prim_offset = 0;
  std::vector<rtg::GeomDesc> descs;
  for (unsigned i = 0; i < numScenes; i++) upload_geometry(sc[i], d, t, i, prim_offset, descs);
  auto store = std::make_shared<rtg::DeviceBVH>();
  {
    std::lock_guard<std::mutex> bl(d->build_mtx);
    rtg::BuildParams p = build_params(d, s0->quality, s0->flags);
    p.builder = rtg::BUILDER_SAH;
    d->monitor((ssize_t)prim_offset * (80 + 48));
    rtg::build_bvh(descs, p, d->scratch, *store, t.stream, numScenes);
  }
  RTG_HIP(hipStreamSynchronize(t.stream));
  for (unsigned i = 0; i < numScenes; i++) {
    sc[i]->bvh = store;
    sc[i]->slot = i;
    build_top(sc[i], d, t);  // no instances: resets any earlier top tree
    sc[i]->committed = true;
  }
  API_END(s0)
}

}  // extern "C"

namespace {

// Validated device-side table of a batched traversal.
struct PreparedBatch {
  Scene* first = nullptr;
  rtg::DevScene sc{};
  std::vector<rtg::BatchItem> items;
  uint32_t total = 0;
  bool aligned = true;
};

PreparedBatch prepare_batch(const RTCGBatchItem* items, unsigned n, bool occluded) {
  if (!items || n == 0) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid argument");
  PreparedBatch pb;
  const size_t minstride = occluded ? sizeof(RTCRay) : sizeof(RTCRayHit);
  const rtg::DeviceBVH* store = nullptr;
  uint64_t total = 0;
  for (unsigned i = 0; i < n; i++) {
    Scene* s = SC(items[i].scene);
    if (!s) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid argument");
    if (!s->committed) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "scene not committed");
    single_level_only(s, "batched traversal");
    if (!pb.first) {
      pb.first = s;
      store = s->bvh.get();
    } else if (s->bvh.get() != store) {
      RTC_THROW(RTC_ERROR_INVALID_OPERATION, "scenes of a batched traversal must come from one rtcgCommitScenes call");
    }
    if (items[i].numRays == 0 || s->bvh->empty(s->slot)) continue;  // nothing to hit: records stay as they are
    if (!items[i].rays) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid argument");
    if (items[i].byteStride < minstride || items[i].byteStride % 4)
      RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid stride");
    if (!is_device_ptr(s->dev, items[i].rays)) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "batched rays must be in device memory");
    rtg::BatchItem b;
    b.rays = (char*)items[i].rays;
    b.stride = items[i].byteStride;
    b.first = (uint32_t)total;
    b.count = items[i].numRays;
    b.root = s->bvh->view(s->slot).root;
    b.pad = 0;
    pb.aligned &= ((uintptr_t)b.rays % 16 == 0) && (b.stride % 16 == 0);
    total += b.count;
    if (total > 0xffffffffull) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "too many rays in one batch");
    pb.items.push_back(b);
  }
  pb.total = (uint32_t)total;
  pb.sc = pb.first->bvh->view(0);
  return pb;
}

}  // namespace

struct BatchGraphH : RefObj {
  Device* dev;
  std::vector<Scene*> scenes;  // retained
  // The captured kernel arguments hold the shared node and triangle arrays'
  // addresses: the graph keeps the arrays alive and remembers the build
  // generation and each scene's slot, and a launch after any scene was
  // committed again (new arrays, or a rebuild in place) is refused.
  std::shared_ptr<rtg::DeviceBVH> store;
  uint64_t generation = 0;
  std::vector<uint32_t> slots;  // per retained scene
  rtg::DevBuf table, counter, err;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  explicit BatchGraphH(Device* d) : dev(d) { ref_inc(d); }
  ~BatchGraphH() override {
    (void)hipSetDevice(dev->ordinal);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (Scene* s : scenes) ref_dec(s);
    ref_dec(dev);
  }
};

Device* dev_of(BatchGraphH* g) { return g ? g->dev : nullptr; }

extern "C" {

RTC_API void rtcgTraceBatch(const RTCGBatchItem* items, unsigned int numItems, int occluded) {
  Scene* s0 = (items && numItems) ? SC(items[0].scene) : nullptr;
  API_BEGIN
  PreparedBatch pb = prepare_batch(items, numItems, occluded != 0);
  if (pb.items.empty()) return;
  Device* d = pb.first->dev;
  d->activate();
  ThreadDev& t = d->td();
  t.batch.reserve(pb.items.size() * sizeof(rtg::BatchItem));
  t.batch_counter.reserve(256);
  RTG_HIP(hipMemcpyAsync(t.batch.ptr, pb.items.data(), pb.items.size() * sizeof(rtg::BatchItem), hipMemcpyHostToDevice,
                         t.stream));
  rtg::launch_trace_batch(pb.sc, t.batch.as<rtg::BatchItem>(), (uint32_t)pb.items.size(), pb.total, occluded != 0,
                          pb.aligned, t.errword.as<uint32_t>(), t.batch_counter.as<uint32_t>(), t.stream, d->trace_params());
  RTG_HIP(hipStreamSynchronize(t.stream));
  check_trace_error(t);
  API_END(s0)
}

RTC_API float rtcgAutotuneBatch(const RTCGBatchItem* items, unsigned int numItems, int occluded, char* out,
                                size_t outSize) {
  Scene* s0 = (items && numItems) ? SC(items[0].scene) : nullptr;
  API_BEGIN
  PreparedBatch pb = prepare_batch(items, numItems, occluded != 0);
  if (pb.items.empty()) return 0.0f;
  // the shared node and triangle arrays stay alive while the tuner traces
  // them, whatever another thread commits meanwhile
  const std::shared_ptr<rtg::DeviceBVH> keep = pb.first->bvh;
  Device* d = pb.first->dev;
  d->activate();
  ThreadDev& t = d->td();
  // the items' rays packed into a pristine copy (src) and a work copy of the
  // same layout, which each timed launch starts from (the caller's records
  // are only read)
  std::vector<rtg::BatchItem> wi = pb.items;
  std::vector<size_t> off(wi.size());
  size_t bytes = 0;
  bool aligned = true;
  for (size_t i = 0; i < wi.size(); i++) {
    off[i] = bytes;
    bytes += (size_t)wi[i].count * wi[i].stride;
    aligned &= wi[i].stride % 16 == 0;
  }
  rtg::DevBuf src, work, table;
  src.reserve(bytes);
  work.reserve(bytes);
  for (size_t i = 0; i < wi.size(); i++) {
    RTG_HIP(hipMemcpyAsync(src.as<char>() + off[i], wi[i].rays, (size_t)wi[i].count * wi[i].stride,
                           hipMemcpyDeviceToDevice, t.stream));
    wi[i].rays = work.as<char>() + off[i];
  }
  table.reserve(wi.size() * sizeof(rtg::BatchItem));
  RTG_HIP(hipMemcpyAsync(table.ptr, wi.data(), wi.size() * sizeof(rtg::BatchItem), hipMemcpyHostToDevice, t.stream));
  t.batch_counter.reserve(256);
  auto time_launch = [&](const rtg::TraceParams& c) {
    float total = 0.0f;
    const int reps = 2;
    for (int i = -1; i < reps; i++) {  // one untimed warm-up
      RTG_HIP(hipMemcpyAsync(work.ptr, src.ptr, bytes, hipMemcpyDeviceToDevice, t.stream));
      RTG_HIP(hipEventRecord(t.ev0, t.stream));
      rtg::launch_trace_batch(pb.sc, table.as<rtg::BatchItem>(), (uint32_t)wi.size(), pb.total, occluded != 0,
                              aligned, t.errword.as<uint32_t>(), t.batch_counter.as<uint32_t>(), t.stream, c);
      RTG_HIP(hipEventRecord(t.ev1, t.stream));
      RTG_HIP(hipEventSynchronize(t.ev1));
      float ms = 0.0f;
      RTG_HIP(hipEventElapsedTime(&ms, t.ev0, t.ev1));
      if (i >= 0) total += ms;
    }
    check_trace_error(t);
    return total / (float)reps;
  };
  const rtg::TraceParams tp = d->trace_params();
  rtg::TraceParams bestt = tp;
  float best = INFINITY;
  for (int slots : {8, 12})
    for (int refill : {8, 12, 16, 24})
      for (int bpc : {4, 6, 8}) {
        rtg::TraceParams c = tp;
        c.lds_slots = slots;
        if (occluded) c.refill_min_occ = refill; else c.refill_min = refill;
        c.blocks_per_cu = bpc;
        const float ms = time_launch(c);
        if (ms < best) {
          best = ms;
          bestt = c;
        }
      }
  {
    std::lock_guard<std::mutex> g(d->settings_mtx);
    d->trace.lds_slots = bestt.lds_slots;
    if (occluded) d->trace.refill_min_occ = bestt.refill_min_occ; else d->trace.refill_min = bestt.refill_min;
    d->trace.blocks_per_cu = bestt.blocks_per_cu;
  }
  if (out && outSize) {
    char buf[128];
    snprintf(buf, sizeof(buf), "trace_stack=%d,%s=%d,blocks_per_cu=%d", bestt.lds_slots,
             occluded ? "refill_min_occ" : "refill_min", occluded ? bestt.refill_min_occ : bestt.refill_min,
             bestt.blocks_per_cu);
    strncpy(out, buf, outSize - 1);
    out[outSize - 1] = 0;
  }
  return (float)pb.total / (best * 1e-3f) / 1e6f;
  API_END(s0)
  return 0.0f;
}

RTC_API RTCGBatchGraph rtcgNewBatchGraph(const RTCGBatchItem* items, unsigned int numItems, int occluded) {
  Scene* s0 = (items && numItems) ? SC(items[0].scene) : nullptr;
  API_BEGIN
  PreparedBatch pb = prepare_batch(items, numItems, occluded != 0);
  Device* d = pb.first->dev;
  d->activate();
  ThreadDev& t = d->td();
  auto* g = new BatchGraphH(d);
  try {
    g->store = pb.first->bvh;
    g->generation = g->store->generation;
    for (unsigned i = 0; i < numItems; i++) {
      Scene* s = SC(items[i].scene);
      ref_inc(s);
      g->scenes.push_back(s);
      g->slots.push_back(s->slot);
    }
    g->table.reserve(pb.items.size() * sizeof(rtg::BatchItem) + 16);
    g->counter.reserve(256);
    g->err.reserve(256);
    if (!pb.items.empty())
      RTG_HIP(hipMemcpyAsync(g->table.ptr, pb.items.data(), pb.items.size() * sizeof(rtg::BatchItem),
                             hipMemcpyHostToDevice, t.stream));
    RTG_HIP(hipMemsetAsync(g->err.ptr, 0, 4, t.stream));
    RTG_HIP(hipStreamSynchronize(t.stream));
    // capture: counter clear (a kernel, not a memset node) + the persistent traversal
    RTG_HIP(hipStreamBeginCapture(t.stream, hipStreamCaptureModeThreadLocal));
    hipError_t ce = hipSuccess;
    try {
      rtg::launch_trace_batch(pb.sc, g->table.as<rtg::BatchItem>(), (uint32_t)pb.items.size(), pb.total,
                              occluded != 0, pb.aligned, g->err.as<uint32_t>(), g->counter.as<uint32_t>(), t.stream,
                              d->trace_params());
    } catch (...) {
      hipGraph_t tmp = nullptr;
      (void)hipStreamEndCapture(t.stream, &tmp);
      if (tmp) (void)hipGraphDestroy(tmp);
      throw;
    }
    ce = hipStreamEndCapture(t.stream, &g->graph);
    RTG_HIP(ce);
    RTG_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  } catch (...) {
    ref_dec(g);
    throw;
  }
  return (RTCGBatchGraph)g;
  API_END(s0)
  return nullptr;
}

RTC_API void rtcgLaunchBatchGraph(RTCGBatchGraph graph, void* hipStream) {
  BatchGraphH* g = (BatchGraphH*)graph;
  API_BEGIN
  VERIFY_HANDLE(graph);
  if (g->store->generation != g->generation)
    RTC_THROW(RTC_ERROR_INVALID_OPERATION, "the batch's arrays were rebuilt since the graph was made: make a new graph");
  for (size_t i = 0; i < g->scenes.size(); i++)
    if (g->scenes[i]->bvh != g->store || g->scenes[i]->slot != g->slots[i])
      RTC_THROW(RTC_ERROR_INVALID_OPERATION,
                "a scene of the batch graph was committed again since the graph was made: make a new graph");
  g->dev->activate();
  ThreadDev& t = g->dev->td();
  hipStream_t st = hipStream ? (hipStream_t)hipStream : t.stream;
  RTG_HIP(hipGraphLaunch(g->exec, st));
  if (!hipStream) {
    uint32_t err = 0;
    RTG_HIP(hipMemcpyAsync(&err, g->err.ptr, 4, hipMemcpyDeviceToHost, st));
    RTG_HIP(hipStreamSynchronize(st));
    if (err) {
      RTG_HIP(hipMemsetAsync(g->err.ptr, 0, 4, st));
      RTG_HIP(hipStreamSynchronize(st));
      RTC_THROW(RTC_ERROR_UNKNOWN, (err & 1u) ? "traversal stack overflow" : "corrupt acceleration structure");
    }
  }
  API_END(g)
}

RTC_API void rtcgReleaseBatchGraph(RTCGBatchGraph graph) {
  BatchGraphH* g = (BatchGraphH*)graph;
  API_BEGIN
  VERIFY_HANDLE(graph);
  ref_dec(g);
  API_END((BatchGraphH*)nullptr)
}

// -------------------------------------------------------------- autotune --

}  // extern "C"

namespace {

// Mean kernel time (ms) of `reps` closest-hit or occlusion launches over a
// scratch copy of the rays (restored before each launch, not timed).
float time_config(Scene* s, ThreadDev& t, const rtg::TraceParams& tp, const void* src, void* work, unsigned M,
                  size_t stride, bool occluded, int reps) {
  float total = 0.0f;
  for (int i = -1; i < reps; i++) {  // one untimed warm-up
    RTG_HIP(hipMemcpyAsync(work, src, (size_t)M * stride, hipMemcpyDeviceToDevice, t.stream));
    RTG_HIP(hipEventRecord(t.ev0, t.stream));
    launch_scene(s, t, work, M, stride, occluded, RTC_INVALID_GEOMETRY_ID, t.stream, tp);
    RTG_HIP(hipEventRecord(t.ev1, t.stream));
    RTG_HIP(hipEventSynchronize(t.ev1));
    float ms = 0.0f;
    RTG_HIP(hipEventElapsedTime(&ms, t.ev0, t.ev1));
    if (i >= 0) total += ms;
  }
  check_trace_error(t);
  return total / (float)reps;
}

// Rebuilds the scene's tree with the given node width, leaf size and
// collapse triangle cost (the device's settings stay untouched).
void rebuild(Scene* s, Device* d, ThreadDev& t, const rtg::BuildParams& shape) {
  std::lock_guard<std::mutex> bl(d->build_mtx);
  rtg::LearnedCost lc{};
  const rtg::BuildParams p = scene_params(s, lc, &shape);
  build_scene_tree(s, d, s->descs, p, t.stream);
  RTG_HIP(hipStreamSynchronize(t.stream));
}

}  // namespace

extern "C" {

RTC_API float rtcgAutotune(RTCScene scene, const void* d_rays, unsigned int M, size_t byteStride, int occluded,
                           int rebuildTree, char* out, size_t outSize) {
  Scene* s = SC(scene);
  API_BEGIN
  VERIFY_HANDLE(scene);
  if (!s->committed) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "scene not committed");
  single_level_only(s, "autotuning");
  if (!d_rays || M == 0) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid argument");
  if (byteStride < (occluded ? sizeof(RTCRay) : sizeof(RTCRayHit)) || byteStride % 4)
    RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid stride");
  if (rebuildTree && (s->bvh.use_count() > 1 || s->slot != 0))
    RTC_THROW(RTC_ERROR_INVALID_OPERATION, "scene shares its BVH with a batch: commit it on its own first");
  Device* d = s->dev;
  d->activate();
  ThreadDev& t = d->td();
  std::lock_guard<std::mutex> lk(s->mtx);
  if (s->bvh->empty(s->slot)) return 0.0f;
  rtg::DevBuf work;
  work.reserve((size_t)M * byteStride);
  const int reps = 2;
  // The sweep works on local copies; the device's settings change once, at
  // the end, under its settings lock, so that commits and queries on other
  // threads never see a trial value.
  const rtg::TraceParams tp = d->trace_params();
  const rtg::BuildParams start = d->build_settings();
  rtg::BuildParams bestb = start;
  float best = time_config(s, t, tp, d_rays, work.ptr, M, byteStride, occluded != 0, reps);
  if (rebuildTree) {
    // tree shape: node width, leaf size, triangle cost of the collapse
    try {
      for (uint32_t w : {4u, 8u})
        for (uint32_t leaf : {4u, 6u, 8u, 12u})
          for (float ct : {0.3f, 0.45f, 0.6f}) {
            rtg::BuildParams c = start;
            c.width = w;
            c.max_leaf = leaf;
            c.cost_tri = ct;
            rebuild(s, d, t, c);
            const float ms = time_config(s, t, tp, d_rays, work.ptr, M, byteStride, occluded != 0, reps);
            if (ms < best) {
              best = ms;
              bestb = c;
            }
          }
      rebuild(s, d, t, bestb);
    } catch (...) {
      rebuild(s, d, t, start);  // the scene keeps the tree of the device's settings
      throw;
    }
    best = time_config(s, t, tp, d_rays, work.ptr, M, byteStride, occluded != 0, reps);
  }
  // traversal: LDS stack entries, refill threshold, resident blocks per CU
  rtg::TraceParams bestt = tp;
  for (int slots : {0, 8, 12, 16})
    for (int refill : {8, 12, 16, 24, 32})
      for (int bpc : {4, 6, 8}) {
        rtg::TraceParams c = tp;
        c.lds_slots = slots;
        if (occluded) c.refill_min_occ = refill; else c.refill_min = refill;
        c.blocks_per_cu = bpc;
        c.persistent = 1;
        const float ms = time_config(s, t, c, d_rays, work.ptr, M, byteStride, occluded != 0, reps);
        if (ms < best) {
          best = ms;
          bestt = c;
        }
      }
  {
    std::lock_guard<std::mutex> g(d->settings_mtx);
    d->build.width = bestb.width;
    d->build.max_leaf = bestb.max_leaf;
    d->build.cost_tri = bestb.cost_tri;
    d->trace = bestt;
  }
  if (out && outSize) {
    char buf[256];
    snprintf(buf, sizeof(buf), "bvh_width=%u,max_leaf=%u,cost_tri=%g,trace_stack=%d,%s=%d,blocks_per_cu=%d",
             bestb.width, bestb.max_leaf, (double)bestb.cost_tri, bestt.lds_slots,
             occluded ? "refill_min_occ" : "refill_min", occluded ? bestt.refill_min_occ : bestt.refill_min,
             bestt.blocks_per_cu);
    strncpy(out, buf, outSize - 1);
    out[outSize - 1] = 0;
  }
  return (float)M / (best * 1e-3f) / 1e6f;
  API_END(s)
  return 0.0f;
}

// ------------------------------------------------- packet and SoA forms --

}  // extern "C"

namespace {

// A run of rays in SoA form: field f of lane i at r[f] + 4 i (ray fields in
// RTCRay order) and h[f] + 4 i (hit fields in RTCHit order; unused for
// occlusion). valid (may be null): a lane is active only where it holds -1,
// as the reference reads the mask (kernels/bvh/bvh_intersector_hybrid.cpp:131,386).
struct SoaRun {
  char* r[12];
  char* h[8];
  unsigned n;
  const int* valid;
};

SoaRun packet_run(void* base, unsigned K, const int* valid, bool with_hit) {
  SoaRun run;
  char* b = (char*)base;
  for (int f = 0; f < 12; f++) run.r[f] = b + (size_t)f * K * 4;
  for (int f = 0; f < 8; f++) run.h[f] = with_hit ? b + (size_t)(12 + f) * K * 4 : nullptr;
  run.n = K;
  run.valid = valid;
  return run;
}

// Gathers the active lanes into AoS records, traces them as one stream and
// writes back what rtcIntersect1M / rtcOccluded1M would have written.
void trace_runs(Scene* sc, const RTCIntersectContext* ctx, const std::vector<SoaRun>& runs, bool occluded) {
  if (!sc->committed) RTC_THROW(RTC_ERROR_INVALID_OPERATION, "scene not committed");
  std::vector<RTCRayHit> aos;
  std::vector<std::pair<uint32_t, uint32_t>> where;
  for (uint32_t ri = 0; ri < runs.size(); ri++) {
    const SoaRun& run = runs[ri];
    for (uint32_t i = 0; i < run.n; i++) {
      if (run.valid && run.valid[i] != -1) continue;
      RTCRayHit rec;
      uint32_t* w = reinterpret_cast<uint32_t*>(&rec.ray);
      for (int f = 0; f < 12; f++) memcpy(&w[f], run.r[f] + 4 * (size_t)i, 4);
      if (!(rec.ray.tnear <= rec.ray.tfar)) continue;
      memset(&rec.hit, 0, sizeof(rec.hit));
      rec.hit.geomID = rec.hit.primID = rec.hit.instID[0] = RTC_INVALID_GEOMETRY_ID;
      aos.push_back(rec);
      where.emplace_back(ri, i);
    }
  }
  if (aos.empty()) return;
  if (occluded) {
    std::vector<RTCRay> rays(aos.size());
    for (size_t k = 0; k < aos.size(); k++) rays[k] = aos[k].ray;
    trace_stream(sc, ctx, rays.data(), (unsigned)rays.size(), sizeof(RTCRay), true);
    for (size_t k = 0; k < rays.size(); k++)
      if (rays[k].tfar == -INFINITY) memcpy(runs[where[k].first].r[8] + 4 * (size_t)where[k].second, &rays[k].tfar, 4);
  } else {
    trace_stream(sc, ctx, aos.data(), (unsigned)aos.size(), sizeof(RTCRayHit), false);
    for (size_t k = 0; k < aos.size(); k++) {
      if (aos[k].hit.geomID == RTC_INVALID_GEOMETRY_ID) continue;
      const SoaRun& run = runs[where[k].first];
      const size_t o = 4 * (size_t)where[k].second;
      memcpy(run.r[8] + o, &aos[k].ray.tfar, 4);
      const uint32_t* hw = reinterpret_cast<const uint32_t*>(&aos[k].hit);
      for (int f = 0; f < 8; f++)
        if (run.h[f]) memcpy(run.h[f] + o, &hw[f], 4);
    }
  }
}

template <unsigned K>
void packet_call(const int* valid, RTCScene scene, RTCIntersectContext* ctx, void* rays, bool occluded) {
  Scene* s = SC(scene);
  VERIFY_HANDLE(scene);
  VERIFY_HANDLE(rays);
  VERIFY_HANDLE(valid);
  trace_runs(s, ctx, {packet_run(rays, K, valid, !occluded)}, occluded);
}

void np_call(RTCScene scene, RTCIntersectContext* ctx, const void* np, unsigned N, bool occluded) {
  Scene* s = SC(scene);
  VERIFY_HANDLE(scene);
  VERIFY_HANDLE(np);
  SoaRun run;
  char* const* p = reinterpret_cast<char* const*>(np);  // RTCRayNp then RTCHitNp: arrays of field pointers
  for (int f = 0; f < 12; f++) run.r[f] = p[f];
  for (int f = 0; f < 8; f++) run.h[f] = occluded ? nullptr : p[12 + f];
  run.n = N;
  run.valid = nullptr;
  trace_runs(s, ctx, {run}, occluded);
}

void nm_call(RTCScene scene, RTCIntersectContext* ctx, void* base, unsigned N, unsigned M, size_t byteStride,
             bool occluded) {
  Scene* s = SC(scene);
  VERIFY_HANDLE(scene);
  if (M && !base) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid argument");
  std::vector<SoaRun> runs;
  for (unsigned j = 0; j < M; j++) runs.push_back(packet_run((char*)base + (size_t)j * byteStride, N, nullptr, !occluded));
  trace_runs(s, ctx, runs, occluded);
}

void mp_call(RTCScene scene, RTCIntersectContext* ctx, void** recs, unsigned M, bool occluded) {
  Scene* s = SC(scene);
  VERIFY_HANDLE(scene);
  if (M && !recs) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid argument");
  // AoS records reached through pointers: each is a run of one lane with field stride 4
  std::vector<SoaRun> runs;
  for (unsigned j = 0; j < M; j++) {
    SoaRun run;
    char* b = (char*)recs[j];
    if (!b) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid argument");
    for (int f = 0; f < 12; f++) run.r[f] = b + 4 * f;
    for (int f = 0; f < 8; f++) run.h[f] = occluded ? nullptr : b + 48 + 4 * f;
    run.n = 1;
    run.valid = nullptr;
    runs.push_back(run);
  }
  trace_runs(s, ctx, runs, occluded);
}

}  // namespace

extern "C" {

RTC_API void rtcIntersect4(const int* valid, RTCScene scene, RTCIntersectContext* context, RTCRayHit4* rayhit) {
  API_BEGIN
  packet_call<4>(valid, scene, context, rayhit, false);
  API_END(SC(scene))
}
RTC_API void rtcIntersect8(const int* valid, RTCScene scene, RTCIntersectContext* context, RTCRayHit8* rayhit) {
  API_BEGIN
  packet_call<8>(valid, scene, context, rayhit, false);
  API_END(SC(scene))
}
RTC_API void rtcIntersect16(const int* valid, RTCScene scene, RTCIntersectContext* context, RTCRayHit16* rayhit) {
  API_BEGIN
  packet_call<16>(valid, scene, context, rayhit, false);
  API_END(SC(scene))
}
RTC_API void rtcOccluded4(const int* valid, RTCScene scene, RTCIntersectContext* context, RTCRay4* ray) {
  API_BEGIN
  packet_call<4>(valid, scene, context, ray, true);
  API_END(SC(scene))
}
RTC_API void rtcOccluded8(const int* valid, RTCScene scene, RTCIntersectContext* context, RTCRay8* ray) {
  API_BEGIN
  packet_call<8>(valid, scene, context, ray, true);
  API_END(SC(scene))
}
RTC_API void rtcOccluded16(const int* valid, RTCScene scene, RTCIntersectContext* context, RTCRay16* ray) {
  API_BEGIN
  packet_call<16>(valid, scene, context, ray, true);
  API_END(SC(scene))
}
RTC_API void rtcIntersect1Mp(RTCScene scene, RTCIntersectContext* context, RTCRayHit** rayhit, unsigned int M) {
  API_BEGIN
  mp_call(scene, context, (void**)rayhit, M, false);
  API_END(SC(scene))
}
RTC_API void rtcOccluded1Mp(RTCScene scene, RTCIntersectContext* context, RTCRay** ray, unsigned int M) {
  API_BEGIN
  mp_call(scene, context, (void**)ray, M, true);
  API_END(SC(scene))
}
RTC_API void rtcIntersectNM(RTCScene scene, RTCIntersectContext* context, RTCRayHitN* rayhit, unsigned int N,
                            unsigned int M, size_t byteStride) {
  API_BEGIN
  nm_call(scene, context, rayhit, N, M, byteStride, false);
  API_END(SC(scene))
}
RTC_API void rtcOccludedNM(RTCScene scene, RTCIntersectContext* context, RTCRayN* ray, unsigned int N, unsigned int M,
                           size_t byteStride) {
  API_BEGIN
  nm_call(scene, context, ray, N, M, byteStride, true);
  API_END(SC(scene))
}
RTC_API void rtcIntersectNp(RTCScene scene, RTCIntersectContext* context, const RTCRayHitNp* rayhit, unsigned int N) {
  API_BEGIN
  np_call(scene, context, rayhit, N, false);
  API_END(SC(scene))
}
RTC_API void rtcOccludedNp(RTCScene scene, RTCIntersectContext* context, const RTCRayNp* ray, unsigned int N) {
  API_BEGIN
  np_call(scene, context, ray, N, true);
  API_END(SC(scene))
}

}  // extern "C"

// ------------------------------------------------------------ user BVH --

namespace {

// Walks a collapsed user tree and calls the caller's callbacks, in the
// reference binned-SAH builder's order per node: createNode, setNodeBounds,
// the children, setNodeChildren (kernels/common/rtcore_builder.cpp:169-188).
struct UserWalk {
  const rtg::UserTree& T;
  const RTCBuildArguments& a;
  RTCBuildPrimitive* prims;  // in tree position order
  RTCThreadLocalAllocator alloc;
  size_t done = 0;

  bool inner(uint32_t c) const { return !(c & rtg::REF_LEAF) && !T.isleaf[c]; }

  void box(uint32_t c, RTCBounds& b) const {
    const float* f = (c & rtg::REF_LEAF) ? &T.pbox[8 * (size_t)(c & ~rtg::REF_LEAF)] : &T.nbox[8 * (size_t)c];
    b.lower_x = f[0]; b.lower_y = f[1]; b.lower_z = f[2]; b.align0 = 0.0f;
    b.upper_x = f[4]; b.upper_y = f[5]; b.upper_z = f[6]; b.align1 = 0.0f;
  }

  void progress(size_t k) {
    done += k;
    if (a.buildProgress && !a.buildProgress(a.userPtr, std::min(1.0, (double)done / (double)T.n)))
      RTC_THROW(RTC_ERROR_CANCELLED, "progress monitor forced termination");
  }

  void* leaf(uint32_t c) {
    uint32_t first, count;
    if (c & rtg::REF_LEAF) {
      first = c & ~rtg::REF_LEAF;
      count = 1;
    } else {
      first = T.range[2 * (size_t)c];
      count = T.range[2 * (size_t)c + 1] - first + 1;
    }
    void* l = a.createLeaf(alloc, prims + first, count, a.userPtr);
    progress(count);
    return l;
  }

  // the children the collapse chose for inner node r (k_collapse's walk)
  int children(uint32_t r, uint32_t ch[8]) const {
    const uint32_t W = T.width;
    uint32_t st_ref[8], st_s[8];
    int sp = 0, n = 0;
    const uint32_t j8 = T.split8[8 * (size_t)r + W - 1] & 7u;
    st_ref[sp] = T.child[2 * (size_t)r + 1]; st_s[sp++] = W - j8;
    st_ref[sp] = T.child[2 * (size_t)r];     st_s[sp++] = j8;
    while (sp > 0 && n < 8) {
      sp--;
      const uint32_t c = st_ref[sp];
      uint32_t slots = st_s[sp];
      if (!(c & rtg::REF_LEAF)) {
        while (slots > 1 && (T.split8[8 * (size_t)c + slots - 1] & 0x80u)) slots--;
      } else {
        slots = 1;
      }
      if (slots <= 1) {
        ch[n++] = c;
      } else {
        if (sp + 2 > 8) RTC_THROW(RTC_ERROR_UNKNOWN, "inconsistent user tree");
        const uint32_t jj = T.split8[8 * (size_t)c + slots - 1] & 7u;
        st_ref[sp] = T.child[2 * (size_t)c + 1]; st_s[sp++] = slots - jj;
        st_ref[sp] = T.child[2 * (size_t)c];     st_s[sp++] = jj;
      }
    }
    return n;
  }

  void* node(uint32_t r, int depth) {
    if (!inner(r)) return leaf(r);
    if (depth > 4096) RTC_THROW(RTC_ERROR_UNKNOWN, "user tree too deep");
    uint32_t ch[8];
    const int n = children(r, ch);
    void* nd = a.createNode(alloc, (unsigned)n, a.userPtr);
    RTCBounds bb[8];
    const RTCBounds* bp[8];
    for (int k = 0; k < n; k++) {
      box(ch[k], bb[k]);
      bp[k] = &bb[k];
    }
    a.setNodeBounds(nd, bp, (unsigned)n, a.userPtr);
    void* cp[8];
    for (int k = 0; k < n; k++) cp[k] = node(ch[k], depth + 1);
    a.setNodeChildren(nd, cp, (unsigned)n, a.userPtr);
    return nd;
  }
};

}  // namespace

extern "C" {

RTC_API RTCBVH rtcNewBVH(RTCDevice device) {
  Device* d = DE(device);
  API_BEGIN
  VERIFY_HANDLE(device);
  return (RTCBVH) new BVHUser(d);
  API_END(d)
  return nullptr;
}

RTC_API void rtcRetainBVH(RTCBVH bvh) {
  BVHUser* b = (BVHUser*)bvh;
  API_BEGIN
  VERIFY_HANDLE(bvh);
  ref_inc(b);
  API_END(b)
}

RTC_API void rtcReleaseBVH(RTCBVH bvh) {
  BVHUser* b = (BVHUser*)bvh;
  API_BEGIN
  VERIFY_HANDLE(bvh);
  ref_dec(b);
  API_END((BVHUser*)nullptr)
}

RTC_API void* rtcThreadLocalAlloc(RTCThreadLocalAllocator allocator, size_t bytes, size_t align) {
  API_BEGIN
  VERIFY_HANDLE(allocator);
  return ((Arena*)allocator)->alloc(bytes, align);
  API_END((BVHUser*)nullptr)
  return nullptr;
}

// Reference kernels/common/rtcore_builder.cpp:355-395.
RTC_API void* rtcBuildBVH(const RTCBuildArguments* args) {
  BVHUser* b = args ? (BVHUser*)args->bvh : nullptr;
  API_BEGIN
  VERIFY_HANDLE(args);
  VERIFY_HANDLE(b);
  VERIFY_HANDLE(args->createNode);
  VERIFY_HANDLE(args->setNodeChildren);
  VERIFY_HANDLE(args->setNodeBounds);
  VERIFY_HANDLE(args->createLeaf);
  if (args->primitiveArrayCapacity < args->primitiveCount)
    RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "primitiveArrayCapacity must be greater or equal to primitiveCount");
  if (args->buildQuality != RTC_BUILD_QUALITY_LOW && args->buildQuality != RTC_BUILD_QUALITY_MEDIUM &&
      args->buildQuality != RTC_BUILD_QUALITY_HIGH)
    RTC_THROW(RTC_ERROR_INVALID_OPERATION, "invalid build quality");
  if (args->primitiveCount && !args->primitives) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "invalid argument");
  if (args->primitiveCount > 0x7fffffffull) RTC_THROW(RTC_ERROR_INVALID_ARGUMENT, "too many primitives");
  std::lock_guard<std::mutex> lk(b->mtx);
  b->arena.reset();
  RTCThreadLocalAllocator alloc = (RTCThreadLocalAllocator)&b->arena;
  if (args->primitiveCount == 0) return args->createLeaf(alloc, args->primitives, 0, args->userPtr);
  std::vector<RTCBuildPrimitive> prims(args->primitives, args->primitives + args->primitiveCount);
  if (args->buildQuality == RTC_BUILD_QUALITY_HIGH && args->splitPrimitive)
    user_spatial_splits(*args, prims);
  const uint32_t n = (uint32_t)prims.size();
  Device* d = b->dev;
  d->activate();
  ThreadDev& t = d->td();
  // boxes to the device as GEOM_INSTANCES primitives (lo xyz, hi xyz)
  std::vector<float> boxes((size_t)n * 6);
  for (uint32_t i = 0; i < n; i++) {
    const RTCBuildPrimitive& p = prims[i];
    float* f = &boxes[6 * (size_t)i];
    f[0] = p.lower_x; f[1] = p.lower_y; f[2] = p.lower_z;
    f[3] = p.upper_x; f[4] = p.upper_y; f[5] = p.upper_z;
  }
  rtg::DevBuf dboxes;
  dboxes.reserve(boxes.size() * 4);
  RTG_HIP(hipMemcpyAsync(dboxes.ptr, boxes.data(), boxes.size() * 4, hipMemcpyHostToDevice, t.stream));
  rtg::GeomDesc gd{};
  gd.vertices = (const char*)dboxes.ptr;
  gd.indices = nullptr;
  gd.vstride = 24;
  gd.istride = 0;
  gd.num_vertices = n;
  gd.num_tris = n;
  gd.kind = rtg::GEOM_INSTANCES;
  rtg::BuildParams p = d->build_settings();
  p.builder = args->buildQuality == RTC_BUILD_QUALITY_LOW ? rtg::BUILDER_MORTON : rtg::BUILDER_SAH;
  p.width = std::min(8u, std::max(2u, args->maxBranchingFactor));
  p.max_leaf = std::min<unsigned>(RTC_BUILD_MAX_PRIMITIVES_PER_LEAF, std::max(1u, args->maxLeafSize));
  p.min_leaf = std::max(1u, std::min(args->minLeafSize, p.max_leaf));
  p.block = std::max(1u, args->sahBlockSize);
  p.cost_node = args->traversalCost;
  p.cost_tri = args->intersectionCost;
  p.learned = nullptr;
  p.robust = 0;
  rtg::UserTree tree;
  rtg::DeviceBVH scratch_out;
  {
    std::lock_guard<std::mutex> bl(d->build_mtx);
    rtg::build_bvh({gd}, p, d->scratch, scratch_out, t.stream, 1, &tree);
  }
  RTG_HIP(hipStreamSynchronize(t.stream));
  // the primitive array in tree order (left-out primitives after the placed ones)
  std::vector<RTCBuildPrimitive> sorted((size_t)n);
  std::vector<uint8_t> placed(n, 0);
  for (uint32_t q = 0; q < tree.n; q++) {
    uint32_t src;
    memcpy(&src, &tree.pbox[8 * (size_t)q + 7], 4);
    if (src >= n || placed[src]) RTC_THROW(RTC_ERROR_UNKNOWN, "inconsistent user tree");
    placed[src] = 1;
    sorted[q] = prims[src];
  }
  uint32_t q = tree.n;
  for (uint32_t i = 0; i < n; i++)
    if (!placed[i]) sorted[q++] = prims[i];
  memcpy(args->primitives, sorted.data(), sorted.size() * sizeof(RTCBuildPrimitive));
  if (tree.n == 0) return args->createLeaf(alloc, args->primitives, 0, args->userPtr);
  UserWalk w{tree, *args, args->primitives, alloc};
  if (args->buildProgress && !args->buildProgress(args->userPtr, 0.0))
    RTC_THROW(RTC_ERROR_CANCELLED, "progress monitor forced termination");
  return w.node(tree.root, 0);
  API_END(b)
  return nullptr;
}

}  // extern "C"
