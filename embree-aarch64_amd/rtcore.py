"""Python mirror of the triangle ray-query C API (ctypes over lib/libembree3.so).

The functions keep the reference API's names, argument meaning and error
behaviour (include/embree3/rtcore_*.h of the reference); numpy structured
arrays stand in for RTCRayHit / RTCRay records. All compute happens in the
HIP library: importing this module without a built library raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libembree3.so")

RTC_INVALID_GEOMETRY_ID = 0xFFFFFFFF

# enum values (reference include/embree3/rtcore_*.h)
RTC_GEOMETRY_TYPE_TRIANGLE = 0
RTC_GEOMETRY_TYPE_QUAD = 1
RTC_GEOMETRY_TYPE_INSTANCE = 121
RTC_BUFFER_TYPE_INDEX = 0
RTC_BUFFER_TYPE_VERTEX = 1
RTC_FORMAT_UINT3 = 0x5003
RTC_FORMAT_FLOAT3 = 0x9003
RTC_FORMAT_FLOAT4 = 0x9004
RTC_FORMAT_FLOAT3X4_ROW_MAJOR = 0x9134
RTC_FORMAT_FLOAT4X4_ROW_MAJOR = 0x9144
RTC_FORMAT_FLOAT3X4_COLUMN_MAJOR = 0x9234
RTC_FORMAT_FLOAT4X4_COLUMN_MAJOR = 0x9244
RTC_BUILD_QUALITY_LOW = 0
RTC_BUILD_QUALITY_MEDIUM = 1
RTC_BUILD_QUALITY_HIGH = 2
RTC_SCENE_FLAG_NONE = 0
RTC_SCENE_FLAG_ROBUST = 4
RTC_ERROR_NONE = 0
RTC_ERROR_UNKNOWN = 1
RTC_ERROR_INVALID_ARGUMENT = 2
RTC_ERROR_INVALID_OPERATION = 3
RTC_ERROR_OUT_OF_MEMORY = 4
RTC_ERROR_UNSUPPORTED_CPU = 5
RTC_ERROR_CANCELLED = 6
RTC_DEVICE_PROPERTY_VERSION = 0
RTC_DEVICE_PROPERTY_RAY_STREAM_SUPPORTED = 35
RTC_DEVICE_PROPERTY_TRIANGLE_GEOMETRY_SUPPORTED = 96

# RTCRayHit (80 B) and RTCRay (48 B), reference include/embree3/rtcore_ray.h:11-49
RAY_FIELDS = [("org_x", "<f4"), ("org_y", "<f4"), ("org_z", "<f4"), ("tnear", "<f4"),
              ("dir_x", "<f4"), ("dir_y", "<f4"), ("dir_z", "<f4"), ("time", "<f4"),
              ("tfar", "<f4"), ("mask", "<u4"), ("id", "<u4"), ("flags", "<u4")]
HIT_FIELDS = [("Ng_x", "<f4"), ("Ng_y", "<f4"), ("Ng_z", "<f4"), ("u", "<f4"), ("v", "<f4"),
              ("primID", "<u4"), ("geomID", "<u4"), ("instID", "<u4")]
RTCRay = np.dtype(RAY_FIELDS, align=True)
RTCRayHit = np.dtype(RAY_FIELDS + HIT_FIELDS, align=True)
assert RTCRay.itemsize == 48 and RTCRayHit.itemsize == 80


class RTCIntersectContext(ctypes.Structure):
    _fields_ = [("flags", ctypes.c_int), ("filter", ctypes.c_void_p), ("instID", ctypes.c_uint * 1)]


RTCBuildPrimitive = np.dtype([("lower_x", "<f4"), ("lower_y", "<f4"), ("lower_z", "<f4"), ("geomID", "<u4"),
                              ("upper_x", "<f4"), ("upper_y", "<f4"), ("upper_z", "<f4"), ("primID", "<u4")])
assert RTCBuildPrimitive.itemsize == 32

CREATE_NODE_FUNC = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p)
SET_NODE_CHILDREN_FUNC = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint,
                                          ctypes.c_void_p)
SET_NODE_BOUNDS_FUNC = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint,
                                        ctypes.c_void_p)
CREATE_LEAF_FUNC = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                    ctypes.c_void_p)
BUILD_PROGRESS_FUNC = ctypes.CFUNCTYPE(ctypes.c_bool, ctypes.c_void_p, ctypes.c_double)
SPLIT_PRIMITIVE_FUNC = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_uint, ctypes.c_float, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p)


class RTCBuildArguments(ctypes.Structure):
    _fields_ = [("byteSize", ctypes.c_size_t), ("buildQuality", ctypes.c_int), ("buildFlags", ctypes.c_int),
                ("maxBranchingFactor", ctypes.c_uint), ("maxDepth", ctypes.c_uint), ("sahBlockSize", ctypes.c_uint),
                ("minLeafSize", ctypes.c_uint), ("maxLeafSize", ctypes.c_uint), ("traversalCost", ctypes.c_float),
                ("intersectionCost", ctypes.c_float), ("bvh", ctypes.c_void_p), ("primitives", ctypes.c_void_p),
                ("primitiveCount", ctypes.c_size_t), ("primitiveArrayCapacity", ctypes.c_size_t),
                ("createNode", CREATE_NODE_FUNC), ("setNodeChildren", SET_NODE_CHILDREN_FUNC),
                ("setNodeBounds", SET_NODE_BOUNDS_FUNC), ("createLeaf", CREATE_LEAF_FUNC),
                ("splitPrimitive", ctypes.c_void_p), ("buildProgress", BUILD_PROGRESS_FUNC),
                ("userPtr", ctypes.c_void_p)]


def default_build_arguments():
    """rtcDefaultBuildArguments (reference rtcore_builder.h:82-108)."""
    a = RTCBuildArguments()
    a.byteSize = ctypes.sizeof(a)
    a.buildQuality = RTC_BUILD_QUALITY_MEDIUM
    a.maxBranchingFactor, a.maxDepth, a.sahBlockSize = 2, 32, 1
    a.minLeafSize, a.maxLeafSize = 1, 32
    a.traversalCost = a.intersectionCost = 1.0
    return a


class RTCBounds(ctypes.Structure):
    _fields_ = [("lower_x", ctypes.c_float), ("lower_y", ctypes.c_float), ("lower_z", ctypes.c_float),
                ("align0", ctypes.c_float), ("upper_x", ctypes.c_float), ("upper_y", ctypes.c_float),
                ("upper_z", ctypes.c_float), ("align1", ctypes.c_float)]


class RTCGSceneStats(ctypes.Structure):
    _fields_ = [("num_prims", ctypes.c_uint), ("num_nodes", ctypes.c_uint), ("max_depth", ctypes.c_uint),
                ("node_bytes", ctypes.c_uint), ("tri_bytes", ctypes.c_uint), ("max_stack", ctypes.c_uint),
                ("sah", ctypes.c_double), ("sah_tri", ctypes.c_double), ("build_ms", ctypes.c_double)]


ERROR_FUNC = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p)
MEMORY_MONITOR_FUNC = ctypes.CFUNCTYPE(ctypes.c_bool, ctypes.c_void_p, ctypes.c_ssize_t, ctypes.c_bool)
PROGRESS_FUNC = ctypes.CFUNCTYPE(ctypes.c_bool, ctypes.c_void_p, ctypes.c_double)

_vp, _u32, _sz, _i32, _f32 = ctypes.c_void_p, ctypes.c_uint, ctypes.c_size_t, ctypes.c_int, ctypes.c_float
# name: (restype, argtypes)
SIGNATURES = {
    "rtcNewDevice": (_vp, [ctypes.c_char_p]),
    "rtcRetainDevice": (None, [_vp]),
    "rtcReleaseDevice": (None, [_vp]),
    "rtcGetDeviceProperty": (ctypes.c_ssize_t, [_vp, _i32]),
    "rtcSetDeviceProperty": (None, [_vp, _i32, ctypes.c_ssize_t]),
    "rtcGetDeviceError": (_i32, [_vp]),
    "rtcSetDeviceErrorFunction": (None, [_vp, ERROR_FUNC, _vp]),
    "rtcSetDeviceMemoryMonitorFunction": (None, [_vp, MEMORY_MONITOR_FUNC, _vp]),
    "rtcNewBuffer": (_vp, [_vp, _sz]),
    "rtcNewSharedBuffer": (_vp, [_vp, _vp, _sz]),
    "rtcGetBufferData": (_vp, [_vp]),
    "rtcRetainBuffer": (None, [_vp]),
    "rtcReleaseBuffer": (None, [_vp]),
    "rtcNewGeometry": (_vp, [_vp, _i32]),
    "rtcRetainGeometry": (None, [_vp]),
    "rtcReleaseGeometry": (None, [_vp]),
    "rtcCommitGeometry": (None, [_vp]),
    "rtcEnableGeometry": (None, [_vp]),
    "rtcDisableGeometry": (None, [_vp]),
    "rtcSetGeometryTimeStepCount": (None, [_vp, _u32]),
    "rtcSetGeometryMask": (None, [_vp, _u32]),
    "rtcSetGeometryBuildQuality": (None, [_vp, _i32]),
    "rtcSetGeometryBuffer": (None, [_vp, _i32, _u32, _i32, _vp, _sz, _sz, _sz]),
    "rtcSetSharedGeometryBuffer": (None, [_vp, _i32, _u32, _i32, _vp, _sz, _sz, _sz]),
    "rtcSetNewGeometryBuffer": (_vp, [_vp, _i32, _u32, _i32, _sz, _sz]),
    "rtcGetGeometryBufferData": (_vp, [_vp, _i32, _u32]),
    "rtcUpdateGeometryBuffer": (None, [_vp, _i32, _u32]),
    "rtcSetGeometryIntersectFilterFunction": (None, [_vp, _vp]),
    "rtcSetGeometryOccludedFilterFunction": (None, [_vp, _vp]),
    "rtcSetGeometryUserData": (None, [_vp, _vp]),
    "rtcGetGeometryUserData": (_vp, [_vp]),
    "rtcSetGeometryInstancedScene": (None, [_vp, _vp]),
    "rtcNewBVH": (_vp, [_vp]),
    "rtcBuildBVH": (_vp, [_vp]),
    "rtcThreadLocalAlloc": (_vp, [_vp, _sz, _sz]),
    "rtcRetainBVH": (None, [_vp]),
    "rtcReleaseBVH": (None, [_vp]),
    "rtcSetGeometryTransform": (None, [_vp, _u32, _i32, _vp]),
    "rtcGetGeometryTransform": (None, [_vp, _f32, _i32, _vp]),
    "rtcNewScene": (_vp, [_vp]),
    "rtcGetSceneDevice": (_vp, [_vp]),
    "rtcRetainScene": (None, [_vp]),
    "rtcReleaseScene": (None, [_vp]),
    "rtcAttachGeometry": (_u32, [_vp, _vp]),
    "rtcAttachGeometryByID": (None, [_vp, _vp, _u32]),
    "rtcDetachGeometry": (None, [_vp, _u32]),
    "rtcGetGeometry": (_vp, [_vp, _u32]),
    "rtcCommitScene": (None, [_vp]),
    "rtcJoinCommitScene": (None, [_vp]),
    "rtcSetSceneProgressMonitorFunction": (None, [_vp, PROGRESS_FUNC, _vp]),
    "rtcSetSceneBuildQuality": (None, [_vp, _i32]),
    "rtcSetSceneFlags": (None, [_vp, _i32]),
    "rtcGetSceneFlags": (_i32, [_vp]),
    "rtcGetSceneBounds": (None, [_vp, ctypes.POINTER(RTCBounds)]),
    "rtcIntersect1": (None, [_vp, _vp, _vp]),
    "rtcIntersect1M": (None, [_vp, _vp, _vp, _u32, _sz]),
    "rtcOccluded1": (None, [_vp, _vp, _vp]),
    "rtcOccluded1M": (None, [_vp, _vp, _vp, _u32, _sz]),
    # extensions (include/embree3/rtcore_mi355x.h)
    "rtcgDeviceMalloc": (_vp, [_vp, _sz]),
    "rtcgDeviceFree": (None, [_vp, _vp]),
    "rtcgMemcpy": (None, [_vp, _vp, _vp, _sz]),
    "rtcgIntersect1MAsync": (None, [_vp, _vp, _vp, _u32, _sz, _vp]),
    "rtcgOccluded1MAsync": (None, [_vp, _vp, _vp, _u32, _sz, _vp]),
    "rtcgQuery1MNoAlloc": (None, [_vp, _vp, _vp, _u32, _sz, _i32, _vp]),
    "rtcgSynchronize": (None, [_vp]),
    "rtcgGetErrorString": (ctypes.c_char_p, []),
    "rtcgTimeTrace": (None, [_vp, _vp, _vp, _u32, _sz, _i32, _i32, _vp]),
    "rtcgTimeBuild": (None, [_vp, _i32, _vp]),
    "rtcgTraceCounted": (None, [_vp, _vp, _u32, _sz, _i32, _vp, _vp]),
    "rtcgGetSceneStats": (None, [_vp, ctypes.POINTER(RTCGSceneStats)]),
    "rtcgReadBVH": (None, [_vp, _vp, _vp]),
    "rtcgGenSphere": (_u32, [_u32, _f32, _f32, _f32, _f32, _vp, _vp, ctypes.POINTER(_u32)]),
    "rtcgGenPinholeRays": (None, [_u32, _u32, _f32, _f32, _f32, _f32, _vp]),
    "rtcgGenSoup": (None, [_u32, ctypes.c_ulonglong, _f32, _vp, _vp]),
    "rtcgGenRandomRays": (None, [_u32, ctypes.c_ulonglong, _f32, _f32, _f32, _vp]),
    "rtcgGenHeightfield": (_u32, [_u32, ctypes.c_ulonglong, _f32, _u32, _vp, _vp]),
    "rtcgGenCameraRays": (None, [_u32, _u32, _vp, _vp, _vp, _f32, _vp]),
    "rtcgGenBounceRays": (_u32, [_vp, _u32, ctypes.c_ulonglong, _f32, _vp]),
    "rtcgGenDisplacedSphere": (_u32, [_u32, ctypes.c_ulonglong, _f32, _vp, _vp, _vp]),
    "rtcgGenSphereRays": (None, [_u32, ctypes.c_ulonglong, _f32, _vp]),
    "rtcgAutotune": (_f32, [_vp, _vp, _u32, _sz, _i32, _i32, ctypes.c_char_p, _sz]),
    "rtcIntersect4": (None, [_vp, _vp, _vp, _vp]),
    "rtcIntersect8": (None, [_vp, _vp, _vp, _vp]),
    "rtcIntersect16": (None, [_vp, _vp, _vp, _vp]),
    "rtcOccluded4": (None, [_vp, _vp, _vp, _vp]),
    "rtcOccluded8": (None, [_vp, _vp, _vp, _vp]),
    "rtcOccluded16": (None, [_vp, _vp, _vp, _vp]),
    "rtcIntersect1Mp": (None, [_vp, _vp, _vp, _u32]),
    "rtcOccluded1Mp": (None, [_vp, _vp, _vp, _u32]),
    "rtcIntersectNM": (None, [_vp, _vp, _vp, _u32, _u32, _sz]),
    "rtcOccludedNM": (None, [_vp, _vp, _vp, _u32, _u32, _sz]),
    "rtcIntersectNp": (None, [_vp, _vp, _vp, _u32]),
    "rtcOccludedNp": (None, [_vp, _vp, _vp, _u32]),
    "rtcgCommitScenes": (None, [_vp, _u32]),
    "rtcgTraceBatch": (None, [_vp, _u32, _i32]),
    "rtcgAutotuneBatch": (_f32, [_vp, _u32, _i32, ctypes.c_char_p, _sz]),
    "rtcgNewBatchGraph": (_vp, [_vp, _u32, _i32]),
    "rtcgLaunchBatchGraph": (None, [_vp, _vp]),
    "rtcgReleaseBatchGraph": (None, [_vp]),
    "rtcgNewSplitModel": (_vp, [_vp, _u32]),
    "rtcgRetainSplitModel": (None, [_vp]),
    "rtcgReleaseSplitModel": (None, [_vp]),
    "rtcgSplitModelParamCount": (_u32, []),
    "rtcgSplitModelSetParams": (None, [_vp, _vp]),
    "rtcgSplitModelGetParams": (None, [_vp, _vp]),
    "rtcgSplitModelPredict": (None, [_vp, _vp, _u32, _vp]),
    "rtcgSplitModelGradient": (_f32, [_vp, _vp, _vp, _vp, _u32, _vp]),
    "rtcgSplitModelCollect": (None, [_vp, _vp, _vp, _u32, _sz]),
    "rtcgSplitModelCollectInstanced": (None, [_vp, _vp, _vp, _vp, _u32, _sz]),
    "rtcgSplitModelSampleCount": (_u32, [_vp]),
    "rtcgSplitModelGetSamples": (None, [_vp, _vp, _vp]),
    "rtcgSplitModelClearSamples": (None, [_vp]),
    "rtcgSplitModelTrain": (_f32, [_vp, _u32, _u32, _f32, _u32]),
    "rtcgSetSplitModelMinCount": (None, [_vp, _u32]),
    "rtcgSetSplitModelSahSlack": (None, [_vp, ctypes.c_float]),
    "rtcgSetSceneSplitModel": (None, [_vp, _vp]),
}

_lib = None


def lib():
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C embree-aarch64_amd`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def __getattr__(name):
    if name.startswith("rtc"):
        return getattr(lib(), name)
    raise AttributeError(name)


def new_rayhits(n):
    r = np.zeros(n, RTCRayHit)
    r["tfar"] = np.inf
    r["mask"] = 0xFFFFFFFF
    r["geomID"] = RTC_INVALID_GEOMETRY_ID
    r["primID"] = RTC_INVALID_GEOMETRY_ID
    r["instID"] = RTC_INVALID_GEOMETRY_ID
    return r


def init_context():
    """rtcInitIntersectContext (reference rtcore_common.h:209-224)."""
    c = RTCIntersectContext()
    c.flags = 0
    c.filter = None
    c.instID[0] = RTC_INVALID_GEOMETRY_ID
    return c


class RTCError(RuntimeError):
    def __init__(self, code, where, message=""):
        super().__init__(f"{where}: RTCError {code}: {message}")
        self.code = code
        self.message = message


def check(device, where=""):
    e = lib().rtcGetDeviceError(device)
    if e != RTC_ERROR_NONE:
        raise RTCError(e, where, (lib().rtcgGetErrorString() or b"").decode(errors="replace"))


# ---------------------------------------------------------------- helpers --

def gen_sphere(num_phi, center=(0.0, 0.0, 0.0), radius=1.0):
    L = lib()
    nv = ctypes.c_uint(0)
    nt = L.rtcgGenSphere(num_phi, center[0], center[1], center[2], radius, None, None, ctypes.byref(nv))
    v = np.zeros((nv.value, 3), np.float32)
    ix = np.zeros((nt, 3), np.uint32)
    L.rtcgGenSphere(num_phi, center[0], center[1], center[2], radius, v.ctypes.data, ix.ctypes.data, None)
    return v, ix


def gen_pinhole(width, height, org=(0.0, 0.0, 3.0), fov=45.0):
    r = np.zeros(width * height, RTCRayHit)
    lib().rtcgGenPinholeRays(width, height, org[0], org[1], org[2], fov, r.ctypes.data)
    return r


def gen_soup(n, seed=1, h=0.005):
    v = np.zeros((3 * n, 3), np.float32)
    ix = np.zeros((n, 3), np.uint32)
    lib().rtcgGenSoup(n, seed, h, v.ctypes.data, ix.ctypes.data)
    return v, ix


def gen_random_rays(n, seed=2, lo=0.0, hi=1.0, tfar=float("inf")):
    r = np.zeros(n, RTCRayHit)
    lib().rtcgGenRandomRays(n, seed, lo, hi, tfar, r.ctypes.data)
    return r


def gen_heightfield(n, seed=1, amplitude=0.2, octaves=6):
    nt = lib().rtcgGenHeightfield(n, seed, amplitude, octaves, None, None)
    v = np.zeros((n * n, 3), np.float32)
    ix = np.zeros((nt, 3), np.uint32)
    lib().rtcgGenHeightfield(n, seed, amplitude, octaves, v.ctypes.data, ix.ctypes.data)
    return v, ix


def gen_camera(width, height, org, target, up=(0.0, 0.0, 1.0), fov=45.0):
    r = np.zeros(width * height, RTCRayHit)
    f3 = lambda a: np.ascontiguousarray(a, np.float32)
    o, t, u = f3(org), f3(target), f3(up)
    lib().rtcgGenCameraRays(width, height, o.ctypes.data, t.ctypes.data, u.ctypes.data, fov, r.ctypes.data)
    return r


def gen_bounce(hits, seed=3, epsilon=1e-4):
    out = np.zeros(len(hits), RTCRayHit)
    hits = np.ascontiguousarray(hits)
    k = lib().rtcgGenBounceRays(hits.ctypes.data, len(hits), seed, epsilon, out.ctypes.data)
    return out[:k].copy()


def gen_bounce_n(hits, n, seed=3, epsilon=1e-4):
    """Exactly n cosine-hemisphere bounce rays from the hit points of `hits`:
    one per hit in order, then again from the first hit with the next seed,
    until n."""
    parts, k, c = [], 0, 0
    while k < n:
        b = gen_bounce(hits, seed=seed + c, epsilon=epsilon)
        if len(b) == 0:
            raise ValueError("no hit to bounce from")
        parts.append(b[: n - k])
        k += len(parts[-1])
        c += 1
    return np.concatenate(parts)


def gen_displaced_sphere(num_phi, seed=1, amplitude=0.1):
    nv = ctypes.c_uint(0)
    nt = lib().rtcgGenDisplacedSphere(num_phi, seed, amplitude, None, None, ctypes.byref(nv))
    v = np.zeros((nv.value, 3), np.float32)
    ix = np.zeros((nt, 3), np.uint32)
    lib().rtcgGenDisplacedSphere(num_phi, seed, amplitude, v.ctypes.data, ix.ctypes.data, None)
    return v, ix


def gen_sphere_rays(n, seed=2, radius=2.0):
    r = np.zeros(n, RTCRayHit)
    lib().rtcgGenSphereRays(n, seed, radius, r.ctypes.data)
    return r


def batch_workload(num_scenes=256, seed=7, tri_range=(1e3, 2e6), ray_range=(4e3, 1e6)):
    """Per-scene (numPhi, triangles, rays) of the batched-scenes workload:
    triangle and ray counts log-uniform in the given ranges; numPhi is the
    UV-sphere ring count whose 4*numPhi*(numPhi-1) triangles come closest."""
    rng = np.random.default_rng(seed)
    tris = np.exp(rng.uniform(np.log(tri_range[0]), np.log(tri_range[1]), num_scenes))
    rays = np.round(np.exp(rng.uniform(np.log(ray_range[0]), np.log(ray_range[1]), num_scenes))).astype(np.int64)
    phis = np.maximum(2, np.round(0.5 + np.sqrt(tris / 4.0))).astype(np.int64)
    return [(int(p), int(4 * p * (p - 1)), int(r)) for p, r in zip(phis, rays)]


def random_transforms(n, seed=5, scale=(0.5, 2.0), extent=50.0):
    """n local-to-world transforms (n x 12 floats, columns vx | vy | vz | p):
    a uniform random rotation (Shoemake's quaternion), a uniform scale in
    `scale` and a translation uniform in [-extent, extent]^3."""
    rng = np.random.default_rng(seed)
    u1, u2, u3 = rng.random((3, n))
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    qx, qy = a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2)
    qz, qw = b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (qy * qy + qz * qz); R[:, 0, 1] = 2 * (qx * qy - qz * qw); R[:, 0, 2] = 2 * (qx * qz + qy * qw)
    R[:, 1, 0] = 2 * (qx * qy + qz * qw); R[:, 1, 1] = 1 - 2 * (qx * qx + qz * qz); R[:, 1, 2] = 2 * (qy * qz - qx * qw)
    R[:, 2, 0] = 2 * (qx * qz - qy * qw); R[:, 2, 1] = 2 * (qy * qz + qx * qw); R[:, 2, 2] = 1 - 2 * (qx * qx + qy * qy)
    sc = rng.uniform(scale[0], scale[1], n)
    t = rng.uniform(-extent, extent, (n, 3))
    out = np.empty((n, 12), np.float32)
    for c in range(3):  # column c = R[:, :, c] * scale
        out[:, 3 * c:3 * c + 3] = R[:, :, c] * sc[:, None]
    out[:, 9:12] = t
    return out


def gen_shadow_rays(hits, light=(0.0, 100.0, 0.0), centre=(0.0, 0.0, 0.0), tnear=1e-4):
    """One occlusion record (RTCRay) per traced closest-hit record, towards
    the point light: from the hit point (org + tfar dir, float32) where the
    ray hit; where it missed, from the point of the ray nearest `centre` (a
    light sample inside the scene's volume, at t >= 0). Direction = light -
    origin, unnormalised, so the light sits at t = 1: tfar = 1."""
    hit = hits["geomID"] != RTC_INVALID_GEOMETRY_ID
    o = np.stack([hits["org_x"], hits["org_y"], hits["org_z"]], 1).astype(np.float32)
    d = np.stack([hits["dir_x"], hits["dir_y"], hits["dir_z"]], 1).astype(np.float32)
    c = np.asarray(centre, np.float32)
    t_near_c = np.maximum(np.float32(0), ((c - o) * d).sum(1, dtype=np.float32) / (d * d).sum(1, dtype=np.float32))
    t = np.where(hit, hits["tfar"], t_near_c).astype(np.float32)
    p = (o + t[:, None] * d).astype(np.float32)
    sh = np.zeros(len(hits), RTCRay)
    sh["org_x"], sh["org_y"], sh["org_z"] = p.T
    dl = (np.asarray(light, np.float32) - p).astype(np.float32)
    sh["dir_x"], sh["dir_y"], sh["dir_z"] = dl.T
    sh["tnear"], sh["tfar"], sh["mask"] = tnear, 1.0, 0xFFFFFFFF
    sh["id"] = np.arange(len(hits), dtype=np.uint32)
    return sh


def instance_workload(seed=11, sizes=None):
    """Base meshes of the instanced-scene workload: two displaced spheres and
    three heightfield terrain patches. sizes = (sphere numPhi, sphere numPhi,
    terrain n, terrain n, terrain n); default about 1M, 200k, 1M, 200k and
    100k triangles."""
    sizes = sizes or (500, 224, 708, 317, 225)
    meshes = [gen_displaced_sphere(sizes[0], seed=seed), gen_displaced_sphere(sizes[1], seed=seed + 1)]
    for k, n in enumerate(sizes[2:]):
        meshes.append(gen_heightfield(n, seed=seed + 2 + k))
    return meshes


class Device:
    def __init__(self, config=""):
        self.h = lib().rtcNewDevice(config.encode())
        if not self.h:
            raise RTCError(lib().rtcGetDeviceError(None), "rtcNewDevice")

    def check(self, where=""):
        check(self.h, where)

    def release(self):
        if self.h:
            lib().rtcReleaseDevice(self.h)
            self.h = None

    def malloc(self, nbytes):
        p = lib().rtcgDeviceMalloc(self.h, nbytes)
        self.check("rtcgDeviceMalloc")
        return p

    def free(self, p):
        lib().rtcgDeviceFree(self.h, p)

    def memcpy(self, dst, src, nbytes):
        lib().rtcgMemcpy(self.h, dst, src, nbytes)
        self.check("rtcgMemcpy")


class Scene:
    """A scene of triangle meshes; geomIDs follow attach order."""

    def __init__(self, device, meshes=(), quality=RTC_BUILD_QUALITY_MEDIUM, flags=0, shared=False):
        L = lib()
        self.device = device
        self.h = L.rtcNewScene(device.h)
        self.keep = []
        if quality != RTC_BUILD_QUALITY_MEDIUM:
            L.rtcSetSceneBuildQuality(self.h, quality)
        if flags:
            L.rtcSetSceneFlags(self.h, flags)
        for v, ix in meshes:
            self.add_mesh(v, ix, shared=shared)
        device.check("Scene")

    def add_mesh(self, v, ix, shared=False):
        L = lib()
        v = np.ascontiguousarray(v, np.float32).reshape(-1, 3)
        ix = np.ascontiguousarray(ix, np.uint32).reshape(-1, 3)
        g = L.rtcNewGeometry(self.device.h, RTC_GEOMETRY_TYPE_TRIANGLE)
        if shared:
            vp = np.zeros((v.shape[0] + 1, 3), np.float32)  # 4 readable bytes past the last vertex
            vp[:-1] = v
            self.keep += [vp, ix]
            L.rtcSetSharedGeometryBuffer(g, RTC_BUFFER_TYPE_VERTEX, 0, RTC_FORMAT_FLOAT3, vp.ctypes.data, 0, 12,
                                         v.shape[0])
            L.rtcSetSharedGeometryBuffer(g, RTC_BUFFER_TYPE_INDEX, 0, RTC_FORMAT_UINT3, ix.ctypes.data, 0, 12,
                                         ix.shape[0])
        else:
            pv = L.rtcSetNewGeometryBuffer(g, RTC_BUFFER_TYPE_VERTEX, 0, RTC_FORMAT_FLOAT3, 12, v.shape[0])
            pi = L.rtcSetNewGeometryBuffer(g, RTC_BUFFER_TYPE_INDEX, 0, RTC_FORMAT_UINT3, 12, ix.shape[0])
            if v.size:
                ctypes.memmove(pv, v.ctypes.data, v.nbytes)
            if ix.size:
                ctypes.memmove(pi, ix.ctypes.data, ix.nbytes)
        L.rtcCommitGeometry(g)
        gid = L.rtcAttachGeometry(self.h, g)
        L.rtcReleaseGeometry(g)
        self.device.check("add_mesh")
        return gid

    def add_instance(self, scene, l2w, enabled=True):
        """Attaches an instance of the committed `scene` with the local-to-world
        transform l2w (3x4: columns vx, vy, vz, p as the rows of a 4x3 array, or
        a 3x4 row-major matrix given as 12 floats in rows); returns its geomID."""
        L = lib()
        m = np.ascontiguousarray(l2w, np.float32).reshape(-1)
        if m.size != 12:
            raise ValueError("l2w needs 12 floats (FLOAT3X4 column major)")
        g = L.rtcNewGeometry(self.device.h, RTC_GEOMETRY_TYPE_INSTANCE)
        L.rtcSetGeometryInstancedScene(g, scene.h)
        L.rtcSetGeometryTransform(g, 0, RTC_FORMAT_FLOAT3X4_COLUMN_MAJOR, m.ctypes.data)
        if not enabled:
            L.rtcDisableGeometry(g)
        L.rtcCommitGeometry(g)
        gid = L.rtcAttachGeometry(self.h, g)
        L.rtcReleaseGeometry(g)
        self.device.check("add_instance")
        return gid

    def commit(self):
        lib().rtcCommitScene(self.h)
        self.device.check("rtcCommitScene")
        return self

    def stats(self):
        s = RTCGSceneStats()
        lib().rtcgGetSceneStats(self.h, ctypes.byref(s))
        return {k: getattr(s, k) for k, _ in RTCGSceneStats._fields_}

    def bounds(self):
        b = RTCBounds()
        lib().rtcGetSceneBounds(self.h, ctypes.byref(b))
        return b

    def intersect1M(self, rays, stride=None, ctx=None):
        ctx = ctx or init_context()
        lib().rtcIntersect1M(self.h, ctypes.byref(ctx), rays.ctypes.data, len(rays), stride or rays.itemsize)
        self.device.check("rtcIntersect1M")

    def occluded1M(self, rays, stride=None, ctx=None):
        ctx = ctx or init_context()
        lib().rtcOccluded1M(self.h, ctypes.byref(ctx), rays.ctypes.data, len(rays), stride or rays.itemsize)
        self.device.check("rtcOccluded1M")

    def read_bvh(self):
        s = self.stats()
        nodes = np.zeros(s["num_nodes"] * 80, np.uint8)
        tris = np.zeros(s["num_prims"] * 48, np.uint8)
        lib().rtcgReadBVH(self.h, nodes.ctypes.data, tris.ctypes.data)
        self.device.check("rtcgReadBVH")
        return nodes, tris

    def release(self):
        if self.h:
            lib().rtcReleaseScene(self.h)
            self.h = None


class SplitModel:
    """Learned split-cost model (FP16 MLP trained on the GPU) on the per-bin
    statistics of split candidates: FEATURES inputs per sample."""

    FEATURES = 12

    def __init__(self, device, seed=1):
        self.device = device
        self.h = lib().rtcgNewSplitModel(device.h, seed)
        device.check("rtcgNewSplitModel")

    def release(self):
        if self.h:
            lib().rtcgReleaseSplitModel(self.h)
            self.h = None

    @staticmethod
    def nparam():
        return int(lib().rtcgSplitModelParamCount())

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparam()
        lib().rtcgSplitModelSetParams(self.h, p.ctypes.data)
        self.device.check("set_params")

    def get_params(self):
        p = np.zeros(self.nparam(), np.float32)
        lib().rtcgSplitModelGetParams(self.h, p.ctypes.data)
        self.device.check("get_params")
        return p

    def predict(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.FEATURES)
        y = np.zeros(len(x), np.float32)
        lib().rtcgSplitModelPredict(self.h, x.ctypes.data, len(x), y.ctypes.data)
        self.device.check("predict")
        return y

    def gradient(self, x, t, w=None):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.FEATURES)
        t = np.ascontiguousarray(t, np.float32)
        w = np.ones(len(x), np.float32) if w is None else np.ascontiguousarray(w, np.float32)
        g = np.zeros(self.nparam(), np.float32)
        loss = lib().rtcgSplitModelGradient(self.h, x.ctypes.data, t.ctypes.data, w.ctypes.data, len(x), g.ctypes.data)
        self.device.check("gradient")
        return float(loss), g

    def collect(self, scene, d_rays, n, stride=80):
        lib().rtcgSplitModelCollect(self.h, scene.h, d_rays, n, stride)
        self.device.check("collect")

    def collect_instanced(self, scene, obj, d_rays, n, stride=80):
        """Samples for a two-level scene's top tree (obj None) or for the
        tree of an instanced scene obj."""
        lib().rtcgSplitModelCollectInstanced(self.h, scene.h, obj.h if obj is not None else None, d_rays, n, stride)
        self.device.check("collect_instanced")

    def samples(self):
        n = int(lib().rtcgSplitModelSampleCount(self.h))
        x = np.zeros((n, self.FEATURES), np.float32)
        t = np.zeros(n, np.float32)
        lib().rtcgSplitModelGetSamples(self.h, x.ctypes.data, t.ctypes.data)
        self.device.check("samples")
        return x, t

    def train(self, steps, batch=65536, lr=1e-3, seed=1):
        loss = lib().rtcgSplitModelTrain(self.h, steps, batch, lr, seed)
        self.device.check("train")
        return float(loss)


RTCGBatchItem = np.dtype([("scene", "<u8"), ("rays", "<u8"), ("numRays", "<u4"), ("byteStride", "<u4")])


def commit_scenes(scenes):
    """One batched build of all scenes (rtcgCommitScenes)."""
    arr = (ctypes.c_void_p * len(scenes))(*[s.h for s in scenes])
    lib().rtcgCommitScenes(arr, len(scenes))
    scenes[0].device.check("rtcgCommitScenes")


def batch_items(streams):
    """streams: (scene, device pointer, number of rays, stride) tuples."""
    it = np.zeros(len(streams), RTCGBatchItem)
    for i, (sc, ptr, n, stride) in enumerate(streams):
        it[i] = (sc.h or 0, ptr or 0, n, stride)
    return it


def build_user_bvh(device, prims, quality=RTC_BUILD_QUALITY_MEDIUM, branching=2, min_leaf=1, max_leaf=32,
                   block=1, traversal_cost=1.0, intersection_cost=1.0, progress=None, capacity=None, count=None,
                   split=None):
    """rtcBuildBVH with Python callbacks that record the tree the library
    builds over `prims` (an RTCBuildPrimitive array, reordered in place).
    Returns (root, nodes, leaves): nodes[id] = {"bounds": [(lo, hi)...],
    "children": [...]}, leaves[id] = (offset into prims, count); ids are the
    void* values the callbacks returned (nodes odd, leaves even). With
    `count`, only the first `count` entries are primitives and the whole
    array is the capacity; `split(prim, dim, pos)` (prim an RTCBuildPrimitive
    record) returns the bounds of the two sides ((lo, hi), (lo, hi)) as the
    splitPrimitive callback."""
    L = lib()
    bvh = L.rtcNewBVH(device.h)
    device.check("rtcNewBVH")
    nodes, leaves = {}, {}
    base = prims.ctypes.data
    counter = [0]

    def create_node(alloc, n, user):
        mem = L.rtcThreadLocalAlloc(alloc, 64, 16)
        counter[0] += 1
        nid = 2 * counter[0] + 1
        nodes[nid] = {"n": n, "bounds": None, "children": None, "mem": mem}
        return nid

    def set_bounds(node, bptr, n, user):
        out = []
        for k in range(n):
            b = ctypes.cast(bptr[k], ctypes.POINTER(RTCBounds)).contents
            out.append(((b.lower_x, b.lower_y, b.lower_z), (b.upper_x, b.upper_y, b.upper_z)))
        nodes[node]["bounds"] = out

    def set_children(node, cptr, n, user):
        nodes[node]["children"] = [cptr[k] for k in range(n)]

    def create_leaf(alloc, pptr, n, user):
        counter[0] += 1
        lid = 2 * counter[0]
        off = ((pptr or base) - base) // 32
        leaves[lid] = (off, n)
        return lid

    def split_prim(pp, dim, pos, lb, rb, user):
        rec = np.frombuffer(ctypes.string_at(pp, prims.itemsize), dtype=prims.dtype)[0]
        (l0, h0), (l1, h1) = split(rec, dim, pos)
        for ptr, (lo, hi) in ((lb, (l0, h0)), (rb, (l1, h1))):
            b = ctypes.cast(ptr, ctypes.POINTER(RTCBounds)).contents
            b.lower_x, b.lower_y, b.lower_z = lo
            b.upper_x, b.upper_y, b.upper_z = hi

    cbs = [CREATE_NODE_FUNC(create_node), SET_NODE_CHILDREN_FUNC(set_children), SET_NODE_BOUNDS_FUNC(set_bounds),
           CREATE_LEAF_FUNC(create_leaf), BUILD_PROGRESS_FUNC(progress or (lambda u, f: True))]
    split_cb = SPLIT_PRIMITIVE_FUNC(split_prim) if split else None
    a = default_build_arguments()
    a.buildQuality = quality
    a.maxBranchingFactor, a.minLeafSize, a.maxLeafSize, a.sahBlockSize = branching, min_leaf, max_leaf, block
    a.traversalCost, a.intersectionCost = traversal_cost, intersection_cost
    a.bvh = bvh
    a.primitives = base
    a.primitiveCount = len(prims) if count is None else count
    a.primitiveArrayCapacity = (len(prims) if count is None else len(prims)) if capacity is None else capacity
    a.createNode, a.setNodeChildren, a.setNodeBounds, a.createLeaf, a.buildProgress = cbs
    if split_cb:
        a.splitPrimitive = ctypes.cast(split_cb, ctypes.c_void_p)
    root = L.rtcBuildBVH(ctypes.byref(a))
    err = L.rtcGetDeviceError(device.h)
    L.rtcReleaseBVH(bvh)
    if err:
        raise RTCError(err, "rtcBuildBVH")
    return root, nodes, leaves
