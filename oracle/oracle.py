"""ctypes binding of the CPU restatement (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker. The product never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")


class OracleMesh(ctypes.Structure):
    _fields_ = [("vertices", ctypes.c_void_p), ("num_vertices", ctypes.c_uint),
                ("indices", ctypes.c_void_p), ("num_triangles", ctypes.c_uint)]


class OracleObject(ctypes.Structure):
    _fields_ = [("meshes", ctypes.c_void_p), ("nmeshes", ctypes.c_uint), ("robust", ctypes.c_int)]


class OracleInstance(ctypes.Structure):
    _fields_ = [("w2l", ctypes.c_float * 12), ("object", ctypes.c_uint), ("instID", ctypes.c_uint),
                ("identity", ctypes.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        vp, u32, sz, i32 = ctypes.c_void_p, ctypes.c_uint, ctypes.c_size_t, ctypes.c_int
        fp = ctypes.POINTER(ctypes.c_float)
        L.oracle_moeller.argtypes = [fp, fp, ctypes.c_float, ctypes.c_float, fp, fp, fp, fp]
        L.oracle_moeller.restype = i32
        L.oracle_moeller_verts.argtypes = [fp, fp, ctypes.c_float, ctypes.c_float, fp, fp, fp, fp]
        L.oracle_moeller_verts.restype = i32
        L.oracle_pluecker.argtypes = [fp, fp, ctypes.c_float, ctypes.c_float, fp, fp, fp, fp]
        L.oracle_pluecker.restype = i32
        L.oracle_intersect_brute.argtypes = [vp, u32, vp, u32, sz, u32, i32]
        L.oracle_intersect_brute_robust.argtypes = [vp, u32, vp, u32, sz, u32, i32]
        L.oracle_occluded_brute_robust.argtypes = [vp, u32, vp, u32, sz, i32]
        L.oracle_intersect_bvh_robust.argtypes = [vp, vp, u32, sz, u32, i32]
        L.oracle_occluded_bvh_robust.argtypes = [vp, vp, u32, sz, i32]
        L.oracle_occluded_brute.argtypes = [vp, u32, vp, u32, sz, i32]
        L.oracle_build_sah.argtypes = [vp, u32]
        L.oracle_build_sah.restype = vp
        L.oracle_build_sah_threads.argtypes = [vp, u32, i32]
        L.oracle_build_sah_threads.restype = vp
        L.oracle_bvh_free.argtypes = [vp]
        L.oracle_bvh_stats.argtypes = [vp, vp]
        L.oracle_bvh_validate.argtypes = [vp]
        L.oracle_bvh_validate.restype = i32
        L.oracle_intersect_bvh.argtypes = [vp, vp, u32, sz, u32, i32]
        L.oracle_occluded_bvh.argtypes = [vp, vp, u32, sz, i32]
        L.oracle_intersect_bvh_counted.argtypes = [vp, vp, u32, sz, vp, vp, i32]
        L.oracle_bit_interleave.argtypes = [u32, u32, u32]
        L.oracle_bit_interleave.restype = u32
        L.oracle_morton_codes.argtypes = [vp, vp, u32, vp]
        L.oracle_half_area.argtypes = [fp, fp]
        L.oracle_half_area.restype = ctypes.c_float
        L.oracle_binned_split.argtypes = [vp, vp, u32, ctypes.POINTER(u32), ctypes.POINTER(ctypes.c_float),
                                          ctypes.POINTER(u32)]
        L.oracle_binned_split.restype = i32
        L.oracle_affine_inverse.argtypes = [fp, fp]
        L.oracle_xfm_point.argtypes = [fp, fp, fp]
        L.oracle_xfm_vector.argtypes = [fp, fp, fp]
        L.oracle_intersect_instances.argtypes = [vp, vp, u32, vp, u32, sz, u32, i32]
        L.oracle_occluded_instances.argtypes = [vp, vp, u32, vp, u32, sz, u32, i32]
        L.oracle_intersect_instances_bvh.argtypes = [vp, vp, vp, u32, vp, u32, sz, u32, i32]
        L.oracle_occluded_instances_bvh.argtypes = [vp, vp, vp, u32, vp, u32, sz, u32, i32]
        _lib = L
    return _lib


def _f3(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def moeller(org, dir, tnear, tfar, v0, e1, e2):
    """Returns (valid, [T, absDen, U, V, Ngx, Ngy, Ngz, t])."""
    o, po = _f3(org); d, pd = _f3(dir); a, pa = _f3(v0); b, pb = _f3(e1); c, pc = _f3(e2)
    out = np.zeros(8, np.float32)
    v = lib().oracle_moeller(po, pd, float(tnear), float(tfar), pa, pb, pc,
                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return bool(v), out


def moeller_verts(org, dir, tnear, tfar, p0, p1, p2):
    o, po = _f3(org); d, pd = _f3(dir); a, pa = _f3(p0); b, pb = _f3(p1); c, pc = _f3(p2)
    out = np.zeros(8, np.float32)
    v = lib().oracle_moeller_verts(po, pd, float(tnear), float(tfar), pa, pb, pc,
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return bool(v), out


def pluecker(org, dir, tnear, tfar, p0, p1, p2):
    """Robust test; returns (valid, [t, U, V, UVW, Ngx, Ngy, Ngz, t or inf])."""
    o, po = _f3(org); d, pd = _f3(dir); a, pa = _f3(p0); b, pb = _f3(p1); c, pc = _f3(p2)
    out = np.zeros(8, np.float32)
    v = lib().oracle_pluecker(po, pd, float(tnear), float(tfar), pa, pb, pc,
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return bool(v), out


class Meshes:
    """Keeps numpy arrays alive for the meshes passed to the oracle."""

    def __init__(self, meshes):
        self.arrays = []
        arr = (OracleMesh * max(1, len(meshes)))()
        for i, (v, ix) in enumerate(meshes):
            v = np.ascontiguousarray(v, dtype=np.float32).reshape(-1, 3)
            ix = np.ascontiguousarray(ix, dtype=np.uint32).reshape(-1, 3)
            self.arrays += [v, ix]
            arr[i].vertices = v.ctypes.data
            arr[i].num_vertices = v.shape[0]
            arr[i].indices = ix.ctypes.data
            arr[i].num_triangles = ix.shape[0]
        self.c = arr
        self.n = len(meshes)


def intersect_brute(meshes, rayhits, inst_id=0xFFFFFFFF, nthreads=0, robust=False):
    m = Meshes(meshes)
    f = lib().oracle_intersect_brute_robust if robust else lib().oracle_intersect_brute
    f(ctypes.addressof(m.c), m.n, rayhits.ctypes.data, len(rayhits), rayhits.itemsize, inst_id, nthreads)


def occluded_brute(meshes, rays, nthreads=0, robust=False):
    m = Meshes(meshes)
    f = lib().oracle_occluded_brute_robust if robust else lib().oracle_occluded_brute
    f(ctypes.addressof(m.c), m.n, rays.ctypes.data, len(rays), rays.itemsize, nthreads)


class BVH:
    """Reference-style binned-SAH BVH8 built by the CPU restatement."""

    def __init__(self, meshes, nthreads=0):
        """nthreads: build threads (0: all); the tree does not depend on it."""
        self.m = Meshes(meshes)
        self.h = lib().oracle_build_sah_threads(ctypes.addressof(self.m.c), self.m.n, nthreads)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_bvh_free(self.h)
            self.h = None

    def stats(self):
        out = np.zeros(8, np.float64)
        lib().oracle_bvh_stats(self.h, out.ctypes.data)
        return {"inner_nodes": int(out[0]), "leaves": int(out[1]), "prims": int(out[2]),
                "max_depth": int(out[3]), "sah": float(out[4]), "root_half_area": float(out[5]),
                "sah_per_tri": float(out[6])}

    def validate(self):
        return lib().oracle_bvh_validate(self.h)

    def intersect(self, rayhits, inst_id=0xFFFFFFFF, nthreads=0, robust=False):
        f = lib().oracle_intersect_bvh_robust if robust else lib().oracle_intersect_bvh
        f(self.h, rayhits.ctypes.data, len(rayhits), rayhits.itemsize, inst_id, nthreads)

    def occluded(self, rays, nthreads=0, robust=False):
        f = lib().oracle_occluded_bvh_robust if robust else lib().oracle_occluded_bvh
        f(self.h, rays.ctypes.data, len(rays), rays.itemsize, nthreads)

    def intersect_counted(self, rayhits, nthreads=0):
        n = np.zeros(len(rayhits), np.uint32)
        t = np.zeros(len(rayhits), np.uint32)
        lib().oracle_intersect_bvh_counted(self.h, rayhits.ctypes.data, len(rayhits), rayhits.itemsize,
                                           n.ctypes.data, t.ctypes.data, nthreads)
        return n, t


def bit_interleave(x, y, z):
    return lib().oracle_bit_interleave(x, y, z)


def morton_codes(lo, hi):
    lo = np.ascontiguousarray(lo, np.float32)
    hi = np.ascontiguousarray(hi, np.float32)
    codes = np.zeros(lo.shape[0], np.uint32)
    lib().oracle_morton_codes(lo.ctypes.data, hi.ctypes.data, lo.shape[0], codes.ctypes.data)
    return codes


def half_area(lo, hi):
    a, pa = _f3(lo); b, pb = _f3(hi)
    return lib().oracle_half_area(pa, pb)


def binned_split(lo, hi):
    lo = np.ascontiguousarray(lo, np.float32)
    hi = np.ascontiguousarray(hi, np.float32)
    pos = ctypes.c_uint(0); sah = ctypes.c_float(0); nb = ctypes.c_uint(0)
    dim = lib().oracle_binned_split(lo.ctypes.data, hi.ctypes.data, lo.shape[0], ctypes.byref(pos),
                                    ctypes.byref(sah), ctypes.byref(nb))
    return dim, pos.value, sah.value, nb.value


def affine_inverse(l2w):
    """world2local of a local-to-world transform (12 floats, columns vx|vy|vz|p)."""
    a, pa = _f3(np.asarray(l2w, np.float32).reshape(12))
    out = np.zeros(12, np.float32)
    lib().oracle_affine_inverse(pa, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return out


def xfm_point(m, p):
    a, pa = _f3(np.asarray(m, np.float32).reshape(12)); b, pb = _f3(p)
    out = np.zeros(3, np.float32)
    lib().oracle_xfm_point(pa, pb, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return out


def xfm_vector(m, v):
    a, pa = _f3(np.asarray(m, np.float32).reshape(12)); b, pb = _f3(v)
    out = np.zeros(3, np.float32)
    lib().oracle_xfm_vector(pa, pb, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return out


class Instances:
    """Two-level scene for the instance restatement: objects are lists of
    (vertices, indices) meshes (geomID = list index) with a robust flag;
    instances are (object, l2w, instID) or the scene's own triangles as an
    identity instance."""

    def __init__(self):
        self.objects, self.insts, self._keep = [], [], []
        self._bvhs = None

    def __del__(self):
        if self._bvhs is not None and lib is not None:
            for b in self._bvhs:
                lib().oracle_bvh_free(b)

    def add_object(self, meshes, robust=False):
        ms = (OracleMesh * max(1, len(meshes)))()
        for k, (v, ix) in enumerate(meshes):
            v = np.ascontiguousarray(v, np.float32)
            ix = np.ascontiguousarray(ix, np.uint32)
            self._keep += [v, ix]
            ms[k] = OracleMesh(v.ctypes.data, len(v), ix.ctypes.data, len(ix))
        self._keep.append(ms)
        self.objects.append(OracleObject(ctypes.cast(ms, ctypes.c_void_p), len(meshes), int(robust)))
        return len(self.objects) - 1

    def add_instance(self, obj, l2w, inst_id):
        i = OracleInstance()
        i.w2l[:] = list(affine_inverse(l2w))
        i.object, i.instID, i.identity = obj, inst_id, 0
        self.insts.append(i)

    def add_identity(self, obj):
        i = OracleInstance()
        i.w2l[:] = [1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0]
        i.object, i.instID, i.identity = obj, 0xFFFFFFFF, 1
        self.insts.append(i)

    def _arrays(self):
        objs = (OracleObject * len(self.objects))(*self.objects)
        insts = (OracleInstance * max(1, len(self.insts)))(*self.insts)
        return objs, insts

    def intersect(self, rays, ctx_inst=0xFFFFFFFF, nthreads=0):
        objs, insts = self._arrays()
        lib().oracle_intersect_instances(ctypes.cast(objs, ctypes.c_void_p), ctypes.cast(insts, ctypes.c_void_p),
                                         len(self.insts), rays.ctypes.data, len(rays), rays.itemsize, ctx_inst,
                                         nthreads)

    def occluded(self, rays, ctx_inst=0xFFFFFFFF, nthreads=0):
        objs, insts = self._arrays()
        lib().oracle_occluded_instances(ctypes.cast(objs, ctypes.c_void_p), ctypes.cast(insts, ctypes.c_void_p),
                                        len(self.insts), rays.ctypes.data, len(rays), rays.itemsize, ctx_inst,
                                        nthreads)

    def _bvh_arrays(self):
        if self._bvhs is None:
            self._bvhs = [lib().oracle_build_sah(o.meshes, o.nmeshes) for o in self.objects]
        bv = (ctypes.c_void_p * len(self._bvhs))(*self._bvhs)
        rb = (ctypes.c_int * len(self.objects))(*[o.robust for o in self.objects])
        return bv, rb

    def intersect_bvh(self, rays, ctx_inst=0xFFFFFFFF, nthreads=0):
        """As intersect(), over per-object oracle BVHs (large scenes)."""
        bv, rb = self._bvh_arrays()
        _, insts = self._arrays()
        lib().oracle_intersect_instances_bvh(ctypes.cast(bv, ctypes.c_void_p), ctypes.cast(rb, ctypes.c_void_p),
                                             ctypes.cast(insts, ctypes.c_void_p), len(self.insts), rays.ctypes.data,
                                             len(rays), rays.itemsize, ctx_inst, nthreads)

    def occluded_bvh(self, rays, ctx_inst=0xFFFFFFFF, nthreads=0):
        bv, rb = self._bvh_arrays()
        _, insts = self._arrays()
        lib().oracle_occluded_instances_bvh(ctypes.cast(bv, ctypes.c_void_p), ctypes.cast(rb, ctypes.c_void_p),
                                            ctypes.cast(insts, ctypes.c_void_p), len(self.insts), rays.ctypes.data,
                                            len(rays), rays.itemsize, ctx_inst, nthreads)
