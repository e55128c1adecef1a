"""rtcgQuery1MNoAlloc, the query entry point that allocates nothing (its
kernels' ray counter and error bits live in a fixed array per GPU), against
the entry points it is a variant of, rtcgIntersect1MAsync and
rtcgOccluded1MAsync: byte-identical records on two calls in a row, the second
with fewer rays, on the null stream (synchronous) and on a caller's stream,
for a one-level and a two-level scene; the one-level records also equal the
oracle's. Host records and more than 2^31 rays are refused on the host."""
import ctypes

import numpy as np
import pytest

from conftest import mismatches

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    bad = mismatches(a, b)
    assert len(bad) == 0, f"{what}: {len(bad)} of {len(a)} records differ, first {bad[:5]}"


def _hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return h


def _rays(rtc, n, seed, occluded):
    r = rtc.gen_random_rays(n, seed=seed)
    r["tfar"][::7] = 0.05
    r["tnear"][3::29] = 2.0  # tnear > tfar: left untouched
    if not occluded:
        return r
    o = np.zeros(n, rtc.RTCRay)
    for k in rtc.RTCRay.names:
        o[k] = r[k]
    return o


def _existing(rtc, dev, sc, rays, occluded):
    L = rtc.lib()
    d = dev.malloc(rays.nbytes)
    try:
        dev.memcpy(d, rays.ctypes.data, rays.nbytes)
        ctx = rtc.init_context()
        f = L.rtcgOccluded1MAsync if occluded else L.rtcgIntersect1MAsync
        f(sc.h, ctypes.byref(ctx), d, len(rays), rays.itemsize, None)
        L.rtcgSynchronize(dev.h)
        dev.check("existing entry point")
        out = np.empty_like(rays)
        dev.memcpy(out.ctypes.data, d, rays.nbytes)
        return out
    finally:
        dev.free(d)


def _no_alloc(rtc, dev, sc, rays, occluded, stream=None):
    L = rtc.lib()
    d = dev.malloc(rays.nbytes)
    try:
        dev.memcpy(d, rays.ctypes.data, rays.nbytes)
        ctx = rtc.init_context()
        L.rtcgQuery1MNoAlloc(sc.h, ctypes.byref(ctx), d, len(rays), rays.itemsize, int(occluded), stream)
        dev.check("rtcgQuery1MNoAlloc")
        if stream is not None:
            assert _hip().hipStreamSynchronize(stream) == 0
        out = np.empty_like(rays)
        dev.memcpy(out.ctypes.data, d, rays.nbytes)
        return out
    finally:
        dev.free(d)


def _two_calls(rtc, dev, sc, occluded, stream, oracle_bvh=None):
    for n, seed in ((300000, 901), (5000, 902)):
        rays = _rays(rtc, n, seed + 10 * int(occluded), occluded)
        want = _existing(rtc, dev, sc, rays, occluded)
        got = _no_alloc(rtc, dev, sc, rays, occluded, stream)
        _same(got, want, f"{n} rays, occluded={occluded}, stream={'caller' if stream else 'null'}")
        if oracle_bvh is not None:
            ref = rays.copy()
            (oracle_bvh.occluded if occluded else oracle_bvh.intersect)(ref)
            _same(got, ref, f"{n} rays against the oracle")


@pytest.mark.parametrize("occluded", [False, True])
def test_no_alloc_matches_async_one_level(rtc, oracle, occluded):
    dev = rtc.Device("")
    try:
        v, ix = rtc.gen_soup(60000, seed=903, h=0.01)
        sc = rtc.Scene(dev, [(v, ix)]).commit()
        bvh = oracle.BVH([(v, ix)])
        _two_calls(rtc, dev, sc, occluded, None, bvh)
        hip = _hip()
        s = ctypes.c_void_p()
        assert hip.hipStreamCreate(ctypes.byref(s)) == 0
        try:
            _two_calls(rtc, dev, sc, occluded, s, bvh)
        finally:
            hip.hipStreamDestroy(s)
        sc.release()
    finally:
        dev.release()


@pytest.mark.parametrize("occluded", [False, True])
def test_no_alloc_matches_async_two_level(rtc, occluded):
    dev = rtc.Device("")
    try:
        v, ix = rtc.gen_sphere(48, center=(0.5, 0.5, 0.5), radius=0.2)
        obj = rtc.Scene(dev, [(v, ix)]).commit()
        T = rtc.random_transforms(40, seed=904, scale=(0.5, 1.5), extent=0.6)
        top = rtc.Scene(dev)
        for k in range(len(T)):
            top.add_instance(obj, T[k])
        top.commit()
        _two_calls(rtc, dev, top, occluded, None)
        top.release()
        obj.release()
    finally:
        dev.release()


def test_no_alloc_refuses_host_records_and_oversize(rtc):
    dev = rtc.Device("")
    try:
        L = rtc.lib()
        v, ix = rtc.gen_soup(1000, seed=905)
        sc = rtc.Scene(dev, [(v, ix)]).commit()
        rays = _rays(rtc, 64, 906, False)
        ctx = rtc.init_context()
        L.rtcgQuery1MNoAlloc(sc.h, ctypes.byref(ctx), rays.ctypes.data, len(rays), 80, 0, None)
        assert L.rtcGetDeviceError(dev.h) == rtc.RTC_ERROR_INVALID_ARGUMENT
        d = dev.malloc(rays.nbytes)
        L.rtcgQuery1MNoAlloc(sc.h, ctypes.byref(ctx), d, (1 << 31) + 1, 80, 0, None)
        assert L.rtcGetDeviceError(dev.h) == rtc.RTC_ERROR_INVALID_ARGUMENT
        assert b"2^31" in L.rtcgGetErrorString()
        dev.free(d)
        sc.release()
    finally:
        dev.release()
