"""The drop-in boundary: the C-ABI library loads, exports every entry point
its headers declare, and the public structs keep the reference's byte layout
(include/embree3/rtcore_ray.h:11-49, rtcore_common.h:157-206 of the reference).
No GPU work happens here."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import PKG, ROOT

INCLUDE = os.path.join(ROOT, "include")
HEADERS = [os.path.join(INCLUDE, "embree3", f) for f in sorted(os.listdir(os.path.join(INCLUDE, "embree3")))]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        for m in re.finditer(r"RTC_API\s+[\w\s\*]+?\b(rtc\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def exported_symbols(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], check=True, capture_output=True, text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_library_exports_every_declared_entry_point(rtc):
    decl = declared_functions()
    assert {"rtcNewDevice", "rtcCommitScene", "rtcIntersect1M", "rtcOccluded1M",
            "rtcgIntersect1MAsync"} <= decl
    exported = exported_symbols(rtc.LIB_PATH)
    assert decl - exported == set()
    # only rtc* symbols leave the library
    assert all(s.startswith("rtc") for s in exported)
    # the Python mirror binds every declared function
    assert decl <= set(rtc.SIGNATURES)


LAYOUT_PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include <embree3/rtcore.h>
#include <embree3/rtcore_mi355x.h>
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(struct RTCRay), sizeof(struct RTCHit), sizeof(struct RTCRayHit),
         sizeof(struct RTCIntersectContext), sizeof(struct RTCBounds));
  printf("%zu %zu %zu %zu %zu %zu %zu\n", offsetof(struct RTCRayHit, ray.tnear), offsetof(struct RTCRayHit, ray.tfar),
         offsetof(struct RTCRayHit, hit.Ng_x), offsetof(struct RTCRayHit, hit.u), offsetof(struct RTCRayHit, hit.primID),
         offsetof(struct RTCRayHit, hit.geomID), offsetof(struct RTCRayHit, hit.instID));
  printf("%zu %zu\n", offsetof(struct RTCIntersectContext, filter), offsetof(struct RTCIntersectContext, instID));
  printf("%d %d %d %d %d %d\n", RTC_FORMAT_UINT3, RTC_FORMAT_FLOAT3, RTC_GEOMETRY_TYPE_INSTANCE,
         RTC_DEVICE_PROPERTY_TRIANGLE_GEOMETRY_SUPPORTED, RTC_SCENE_FLAG_ROBUST, RTC_ERROR_CANCELLED);
  struct RTCIntersectContext ctx; rtcInitIntersectContext(&ctx);
  printf("%d %u\n", (int)ctx.flags, ctx.instID[0]);
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(struct RTCRayHit4), sizeof(struct RTCRayHit8), sizeof(struct RTCRayHit16),
         offsetof(struct RTCRayHit4, hit), offsetof(struct RTCRayHit8, hit.u), offsetof(struct RTCRayHit16, hit.instID));
  printf("%zu %zu\n", sizeof(struct RTCRayHitNp), offsetof(struct RTCRayHitNp, hit.geomID));
  printf("%zu %zu %zu %zu\n", sizeof(struct RTCBuildPrimitive), offsetof(struct RTCBuildPrimitive, geomID),
         offsetof(struct RTCBuildPrimitive, upper_x), offsetof(struct RTCBuildPrimitive, primID));
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(struct RTCBuildArguments),
         offsetof(struct RTCBuildArguments, buildQuality), offsetof(struct RTCBuildArguments, maxLeafSize),
         offsetof(struct RTCBuildArguments, intersectionCost), offsetof(struct RTCBuildArguments, bvh),
         offsetof(struct RTCBuildArguments, primitiveArrayCapacity), offsetof(struct RTCBuildArguments, createNode),
         offsetof(struct RTCBuildArguments, userPtr));
  struct RTCBuildArguments a = rtcDefaultBuildArguments();
  printf("%zu %d %u %u %u %u %u %g %g\n", a.byteSize, (int)a.buildQuality, a.maxBranchingFactor, a.maxDepth,
         a.sahBlockSize, a.minLeafSize, a.maxLeafSize, a.traversalCost, a.intersectionCost);
  return 0;
}
"""


@pytest.mark.parametrize("compiler", ["gcc", "g++"])
def test_struct_layout_and_enums(tmp_path, compiler):
    src = tmp_path / ("probe.c" if compiler == "gcc" else "probe.cpp")
    src.write_text(LAYOUT_PROBE)
    exe = tmp_path / "probe"
    subprocess.run([compiler, "-I", INCLUDE, str(src), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    assert lines[0].split() == ["48", "32", "80", "24", "32"]
    assert lines[1].split() == ["12", "32", "48", "60", "68", "72", "76"]
    assert lines[2].split() == ["8", "16"]
    assert lines[3].split() == [str(0x5003), str(0x9003), "121", "96", "4", "6"]
    assert lines[4].split() == ["0", str(0xFFFFFFFF)]
    # packets: 12 ray + 8 hit fields of K lanes each; pointer SoA: 20 pointers
    assert lines[5].split() == ["320", "640", "1280", "192", str(12 * 32 + 3 * 32), str(12 * 64 + 7 * 64)]
    assert lines[6].split() == ["160", str(8 * 18)]
    # user builder (reference rtcore_builder.h:14-108)
    assert lines[7].split() == ["32", "12", "16", "28"]
    assert lines[8].split() == ["136", "8", "32", "40", "48", "72", "80", "128"]
    assert lines[9].split() == ["136", "1", "2", "32", "1", "1", "32", "1", "1"]


def test_c_program_links_against_library(tmp_path, rtc):
    """A C caller builds and links with -lembree3 exactly as against the reference."""
    src = tmp_path / "app.c"
    src.write_text("#include <embree3/rtcore.h>\nint main(void){ RTCDevice d = rtcNewDevice(\"verbose=0\");"
                   " if (d) rtcReleaseDevice(d); return (int)rtcGetDeviceError(NULL) == 99; }\n")
    libdir = os.path.dirname(rtc.LIB_PATH)
    subprocess.run(["gcc", "-I", INCLUDE, str(src), "-L", libdir, "-l:libembree3.so", "-o", str(tmp_path / "app")],
                   check=True)
    # the program needs the library by its soname, libembree3.so.3, as a
    # program linked against the reference does
    env = dict(os.environ, LD_LIBRARY_PATH=libdir)
    subprocess.run([str(tmp_path / "app")], check=True, env=env, timeout=120)


def test_integration_example_compiles_as_c(tmp_path, rtc):
    """The caller program INTEGRATION.md shows is strict C (gcc -Wall -Werror)
    against these headers; the GPU test runs it."""
    import re
    text = open(os.path.join(os.path.dirname(INCLUDE), "INTEGRATION.md")).read()
    m = re.search(r"```c example-app\n(.*?)```", text, re.S)
    assert m, "INTEGRATION.md has no example-app block"
    src = tmp_path / "app.c"
    src.write_text(m.group(1))
    libdir = os.path.dirname(rtc.LIB_PATH)
    subprocess.run(["gcc", "-std=c11", "-O2", "-Wall", "-Werror", "-I", INCLUDE, str(src), "-L", libdir,
                    "-l:libembree3.so", "-o", str(tmp_path / "app")], check=True)


def _gpu_visible():
    return os.path.exists("/dev/kfd") and any(n.startswith("renderD") for n in os.listdir("/dev/dri")) \
        if os.path.isdir("/dev/dri") else False


@pytest.mark.skipif(_gpu_visible(), reason="checks the no-GPU error path")
def test_new_device_without_gpu_reports_error(rtc):
    L = rtc.lib()
    d = L.rtcNewDevice(b"")
    assert not d
    assert L.rtcGetDeviceError(None) == rtc.RTC_ERROR_UNSUPPORTED_CPU
    assert L.rtcGetDeviceError(None) == rtc.RTC_ERROR_NONE  # read clears


def test_null_handles_report_invalid_argument(rtc):
    L = rtc.lib()
    L.rtcCommitScene(None)
    assert L.rtcGetDeviceError(None) == rtc.RTC_ERROR_INVALID_ARGUMENT
    L.rtcIntersect1M(None, None, None, 0, 80)
    assert L.rtcGetDeviceError(None) == rtc.RTC_ERROR_INVALID_ARGUMENT
    assert L.rtcGetDeviceError(None) == rtc.RTC_ERROR_NONE
