"""Pins the CPU restatement (oracle/) before it is trusted as the checker.

Each test checks the oracle against something it does not compute itself:
the reference's recorded answers (tests/golden/reference_probe.json), the
reference's own known-answer and watertightness tests
(tutorials/verify/verify.cpp:2339-2424, :2898-2979), closed forms, brute
force, or structural invariants of the reference's builder
(kernels/builders/bvh_builder_sah.h:222-330).
"""
import math

import numpy as np
import pytest

from conftest import mismatches, records_equal

RAYHIT = np.dtype([("org_x", "<f4"), ("org_y", "<f4"), ("org_z", "<f4"), ("tnear", "<f4"),
                   ("dir_x", "<f4"), ("dir_y", "<f4"), ("dir_z", "<f4"), ("time", "<f4"),
                   ("tfar", "<f4"), ("mask", "<u4"), ("id", "<u4"), ("flags", "<u4"),
                   ("Ng_x", "<f4"), ("Ng_y", "<f4"), ("Ng_z", "<f4"), ("u", "<f4"), ("v", "<f4"),
                   ("primID", "<u4"), ("geomID", "<u4"), ("instID", "<u4")])
INV = 0xFFFFFFFF


def rays(org, dirs, tnear=0.0, tfar=np.inf):
    org = np.broadcast_to(np.asarray(org, np.float32), (len(dirs), 3))
    r = np.zeros(len(dirs), RAYHIT)
    r["org_x"], r["org_y"], r["org_z"] = org[:, 0], org[:, 1], org[:, 2]
    d = np.asarray(dirs, np.float32)
    r["dir_x"], r["dir_y"], r["dir_z"] = d[:, 0], d[:, 1], d[:, 2]
    r["tnear"] = tnear
    r["tfar"] = tfar
    r["mask"] = INV
    r["geomID"] = r["primID"] = r["instID"] = INV
    return r


def soup(n, seed, h=0.02):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 1, 3), dtype=np.float32)
    v = (c + (rng.random((n, 3, 3), dtype=np.float32) * 2 - 1) * h).reshape(-1, 3).astype(np.float32)
    return v, np.arange(3 * n, dtype=np.uint32).reshape(n, 3)


def random_rays(n, seed, lo=0.0, hi=1.0, tfar=np.inf):
    rng = np.random.default_rng(seed)
    o = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    r = rays(o, d, 0.0, tfar)
    return r


def uv_sphere(num_phi):
    """UV sphere with the reference tutorial topology (geometry_creation.cpp:121-175)."""
    nt = 2 * num_phi
    phi = np.arange(num_phi + 1, dtype=np.float32)[:, None]
    th = np.arange(nt, dtype=np.float32)[None, :]
    pf = phi * np.float32(np.pi) / np.float32(num_phi)
    tf = th * np.float32(2.0) * np.float32(np.pi) / np.float32(nt)
    v = np.stack([np.sin(pf) * np.sin(tf), np.cos(pf) * np.ones_like(tf), np.sin(pf) * np.cos(tf)], -1)
    v = v.reshape(-1, 3).astype(np.float32)
    tris = []
    for p in range(1, num_phi + 1):
        for t in range(1, nt + 1):
            p00 = (p - 1) * nt + t - 1
            p01 = (p - 1) * nt + t % nt
            p10 = p * nt + t - 1
            p11 = p * nt + t % nt
            if p == 1:
                tris.append((p10, nt - 1, p11))
            elif p == num_phi:
                tris.append((num_phi * nt, p00, p01))
            else:
                tris.append((p10, p00, p11))
                tris.append((p01, p11, p00))
    return v, np.array(tris, np.uint32)


# ------------------------------------------------ reference recorded answers --

def test_single_triangle_matches_reference_answer(oracle, golden_probe):
    g = golden_probe["single_triangle"]
    v = np.array(g["vertices"], np.float32)
    ix = np.array(g["indices"], np.uint32)
    r = rays(g["ray"]["org"], [g["ray"]["dir"]] * 2)
    oracle.intersect_brute([(v, ix)], r)
    e = g["expected"]
    for k in range(2):
        assert r["geomID"][k] == e["geomID"] and r["primID"][k] == e["primID"]
        assert abs(r["tfar"][k] - e["t"]) <= 1e-6 * e["t"]
        assert abs(r["u"][k] - e["u"]) <= 1e-5 and abs(r["v"][k] - e["v"]) <= 1e-5
        ng = np.array([r["Ng_x"][k], r["Ng_y"][k], r["Ng_z"][k]], np.float32)
        # bitwise, including the sign of the zero component
        assert ng.tobytes() == np.array(e["Ng"], np.float32).tobytes()
    # IEEE division gives the exact answer
    assert r["tfar"][0] == np.float32(1.0) and r["u"][0] == np.float32(0.25)


@pytest.fixture(scope="module")
def c1_sphere(rtc):
    v, ix = rtc.gen_sphere(128)
    return v, ix


def test_c1_sphere_generator_topology(c1_sphere, golden_probe):
    v, ix = c1_sphere
    g = golden_probe["uv_sphere_pinhole"]
    assert v.shape[0] == g["num_vertices"] and ix.shape[0] == g["num_triangles"]
    assert np.allclose(np.linalg.norm(v.astype(np.float64), axis=1), 1.0, atol=1e-6)
    # numpy's float32 sin/cos may differ from the C library by an ulp; topology is identical
    nv, nix = uv_sphere(128)
    assert np.array_equal(nix, ix)
    assert np.abs(nv - v).max() < 1e-6


def test_c1_hit_count_matches_reference(oracle, rtc, c1_sphere, golden_probe):
    v, ix = c1_sphere
    r = rtc.gen_pinhole(1024, 1024)
    bvh = oracle.BVH([(v, ix)])
    bvh.intersect(r)
    assert int((r["geomID"] != INV).sum()) == golden_probe["uv_sphere_pinhole"]["expected_hits"]
    # hit points lie on the unit sphere's tessellation (inside radius 1)
    hit = r["geomID"] != INV
    p = np.stack([r["org_x"] + r["tfar"] * r["dir_x"], r["org_y"] + r["tfar"] * r["dir_y"],
                  r["org_z"] + r["tfar"] * r["dir_z"]], -1)[hit].astype(np.float64)
    rad = np.linalg.norm(p, axis=1)
    assert rad.max() < 1.0 + 1e-5 and rad.min() > math.cos(math.pi / 128) - 1e-4


# --------------------------------------------- reference known-answer tests --

def test_triangle_hit_test_known_answer(oracle):
    """verify.cpp:2339-2424: rays from (0,0,-1) to interior points of the unit
    triangle hit it at t=1 with the point's barycentrics, within 16 ulp."""
    rng = np.random.default_rng(7)
    u0 = rng.uniform(0.05, 0.9, 256).astype(np.float32)
    v0 = (rng.uniform(0.05, 0.9, 256) * (1 - u0) * 0.95).astype(np.float32)
    org = np.array([0.0, 0.0, -1.0], np.float32)
    d = np.stack([u0, v0, np.ones_like(u0)], -1).astype(np.float32)
    r = rays(org, d)
    tri = (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]], np.uint32))
    oracle.intersect_brute([tri], r)
    ulp = 16 * np.finfo(np.float32).eps
    assert (r["geomID"] == 0).all() and (r["primID"] == 0).all()
    assert np.abs(r["tfar"] - 1).max() <= ulp
    assert np.abs(r["u"] - u0).max() <= ulp and np.abs(r["v"] - v0).max() <= ulp
    assert (r["Ng_x"] == 0).all() and (r["Ng_y"] == 0).all() and (r["Ng_z"] == 1).all()


@pytest.mark.parametrize("segments", [50, 500])
def test_watertight_closed_sphere(oracle, segments):
    """verify.cpp:2898-2979: rays from inside a closed sphere miss at most 0.002%."""
    v, ix = uv_sphere(segments)
    r = random_rays(20000, 11, -0.5, 0.5)
    bvh = oracle.BVH([(v, ix)])
    bvh.intersect(r)
    assert (r["geomID"] == INV).mean() <= 2e-5


def test_moeller_special_cases(oracle):
    v0 = np.array([0, 0, 0], np.float32)
    p1 = np.array([1, 0, 0], np.float32)
    p2 = np.array([0, 1, 0], np.float32)
    # parallel ray: den == 0 -> miss
    ok, _ = oracle.moeller_verts([0.2, 0.2, 1], [1, 0, 0], 0, np.inf, v0, p1, p2)
    assert not ok
    # through a vertex and along an edge: U = 0 / V = 0 / U+V = absDen are inclusive
    for o in ([0, 0, -1], [0.5, 0, -1], [0, 0.5, -1], [0.5, 0.5, -1]):
        ok, h = oracle.moeller_verts(o, [0, 0, 1], 0, np.inf, v0, p1, p2)
        assert ok, o
    ok, _ = oracle.moeller_verts([0.51, 0.5, -1], [0, 0, 1], 0, np.inf, v0, p1, p2)
    assert not ok
    # depth test: absDen*tnear < T (strict), T <= absDen*tfar (inclusive)
    ok, h = oracle.moeller_verts([0.2, 0.2, -1], [0, 0, 1], 1.0, np.inf, v0, p1, p2)
    assert not ok
    ok, h = oracle.moeller_verts([0.2, 0.2, -1], [0, 0, 1], 0.0, 1.0, v0, p1, p2)
    assert ok and h[7] == 1.0
    ok, h = oracle.moeller_verts([0.2, 0.2, -1], [0, 0, 1], 0.0, np.nextafter(np.float32(1), np.float32(0)),
                                 v0, p1, p2)
    assert not ok
    # back faces are hit (no culling); the geometric normal is not flipped
    okf, hf = oracle.moeller_verts([0.2, 0.2, -1], [0, 0, 1], 0, np.inf, v0, p1, p2)
    okb, hb = oracle.moeller_verts([0.2, 0.2, 1], [0, 0, -1], 0, np.inf, v0, p1, p2)
    assert okf and okb and hf[6] == hb[6] == 1.0
    # a NaN direction never hits
    ok, _ = oracle.moeller_verts([0.2, 0.2, -1], [np.nan, 0, 1], 0, np.inf, v0, p1, p2)
    assert not ok


def test_invalid_triangles_are_skipped(oracle):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [np.nan, 0, 0], [2e18, 0, 0]], np.float32)
    ix = np.array([[0, 1, 2], [0, 1, 9], [3, 1, 2], [4, 1, 2], [0, 1, 2]], np.uint32)
    r = rays([0.2, 0.2, -1], [[0, 0, 1]])
    oracle.intersect_brute([(v, ix)], r)
    # primitives 1-3 are invalid; 0 and 4 are the same triangle -> tie -> smaller primID
    assert r["primID"][0] == 0
    bvh = oracle.BVH([(v, ix)])
    assert bvh.stats()["prims"] == 2 and bvh.validate() == 0


def test_ties_resolve_to_smaller_ids(oracle):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    ix = np.array([[0, 1, 2]] * 5, np.uint32)
    r = rays([0.2, 0.2, -1], [[0, 0, 1]] * 3)
    oracle.intersect_brute([(v, ix), (v, ix)], r)
    assert (r["geomID"] == 0).all() and (r["primID"] == 0).all()
    bvh = oracle.BVH([(v, ix), (v, ix)])
    r2 = rays([0.2, 0.2, -1], [[0, 0, 1]] * 3)
    bvh.intersect(r2)
    assert records_equal(r, r2)


def test_inactive_rays_untouched(oracle):
    v, ix = soup(200, 3)
    r = random_rays(300, 4)
    r["tnear"][::3] = 2.0
    r["tfar"][::3] = 1.0                          # tnear > tfar
    r["tnear"][1::7] = np.nan                     # invalid interval
    before = r.copy()
    bvh = oracle.BVH([(v, ix)])
    bvh.intersect(r)
    inactive = np.zeros(len(r), bool)
    inactive[::3] = True
    inactive[1::7] = True
    assert records_equal(r[inactive], before[inactive])


# ------------------------------------------------- BVH == brute force -------

@pytest.mark.parametrize("seed", [1, 2, 3])
def test_bvh_closest_equals_brute_force(oracle, seed):
    v, ix = soup(3000, seed, h=0.06)
    r = random_rays(4000, seed + 100, -0.1, 1.1)
    r["tfar"][::5] = 0.05
    r["tnear"][::7] = 0.01
    a = r.copy()
    b = r.copy()
    oracle.intersect_brute([(v, ix)], a)
    bvh = oracle.BVH([(v, ix)])
    assert bvh.validate() == 0
    bvh.intersect(b)
    assert len(mismatches(a, b)) == 0
    assert (a["geomID"] != INV).mean() > 0.3


def test_bvh_occluded_equals_brute_force(oracle):
    v, ix = soup(3000, 5, h=0.06)
    r = random_rays(4000, 6, tfar=0.1).astype(RAYHIT)
    a = r.copy()
    b = r.copy()
    oracle.occluded_brute([(v, ix)], a)
    bvh = oracle.BVH([(v, ix)])
    bvh.occluded(b)
    assert records_equal(a, b)
    assert 0.05 < np.isneginf(a["tfar"]).mean() < 0.95


def test_bvh_multi_geometry_and_nan_rays(oracle):
    v1, i1 = soup(500, 8)
    v2, i2 = uv_sphere(20)
    v2 = (v2 * 0.3 + 0.5).astype(np.float32)
    r = random_rays(2000, 9)
    r["dir_x"][::11] = np.nan
    r["org_y"][5::13] = np.inf
    a, b = r.copy(), r.copy()
    oracle.intersect_brute([(v1, i1), (v2, i2)], a)
    bvh = oracle.BVH([(v1, i1), (v2, i2)])
    bvh.intersect(b)
    assert records_equal(a, b)
    assert set(np.unique(a["geomID"])) >= {0, 1, INV}


# ------------------------------------------------------ builder properties --

def test_sah_builder_structure(oracle):
    for n, seed in ((1, 1), (4, 2), (29, 3), (5000, 4)):
        v, ix = soup(n, seed)
        bvh = oracle.BVH([(v, ix)])
        s = bvh.stats()
        assert bvh.validate() == 0, n
        assert s["prims"] == n and s["max_depth"] <= 40
        if n > 28:
            assert s["inner_nodes"] >= 1
    # identical centroids: no valid binned split, median fallback still builds
    v = np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), (300, 1))
    ix = np.arange(900, dtype=np.uint32).reshape(300, 3)
    bvh = oracle.BVH([(v, ix)])
    assert bvh.validate() == 0 and bvh.stats()["prims"] == 300


def test_parallel_build_is_the_serial_tree(oracle):
    """Subtrees built concurrently (per-subtree arenas appended in child
    order) give the serial build's tree: the same node, leaf and depth
    counts, the same SAH statistic to the last bit, and every ray visits
    the same number of nodes and triangles in both."""
    v, ix = soup(200000, 11)
    trees = [oracle.BVH([(v, ix)], nthreads=t) for t in (1, 3, 8)]
    st = [b.stats() for b in trees]
    assert st[0] == st[1] == st[2]
    assert all(b.validate() == 0 for b in trees)
    rays = random_rays(20000, 12)
    counts = [b.intersect_counted(rays.copy()) for b in trees]
    for n, t in counts[1:]:
        assert np.array_equal(n, counts[0][0]) and np.array_equal(t, counts[0][1])


def test_sah_statistic_closed_form(oracle):
    """The SAH statistic (reference kernels/bvh/bvh_statistics.cpp:40-56,
    bvh_statistics.h:36-38,99-101: half area of every inner node, plus half
    area times primitive blocks of every leaf, over the root's half area) on
    a tree whose shape the build cannot miss: 8 tight clusters of 4
    triangles at the corners of a cube make one inner node with 8 leaves of
    one block each, so SAH = 1 + sum(A_cluster) / A_root exactly."""
    rng = np.random.default_rng(21)
    verts, boxes = [], []
    for c in range(8):
        corner = np.array([(c >> 0) & 1, (c >> 1) & 1, (c >> 2) & 1], np.float32) * 10.0
        tv = (corner + rng.uniform(0.0, 0.5, (12, 3))).astype(np.float32)
        verts.append(tv)
        boxes.append((tv.min(0).astype(np.float64), tv.max(0).astype(np.float64)))
    v = np.concatenate(verts)
    ix = np.arange(len(v), dtype=np.uint32).reshape(-1, 3)
    bvh = oracle.BVH([(v, ix)])
    s = bvh.stats()
    assert (s["inner_nodes"], s["leaves"], s["prims"]) == (1, 8, 32)

    def ha(lo, hi):
        d = hi - lo
        return d[0] * d[1] + d[1] * d[2] + d[2] * d[0]

    root = ha(v.min(0).astype(np.float64), v.max(0).astype(np.float64))
    expect = 1.0 + sum(ha(lo, hi) for lo, hi in boxes) / root
    assert abs(s["sah"] - expect) <= 1e-6 * expect, (s["sah"], expect)
    # leaf cost per triangle counts 4 per leaf instead of 1 block
    expect_tri = 1.0 + 4.0 * sum(ha(lo, hi) for lo, hi in boxes) / root
    assert abs(s["sah_per_tri"] - expect_tri) <= 1e-6 * expect_tri


def test_binned_split_cost_matches_partition(oracle):
    """The split the binner returns costs exactly what its partition costs
    under halfArea * ceil(n/4), and separates two clusters."""
    rng = np.random.default_rng(3)
    c = np.concatenate([rng.random((40, 3)) * 0.1, rng.random((60, 3)) * 0.1 + [0.8, 0.0, 0.0]])
    lo = (c - 0.01).astype(np.float32)
    hi = (c + 0.01).astype(np.float32)
    dim, pos, sah, nbins = oracle.binned_split(lo, hi)
    assert dim == 0 and nbins == min(32, int(4 + 0.05 * 100))
    left = (lo[:, 0] + hi[:, 0]) < 0.8
    assert left.sum() == 40

    def half_area(l, h):
        d = (h.max(0) - l.min(0)).astype(np.float64)
        return d[0] * d[1] + d[1] * d[2] + d[2] * d[0]

    expect = half_area(lo[left], hi[left]) * math.ceil(40 / 4) + half_area(lo[~left], hi[~left]) * math.ceil(60 / 4)
    assert abs(sah - expect) <= 1e-5 * expect


def test_half_area_closed_form(oracle):
    rng = np.random.default_rng(1)
    for _ in range(50):
        lo = rng.random(3).astype(np.float32)
        hi = (lo + rng.random(3)).astype(np.float32)
        d = hi.astype(np.float64) - lo.astype(np.float64)
        exact = d[0] * d[1] + d[1] * d[2] + d[2] * d[0]
        assert abs(oracle.half_area(lo, hi) - exact) <= 4e-7 * exact


def test_morton_interleave_textbook(oracle):
    rng = np.random.default_rng(2)
    for x, y, z in rng.integers(0, 1024, (200, 3)):
        expect = 0
        for b in range(10):
            expect |= ((int(x) >> b) & 1) << (3 * b)
            expect |= ((int(y) >> b) & 1) << (3 * b + 1)
            expect |= ((int(z) >> b) & 1) << (3 * b + 2)
        assert oracle.bit_interleave(int(x), int(y), int(z)) == expect
    # codes of boxes order like the lattice cells they fall in
    lo = np.array([[0, 0, 0], [1, 1, 1], [0.5, 0, 0]], np.float32)
    codes = oracle.morton_codes(lo, lo)
    assert codes[0] == 0 and codes[1] == oracle.bit_interleave(1013, 1013, 1013)


# ---------------------------------------------------- Pluecker (robust mode) --

def test_pluecker_known_answer(oracle):
    """TriangleHitTest (verify.cpp:2339-2424) with the robust intersector: rays
    from (0,0,-1) through interior points of the unit triangle hit at t=1 with
    the point's barycentrics; U+V+W and both doubled terms are exact here."""
    rng = np.random.default_rng(8)
    u0 = rng.uniform(0.05, 0.9, 256).astype(np.float32)
    v0 = (rng.uniform(0.05, 0.9, 256) * (1 - u0) * 0.95).astype(np.float32)
    d = np.stack([u0, v0, np.ones_like(u0)], -1).astype(np.float32)
    r = rays([0.0, 0.0, -1.0], d)
    tri = (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]], np.uint32))
    oracle.intersect_brute([tri], r, robust=True)
    ulp = 16 * np.finfo(np.float32).eps
    assert (r["geomID"] == 0).all() and (r["primID"] == 0).all()
    assert (r["tfar"] == 1.0).all()
    assert np.abs(r["u"] - u0).max() <= ulp and np.abs(r["v"] - v0).max() <= ulp
    assert (r["Ng_x"] == 0).all() and (r["Ng_y"] == 0).all() and (r["Ng_z"] == 1).all()
    # by hand for the point (0.25, 0.25): U = V = 0.5, W = 1, UVW = 2, den = T = 2
    ok, h = oracle.pluecker([0.25, 0.25, -1.0], [0, 0, 1], 0.0, np.inf, *tri[0])
    assert ok and list(h[:4]) == [1.0, 0.5, 0.5, 2.0] and list(h[4:7]) == [0.0, 0.0, 1.0]


def test_pluecker_special_cases(oracle):
    tri = [np.array(p, np.float32) for p in ([0, 0, 1], [1, 0, 1], [0, 1, 1])]
    # t == tnear and t == tfar are both inside (tnear <= t <= tfar), unlike
    # Moeller's strict lower bound
    ok, h = oracle.pluecker([0.2, 0.2, 0.0], [0, 0, 1], 1.0, 1.0, *tri)
    assert ok and h[0] == 1.0
    assert not oracle.moeller_verts([0.2, 0.2, 0.0], [0, 0, 1], 1.0, 1.0, *tri)[0]
    assert not oracle.pluecker([0.2, 0.2, 0.0], [0, 0, 1], 1.0000001, np.inf, *tri)[0]
    # parallel ray (den == 0) misses; back side hits (no culling)
    assert not oracle.pluecker([0.2, 0.2, 0.0], [1, 0, 0], 0.0, np.inf, *tri)[0]
    ok, h = oracle.pluecker([0.2, 0.2, 2.0], [0, 0, -1], 0.0, np.inf, *tri)
    assert ok and h[0] == 1.0
    # a ray through a shared edge hits both triangles of a quad; through the
    # shared vertex of a fan it hits at least one
    quad = [np.array(p, np.float32) for p in ([0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1])]
    for org in ([0.5, 0.5, 0.0], [0.25, 0.25, 0.0], [0.0, 0.0, 0.0], [1.0, 1.0, 0.0]):
        a = oracle.pluecker(org, [0, 0, 1], 0.0, np.inf, quad[0], quad[1], quad[2])[0]
        b = oracle.pluecker(org, [0, 0, 1], 0.0, np.inf, quad[0], quad[2], quad[3])[0]
        assert a and b, org


@pytest.mark.parametrize("segments", [50, 200])
def test_pluecker_watertight_vertices_and_edges(oracle, segments):
    """Rays from the centre of the closed UV sphere through every edge
    midpoint never miss with the robust test: the two triangles sharing an
    edge compute its edge function as exact negatives, so one side always
    accepts (the watertightness the reference claims for Pluecker,
    triangle_intersector_pluecker.h:11-16). Through vertices there is no such
    guarantee; misses stay rare and far below Moeller's."""
    v, ix = uv_sphere(segments)
    e = np.concatenate([ix[:, [0, 1]], ix[:, [1, 2]], ix[:, [2, 0]]])
    mids = (0.5 * (v[e[:, 0]] + v[e[:, 1]])).astype(np.float32)
    verts = v[np.unique(ix)]
    bvh = oracle.BVH([(v, ix)])
    rm = rays([0.0, 0.0, 0.0], mids)
    bvh.intersect(rm, robust=True)
    assert (rm["geomID"] == 0).all()
    rv = rays([0.0, 0.0, 0.0], verts)
    bvh.intersect(rv, robust=True)
    robust_vertex_misses = int((rv["geomID"] == INV).sum())
    assert robust_vertex_misses <= 0.05 * len(verts)
    mo = rays([0.0, 0.0, 0.0], np.concatenate([mids, verts]))
    bvh.intersect(mo)
    assert int((mo["geomID"] == INV).sum()) > 10 * max(1, robust_vertex_misses)
    rb = rays([0.0, 0.0, 0.0], mids[::97])
    oracle.intersect_brute([(v, ix)], rb, robust=True)
    assert records_equal(rm[::97].copy(), rb)


def test_pluecker_agrees_with_moeller_off_edges(oracle):
    """Two different formulas for the same hit: same triangle for rays not
    grazing an edge, t within 1e-6 relative, u and v within 1e-5."""
    v, ix = soup(3000, 31, h=0.05)
    r = random_rays(20000, 32)
    a, b = r.copy(), r.copy()
    oracle.intersect_brute([(v, ix)], a)
    oracle.intersect_brute([(v, ix)], b, robust=True)
    both = (a["geomID"] == 0) & (b["geomID"] == 0)
    same = both & (a["primID"] == b["primID"])
    assert same.sum() >= 0.999 * both.sum() and both.sum() > 1000
    assert ((a["geomID"] == INV) == (b["geomID"] == INV)).mean() > 0.999
    # both carry an absolute error of a few coordinate ulps, so t is compared
    # relatively except where the origin nearly touches the triangle
    ta, tb = a["tfar"][same], b["tfar"][same]
    assert (np.abs(ta - tb) <= 1e-6 * ta + 1e-6).all()
    assert np.percentile(np.abs(ta - tb) / ta, 99) <= 1e-6
    du, dv = np.abs(a["u"][same] - b["u"][same]), np.abs(a["v"][same] - b["v"][same])
    assert du.max() <= 1e-4 and dv.max() <= 1e-4
    assert np.percentile(du, 99) <= 1e-5 and np.percentile(dv, 99) <= 1e-5
    # Ng: Moeller cross(e2,e1) and the stable normal are the same vector up to rounding
    na = np.stack([a["Ng_x"], a["Ng_y"], a["Ng_z"]], -1)[same]
    nb = np.stack([b["Ng_x"], b["Ng_y"], b["Ng_z"]], -1)[same]
    assert np.abs(na - nb).max() <= 1e-5 * np.abs(na).max()


@pytest.mark.parametrize("occluded", [False, True])
def test_robust_bvh_equals_brute_force(oracle, occluded):
    v, ix = soup(5000, 33, h=0.03)
    r = random_rays(20000, 34, tfar=0.2 if occluded else np.inf)
    bvh = oracle.BVH([(v, ix)])
    a, b = r.copy(), r.copy()
    if occluded:
        bvh.occluded(a, robust=True)
        oracle.occluded_brute([(v, ix)], b, robust=True)
        assert 0.05 < np.isneginf(a["tfar"]).mean() < 0.95
    else:
        bvh.intersect(a, robust=True)
        oracle.intersect_brute([(v, ix)], b, robust=True)
    assert records_equal(a, b)


# -------------------------------------------------------------- instances --
# The instance restatement is pinned by closed forms of the affine inverse,
# by W*L = I on random transforms (a transposed adjoint, a dropped cofactor
# or a sign slip breaks it), by translation invariance on an exact grid
# (an instance translated by t must hit exactly what the mesh moved by t
# hits), and by the per-object BVH path agreeing with brute force.

def random_l2w(n, seed):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w)], -1),
                  np.stack([2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w)], -1),
                  np.stack([2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y)], -1)], 1)
    s = rng.uniform(0.5, 2.0, (n, 3))  # non-uniform scale: the inverse is not a transpose
    out = np.empty((n, 12), np.float32)
    out[:, :9] = (R * s[:, :, None]).reshape(n, 9)  # columns vx, vy, vz
    out[:, 9:] = rng.uniform(-50, 50, (n, 3))
    return out


def test_affine_inverse_closed_forms(oracle):
    m = np.array([2, 0, 0, 0, 0.25, 0, 0, 0, 8, 1.5, -3, 64], np.float32)
    assert oracle.affine_inverse(m).tolist() == [0.5, 0, 0, 0, 4, 0, 0, 0, 0.125, -0.75, 12, -8]
    # a cyclic axis permutation: x -> y, y -> z, z -> x, then translate
    p = np.array([0, 1, 0, 0, 0, 1, 1, 0, 0, 5, 6, 7], np.float32)
    w = oracle.affine_inverse(p)
    assert w.tolist() == [0, 0, 1, 1, 0, 0, 0, 1, 0, -6, -7, -5]
    for pt in ([1, 2, 3], [-0.5, 0.25, 10]):
        assert oracle.xfm_point(w, oracle.xfm_point(p, pt)).tolist() == pt


def test_affine_inverse_random_transforms(oracle):
    for m in random_l2w(64, 1):
        w = oracle.affine_inverse(m)
        L = m[:9].reshape(3, 3).T.astype(np.float64)   # columns -> matrix
        W = w[:9].reshape(3, 3).T.astype(np.float64)
        assert np.abs(W @ L - np.eye(3)).max() < 2e-6
        assert np.abs(W @ m[9:] + w[9:]).max() < 2e-4
        p = np.array([0.3, -7.0, 11.0], np.float32)
        back = oracle.xfm_point(m, oracle.xfm_point(w, p))
        assert np.abs(back - p).max() < 1e-4
        v = np.array([1.0, 2.0, -3.0], np.float32)
        assert np.allclose(oracle.xfm_vector(m, v), L @ v, rtol=1e-6, atol=1e-5)  # no translation


def test_instance_translation_invariance(oracle):
    # vertices and ray origins on a 2^-8 grid in [-4, 4), translation by
    # powers of two: every subtraction in the Moeller test is exact, so the
    # instance must report bit-identical t, u, v, Ng to the moved mesh
    rng = np.random.default_rng(3)
    n = 400
    v = np.round(rng.uniform(-2, 2, (3 * n, 3)) * 256) / 256
    v = v.astype(np.float32)
    ix = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    t = np.array([8.0, -16.0, 4.0], np.float32)
    org = (np.round(rng.uniform(-4, 4, (3000, 3)) * 256) / 256).astype(np.float32)
    tgt = v[rng.integers(0, 3 * n, 3000)] + rng.uniform(-0.05, 0.05, (3000, 3)).astype(np.float32)
    r = rays(org + t, tgt - org)
    I = oracle.Instances()
    I.add_instance(I.add_object([(v, ix)]), np.concatenate([np.eye(3, dtype=np.float32).reshape(-1), t]), 5)
    a = r.copy()
    I.intersect(a)
    b = r.copy()
    oracle.intersect_brute([(v + t, ix)], b)
    hit = b["geomID"] != INV
    assert hit.mean() > 0.3
    assert (a["geomID"] != INV).tolist() == hit.tolist()
    for k in ("tfar", "u", "v", "Ng_x", "Ng_y", "Ng_z", "primID"):
        assert a[k].tobytes() == b[k].tobytes(), k
    assert (a["instID"][hit] == 5).all()


def test_instance_order_and_identity(oracle):
    v, ix = soup(300, 21, h=0.05)
    I = oracle.Instances()
    ob = I.add_object([(v, ix)])
    own = I.add_object([soup(200, 22, h=0.05)])
    I.add_identity(own)
    m = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], np.float32)
    # two instances of the same object at the same place: equal t everywhere,
    # the smaller instID wins
    I.add_instance(ob, m, 9)
    I.add_instance(ob, m, 4)
    r = random_rays(3000, 23, -0.1, 1.1)
    a = r.copy()
    I.intersect(a)
    b = r.copy()
    oracle.intersect_brute([(v, ix)], b, inst_id=4)
    c = r.copy()
    oracle.intersect_brute([soup(200, 22, h=0.05)], c)
    for i in range(len(r)):
        cands = [(x["tfar"], x["instID"]) for x in (b[i], c[i]) if x["geomID"] != INV]
        if not cands:
            assert a["geomID"][i] == INV
            continue
        assert (a["tfar"][i], a["instID"][i]) == min(cands)
    # a full instance stack keeps only the identity record
    d = r.copy()
    I.intersect(d, ctx_inst=3)
    c3 = r.copy()
    oracle.intersect_brute([soup(200, 22, h=0.05)], c3, inst_id=3)
    assert records_equal(d, c3)


def test_instance_bvh_equals_brute_force(oracle):
    I = oracle.Instances()
    objs = [I.add_object([soup(400, 30 + k, h=0.05)], robust=(k == 1)) for k in range(3)]
    for k, m in enumerate(random_l2w(25, 7)):
        m = m.copy()
        m[9:] *= 0.02
        I.add_instance(objs[k % 3], m, k)
    r = random_rays(2000, 31, -1.5, 1.5)
    a, b = r.copy(), r.copy()
    I.intersect(a)
    I.intersect_bvh(b)
    assert records_equal(a, b)
    assert (a["geomID"] != INV).mean() > 0.2
    r["tfar"] = 0.5
    a, b = r.copy(), r.copy()
    I.occluded(a)
    I.occluded_bvh(b)
    assert records_equal(a, b)
    assert 0 < np.isneginf(a["tfar"]).sum() < len(r)
