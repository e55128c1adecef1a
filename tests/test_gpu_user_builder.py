"""rtcBuildBVH (the user builder API) on the GPU, through the C-ABI with
Python callbacks (reference kernels/common/rtcore_builder.cpp:141-188,
355-395).

The tree is the library's own (GPU binned SAH or Morton tree with the
SAH-optimal collapse), so the checks are the properties the reference's
builders guarantee and that a caller relies on: every primitive in exactly
one leaf, leaves pointing into the caller's reordered primitive array, the
bounds given to setNodeBounds equal to the union of the boxes below, child
counts within [2, maxBranchingFactor], leaf sizes within maxLeafSize, no
inner node over at most minLeafSize primitives, and the SAH cost no worse
than a plain object-median tree over the same boxes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def random_prims(rtc, n, seed):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3), dtype=np.float32)
    e = rng.random((n, 3), dtype=np.float32) * np.float32(0.02)
    p = np.zeros(n, rtc.RTCBuildPrimitive)
    p["lower_x"], p["lower_y"], p["lower_z"] = (c - e).T
    p["upper_x"], p["upper_y"], p["upper_z"] = (c + e).T
    p["geomID"] = 7
    p["primID"] = np.arange(n, dtype=np.uint32)
    return p


def boxes_of(prims):
    lo = np.stack([prims["lower_x"], prims["lower_y"], prims["lower_z"]], 1)
    hi = np.stack([prims["upper_x"], prims["upper_y"], prims["upper_z"]], 1)
    return lo, hi


def half_area(lo, hi):
    d = np.maximum(np.asarray(hi, np.float64) - np.asarray(lo, np.float64), 0)
    return d[..., 0] * (d[..., 1] + d[..., 2]) + d[..., 1] * d[..., 2]


def check_tree(prims, root, nodes, leaves, branching, min_leaf, max_leaf):
    """Walks the recorded tree; returns its SAH (nodes cost 1, primitives 1)."""
    lo, hi = boxes_of(prims)
    seen = np.zeros(len(prims), np.int32)
    sah = [0.0]

    def walk(ref):
        if ref in leaves:
            off, cnt = leaves[ref]
            assert 1 <= cnt <= max_leaf
            seen[off:off + cnt] += 1
            l, h = lo[off:off + cnt].min(0), hi[off:off + cnt].max(0)
            sah[0] += half_area(l, h) * cnt
            return l, h, cnt
        nd = nodes[ref]
        n = nd["n"]
        assert 2 <= n <= branching and len(nd["children"]) == n and len(nd["bounds"]) == n
        ls, hs, tot = [], [], 0
        for k, c in enumerate(nd["children"]):
            l, h, cnt = walk(c)
            assert tuple(np.float32(x) for x in l) == nd["bounds"][k][0]
            assert tuple(np.float32(x) for x in h) == nd["bounds"][k][1]
            ls.append(l); hs.append(h); tot += cnt
        assert tot > min_leaf
        l, h = np.min(ls, 0), np.max(hs, 0)
        sah[0] += half_area(l, h)
        return l, h, tot

    l, h, tot = walk(root)
    assert tot == len(prims) and (seen == 1).all()
    return sah[0] / half_area(l, h)


def median_tree_sah(prims, max_leaf):
    lo, hi = boxes_of(prims)
    c = (lo + hi) * 0.5
    total = [0.0]

    def rec(idx):
        l, h = lo[idx].min(0), hi[idx].max(0)
        if len(idx) <= max_leaf:
            total[0] += half_area(l, h) * len(idx)
            return
        total[0] += half_area(l, h)
        ax = int(np.argmax(h - l))
        o = idx[np.argsort(c[idx, ax], kind="stable")]
        rec(o[:len(o) // 2])
        rec(o[len(o) // 2:])

    rec(np.arange(len(prims)))
    return total[0] / half_area(lo.min(0), hi.max(0))


@pytest.mark.parametrize("quality,branching,min_leaf,max_leaf", [
    (1, 2, 1, 32), (1, 4, 1, 8), (1, 8, 2, 4), (0, 8, 1, 8), (2, 4, 4, 16), (1, 16, 1, 32)])
def test_user_builder_tree(rtc, device, quality, branching, min_leaf, max_leaf):
    prims = random_prims(rtc, 20000, seed=quality * 10 + branching)
    orig = prims.copy()
    root, nodes, leaves = rtc.build_user_bvh(device, prims, quality=quality, branching=branching, min_leaf=min_leaf,
                                             max_leaf=max_leaf)
    # the caller's array is reordered in place: a permutation of the input
    assert sorted(prims["primID"].tolist()) == list(range(len(prims)))
    assert (prims["geomID"] == 7).all()
    back = prims[np.argsort(prims["primID"])]
    assert back.tobytes() == orig.tobytes()
    sah = check_tree(prims, root, nodes, leaves, min(branching, 8), min_leaf, max_leaf)
    ref = median_tree_sah(orig, max_leaf)
    # binned SAH beats an object-median tree; the Morton tree comes close
    assert sah <= ref * (1.3 if quality == 0 else 1.05), (sah, ref)


def test_user_builder_small_and_empty(rtc, device):
    for n in (1, 2, 3, 33):
        prims = random_prims(rtc, n, seed=n)
        root, nodes, leaves = rtc.build_user_bvh(device, prims, branching=4, max_leaf=2)
        check_tree(prims, root, nodes, leaves, 4, 1, 2)
    prims = np.zeros(0, rtc.RTCBuildPrimitive)
    root, nodes, leaves = rtc.build_user_bvh(device, prims)
    assert root in leaves and leaves[root][1] == 0 and not nodes


def test_user_builder_errors_and_cancel(rtc, device):
    prims = random_prims(rtc, 100, seed=3)
    with pytest.raises(rtc.RTCError) as e:
        rtc.build_user_bvh(device, prims, capacity=50)
    assert e.value.code == rtc.RTC_ERROR_INVALID_ARGUMENT
    with pytest.raises(rtc.RTCError) as e:
        rtc.build_user_bvh(device, prims, progress=lambda u, f: False)
    assert e.value.code == rtc.RTC_ERROR_CANCELLED
    calls = []
    rtc.build_user_bvh(device, prims, progress=lambda u, f: calls.append(f) or True)
    assert calls[0] == 0.0 and calls[-1] == 1.0 and calls == sorted(calls)


def test_user_builder_sah_block_and_costs(rtc, device):
    # a larger leaf cost per block makes leaves smaller; cheaper nodes likewise
    prims = random_prims(rtc, 8000, seed=9)
    _, _, leaves_a = rtc.build_user_bvh(device, prims.copy(), branching=4, max_leaf=16, intersection_cost=0.2)
    _, _, leaves_b = rtc.build_user_bvh(device, prims.copy(), branching=4, max_leaf=16, intersection_cost=5.0)
    mean = lambda lv: np.mean([c for _, c in lv.values()])
    assert mean(leaves_a) > mean(leaves_b)
    _, _, leaves_c = rtc.build_user_bvh(device, prims.copy(), branching=4, max_leaf=16, block=4)
    _, _, leaves_d = rtc.build_user_bvh(device, prims.copy(), branching=4, max_leaf=16, block=1)
    assert mean(leaves_c) >= mean(leaves_d)
