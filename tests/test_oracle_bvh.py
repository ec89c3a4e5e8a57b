"""The reference's BVH tests (tests/bvh.test.ts) restated against the CPU oracle, plus pins of
the builder's exact arrays, which the reference's tests leave open: node layouts derived by hand
from bvh.ts (SAH tie-break, right-first numbering, two-cursor partition, coincident and SAH-leaf
rules) and a node-by-node numpy replay of buildBvh against the oracle's output on every test
input, which must also reject builds with a planted numbering or order mistake."""
import numpy as np
import pytest

import oracle
from helpers import probe_rays_400, random_triangles, reorder
from scenes_common import BVH_CASES, geometry_from
from meercadraytracer_amd.types import CompiledBvh, CompiledGeometry


@pytest.fixture(scope='module')
def lcg():
    raw = random_triangles(3000, 7)
    g, order = geometry_from(raw, 3000, oracle.build_bvh)
    return raw, g, order


def test_order_is_permutation_and_leaves_contain_triangles(lcg):  # bvh.test.ts:62-80
    raw, g, order = lcg
    assert sorted(order.tolist()) == list(range(3000))
    data = g.triangles.reshape(-1, 9).astype(np.float64)
    covered = 0
    nodes = g.bvh.nodes.reshape(-1, 3)
    bounds = g.bvh.bounds.reshape(-1, 6)
    for n in range(g.bvh.nodeCount):
        first, c = nodes[n, 0], nodes[n, 1]
        if c == 0:
            continue
        covered += c
        for k in range(first, first + c):
            for v in range(3):
                p = data[k, 0:3] + (data[k, v * 3:v * 3 + 3] if v > 0 else 0)
                assert np.all(p >= bounds[n, :3] - 1e-4) and np.all(p <= bounds[n, 3:] + 1e-4)
    assert covered == 3000


def test_closest_and_all_hits_match_brute_force(lcg):  # bvh.test.ts:82-107
    raw, g, order = lcg
    rays = probe_rays_400()
    tri, t, u, v, front, _ = oracle.intersect_closest(g, rays)
    btri, bt, ball = oracle.brute_force(g.triangles, 3000, rays)
    hits = tri >= 0
    assert np.array_equal(hits, btri >= 0)
    assert np.array_equal(tri[hits], btri[hits])
    assert np.all(np.abs(t[hits] - bt[hits]) < 5e-4)  # toBeCloseTo(expected, 3)
    count, _ = oracle.intersect_any_count(g, rays)
    assert np.array_equal(count, ball)
    assert hits.sum() > 100


def _one(side):
    tri = np.array([-1, -1, 0, 2, 0, 0, 0, 2, 0], np.float32)
    g, _ = geometry_from(tri, 1, oracle.build_bvh, sides=side)
    return g


def test_culling():  # bvh.test.ts:109-123
    from_front = np.array([[0, 0, 5, 0, 0, -1]], float)
    from_back = np.array([[0, 0, -5, 0, 0, 1]], float)
    hit = lambda side, r: oracle.intersect_closest(_one(side), r)
    tri, t, u, v, front, _ = hit(0, from_front)
    assert tri[0] == 0 and front[0] == 1
    assert hit(0, from_back)[0][0] == -1
    assert hit(1, from_front)[0][0] == -1
    tri, t, u, v, front, _ = hit(1, from_back)
    assert tri[0] == 0 and front[0] == 0
    assert hit(2, from_front)[0][0] == 0 and hit(2, from_back)[0][0] == 0


def test_flat_box_ray_along_face():  # bvh.test.ts:125-133
    quad = np.array([-1, -1, 0, 2, 0, 0, 2, 2, 0, -1, -1, 0, 2, 2, 0, 0, 2, 0], np.float32)
    g, _ = geometry_from(quad, 2, oracle.build_bvh)
    tri, t, *_ = oracle.intersect_closest(g, np.array([[-1, 0, 3, 0, 0, -1]], float))
    assert tri[0] >= 0 and abs(t[0] - 3) < 5e-3


def test_coincident_triangles():  # bvh.test.ts:135-142
    raw, n = BVH_CASES['coincident500']()
    g, _ = geometry_from(raw, n, oracle.build_bvh)
    assert g.bvh.nodeCount > 1
    tri, *_ = oracle.intersect_closest(g, np.array([[0.2, 0.2, 1, 0, 0, -1]], float))
    assert tri[0] >= 0


def test_empty_scene():  # bvh.test.ts:144-148
    bounds, nodes, order, nc = oracle.build_bvh(np.zeros(0, np.float32), 0)
    assert nc == 0
    g = CompiledGeometry(0, np.zeros(0, np.float32), *(np.zeros(0, np.float32),) * 4, np.zeros(0, np.uint32),
                         np.zeros(0, np.uint8), np.zeros(0, np.uint8), CompiledBvh(bounds, nodes, 0), 1e-5)
    tri, *_ = oracle.intersect_closest(g, np.array([[0, 0, 0, 0, 0, 1]], float))
    assert tri[0] == -1


def _spaced(k, step=2.0):
    """k unit right triangles along x: triangle i spans x in [step*i, step*i + 1], y in [0, 1]."""
    tris = []
    for i in range(k):
        tris += [step * i, 0, 0, 1, 0, 0, 0, 1, 0]
    return np.array(tris, np.float32)


def test_small_build_known_layout():
    """Hand-derived from bvh.ts:88-171 for five triangles spaced along x (centroids 2i+0.5, so
    bins 4i and 15 at the root): a split after k triangles costs k(2k-1) + (5-k)(9-2k), i.e. 29,
    21, 21, 29 for k = 1..4. The 2|3 split (first reached at split index 4) and the 3|2 split
    (index 8) tie, and the first strict minimum keeps 2|3. The two-cursor loop of :146-157 then
    swaps triangles 2 and 4, then 4 and 3, giving order [0, 1, 3, 4, 2]; the pair (1, 2) is
    allocated at :165-168 and both children are leaves (n <= 4)."""
    bounds, nodes, order, nc = oracle.build_bvh(_spaced(5), 5)
    assert nc == 3
    assert nodes.reshape(-1, 3).tolist() == [[1, 0, 0], [0, 2, 0], [2, 3, 0]]
    assert order.tolist() == [0, 1, 3, 4, 2]
    assert bounds.reshape(-1, 6).tolist() == [[0, 0, 0, 9, 1, 0], [0, 0, 0, 3, 1, 0], [4, 0, 0, 9, 1, 0]]


def test_build_known_layout_right_child_first():
    """Hand-derived for ten triangles spaced along x. Root: costs k(2k-1) + (10-k)(19-2k) have a
    unique minimum 90 at k = 5; the loop leaves order [0..4, 6, 7, 8, 9, 5]. The right child
    (node 2) is popped first (bvh.ts:169-170 push it last), so it takes the pair (3, 4) and the
    left child (node 1) the pair (5, 6). Each child holds five spaced triangles and splits 2|3
    as in the five-triangle case: node 2 reorders [6, 7, 8, 9, 5] to [6, 5, 9, 8, 7], node 1
    reorders [0..4] to [0, 1, 3, 4, 2]."""
    bounds, nodes, order, nc = oracle.build_bvh(_spaced(10), 10)
    assert nc == 7
    assert nodes.reshape(-1, 3).tolist() == [[1, 0, 0], [5, 0, 0], [3, 0, 0], [5, 2, 0], [7, 3, 0],
                                             [0, 2, 0], [2, 3, 0]]
    assert order.tolist() == [0, 1, 3, 4, 2, 6, 5, 9, 8, 7]
    assert bounds.reshape(-1, 6).tolist() == [
        [0, 0, 0, 19, 1, 0], [0, 0, 0, 9, 1, 0], [10, 0, 0, 19, 1, 0], [10, 0, 0, 13, 1, 0],
        [14, 0, 0, 19, 1, 0], [0, 0, 0, 3, 1, 0], [4, 0, 0, 9, 1, 0]]


def test_build_coincident_centroids_split_in_half():
    """bvh.ts:158-163: nine identical triangles have no axis with extent > 0, so every node with
    more than four splits its run in half (mid = first + (n >> 1)) on axis 0: root 4|5, then the
    right child (five, popped first) 2|3."""
    same = np.tile(np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32), 9)
    bounds, nodes, order, nc = oracle.build_bvh(same, 9)
    assert nodes.reshape(-1, 3).tolist() == [[1, 0, 0], [0, 4, 0], [3, 0, 0], [4, 2, 0], [6, 3, 0]]
    assert order.tolist() == list(range(9))
    assert bounds.reshape(-1, 6).tolist() == [[0, 0, 0, 1, 1, 0]] * 5


def _overlapping(k):
    """k large triangles whose boxes nearly coincide (x offsets of 0.01 i): every split costs
    about as much as the parent."""
    tris = []
    for i in range(k):
        tris += [0.01 * i, 0, 0, 10, 0, 0, 0, 10, 0]
    return np.array(tris, np.float32)


def test_build_sah_leaf_rule():
    """bvh.ts:136-144: a split is refused when 1 + cost / area(parent) >= n, but only for n <= 16.
    For k overlapping 10x10 triangles every split costs about k * 100 against a parent area of
    about 100, so 1 + cost/area > k: k = 6 and k = 16 stay one leaf (order untouched), k = 17
    must split."""
    for k in (6, 16):
        tris = _overlapping(k)
        bounds, nodes, order, nc = oracle.build_bvh(tris, k)
        assert nc == 1 and nodes.tolist() == [0, k, 0] and order.tolist() == list(range(k))
        x1 = np.float32(np.float64(tris[(k - 1) * 9]) + 10.0)
        assert bounds.tolist() == [0, 0, 0, float(x1), 10, 0]
    bounds, nodes, order, nc = oracle.build_bvh(_overlapping(17), 17)
    assert nc > 1 and nodes[1] == 0 and nodes[2] == 0


def _prim_boxes(tris, count):
    """bvh.ts:22-33: vertices v1 = v0 + e1, v2 = v0 + e2 in FP64; boxes stored as f32; centroids
    (min + max) * 0.5 from the unrounded FP64 min and max, stored as f32."""
    t = tris.reshape(-1, 9)[:count].astype(np.float64)
    v0 = t[:, 0:3]
    v1 = v0 + t[:, 3:6]
    v2 = v0 + t[:, 6:9]
    mn = np.minimum(np.minimum(v0, v1), v2)
    mx = np.maximum(np.maximum(v0, v1), v2)
    return mn.astype(np.float32), mx.astype(np.float32), ((mn + mx) * 0.5).astype(np.float32)


def _area(lo, hi):  # bvh.ts:180-183
    d = hi.astype(np.float64) - lo.astype(np.float64)
    dx, dy, dz = d[..., 0], d[..., 1], d[..., 2]
    return dx * dy + dy * dz + dz * dx


def _split_costs(bmin, bmax, cen, ids, cmin, cmax):
    """Every candidate of bvh.ts:88-135 for the triangles `ids`: costs[axis, split] (inf where the
    axis has no extent or a side is empty), evaluated in FP64 as the reference does."""
    costs = np.full((3, 15), np.inf)
    for axis in range(3):
        lo = np.float64(cmin[axis])
        extent = np.float64(cmax[axis]) - lo
        if not extent > 0:
            continue
        scale = 16 / extent
        bins = np.minimum(15, np.floor((cen[ids, axis].astype(np.float64) - lo) * scale)).astype(np.int64)
        cnt = np.bincount(bins, minlength=16)
        blo = np.full((16, 3), np.inf, np.float32)
        bhi = np.full((16, 3), -np.inf, np.float32)
        srt = np.argsort(bins, kind='stable')
        sb = bins[srt]
        starts = np.flatnonzero(np.r_[True, sb[1:] != sb[:-1]])
        used = sb[starts]
        blo[used] = np.minimum.reduceat(bmin[ids][srt], starts, axis=0)
        bhi[used] = np.maximum.reduceat(bmax[ids][srt], starts, axis=0)
        llo = np.minimum.accumulate(blo, 0)
        lhi = np.maximum.accumulate(bhi, 0)
        rlo = np.minimum.accumulate(blo[::-1], 0)[::-1]
        rhi = np.maximum.accumulate(bhi[::-1], 0)[::-1]
        lc = np.cumsum(cnt)
        rc = np.cumsum(cnt[::-1])[::-1]
        for i in range(15):
            if lc[i] == 0 or rc[i + 1] == 0:
                continue
            costs[axis, i] = lc[i] * _area(llo[i], lhi[i]) + rc[i + 1] * _area(rlo[i + 1], rhi[i + 1])
    return costs


def _replay_build(tris, count):
    """Replays buildBvh (bvh.ts:16-178) node by node in its own pop order against the oracle's
    output: each node's bounds; the leaf rules (n <= 4, depth >= 100, SAH leaf for n <= 16); the
    chosen axis as the first strict minimum over (axis 0..2, split 0..14) of the binned-SAH
    costs, recomputed here with numpy from the primitives; the two-cursor partition loop of
    :146-157 on the working order; the half split of coincident centroids; and the child pair
    numbering of a LIFO that pushes left then right (:165-171). The final working order must be
    the oracle's `order`. Returns the node count."""
    bounds, nodes, order, nc = oracle.build_bvh(tris, count)
    nodes = nodes.reshape(-1, 3)
    bounds = bounds.reshape(-1, 6)
    bmin, bmax, cen = _prim_boxes(tris, count)
    cur = np.arange(count, dtype=np.int64)
    stack = [(0, 0, count, 0)]
    next_node = 1
    while stack:
        node, first, n, depth = stack.pop()
        ids = cur[first:first + n]
        nlo, nhi = bmin[ids].min(0), bmax[ids].max(0)
        assert np.array_equal(bounds[node, :3], nlo) and np.array_equal(bounds[node, 3:], nhi), node
        if n <= 4 or depth >= 100:
            assert nodes[node].tolist() == [first, n, 0], node
            continue
        cmin, cmax = cen[ids].min(0), cen[ids].max(0)
        costs = _split_costs(bmin, bmax, cen, ids, cmin, cmax)
        best = int(np.argmin(costs))  # the first of equal minima, in (axis, split) order
        axis, split = divmod(best, 15)
        if np.isfinite(costs.flat[best]):
            split_cost = 1 + costs.flat[best] / max(_area(nlo, nhi), 1e-30)
            if split_cost >= n and n <= 16:
                assert nodes[node].tolist() == [first, n, 0], node
                continue
            lo = np.float64(cmin[axis])
            scale = 16 / (np.float64(cmax[axis]) - lo)
            i, j = first, first + n - 1
            while i <= j:
                b = min(15, int(np.floor((np.float64(cen[cur[i], axis]) - lo) * scale)))
                if b <= split:
                    i += 1
                else:
                    cur[i], cur[j] = cur[j], cur[i]
                    j -= 1
            mid = i
        else:
            axis = 0
            mid = first + (n >> 1)
        if mid == first or mid == first + n:
            mid = first + (n >> 1)
        assert nodes[node].tolist() == [next_node, 0, axis], node
        left = next_node
        next_node += 2
        stack.append((left, first, mid - first, depth + 1))
        stack.append((left + 1, mid, first + n - mid, depth + 1))
    assert next_node == nc
    assert np.array_equal(cur, order.astype(np.int64))
    return nc


@pytest.mark.parametrize('case', sorted(BVH_CASES))
def test_build_replay(case):
    tris, n = BVH_CASES[case]()
    assert _replay_build(tris, n) >= 1


def test_build_replay_hand_cases():
    for tris, n in ((_spaced(5), 5), (_spaced(10), 10), (_spaced(37, 1.5), 37), (_overlapping(17), 17),
                    (_overlapping(40), 40)):
        assert _replay_build(tris, n) >= 1


def test_replay_detects_planted_mistakes():
    """The replay must fail on a build that breaks the reference's rules: the children of one inner
    node swapped (right-first numbering), and two triangles of one leaf exchanged (the partition
    permutation)."""
    tris, n = BVH_CASES['lcg3000']()
    bounds, nodes, order, nc = oracle.build_bvh(tris, n)
    real = oracle.build_bvh

    def planted(kind):
        def build(t, c):
            b, nd, o, k = real(t, c)
            nd = nd.copy().reshape(-1, 3)
            o = o.copy()
            if kind == 'numbering':
                nd[0, 0] = nd[0, 0] + 2
            else:
                leaf = next(i for i in range(k) if nd[i, 1] >= 2)
                f = nd[leaf, 0]
                o[f], o[f + 1] = o[f + 1], o[f]
            return b, nd.reshape(-1), o, k
        return build
    for kind in ('numbering', 'order'):
        oracle.build_bvh = planted(kind)
        try:
            with pytest.raises(AssertionError):
                _replay_build(tris, n)
        finally:
            oracle.build_bvh = real
