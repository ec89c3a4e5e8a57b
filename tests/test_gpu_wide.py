"""Precision form on the GPU: the compressed 8-wide BVH built on the device from the binary BVH and
queried in FP32 (mcr_intersect_closest_wide) must return exactly the closest hit of the FP32
watertight checker over every triangle (oracle.wide_brute). Its child boxes are rounded outward
and its slab tests widened by their rounding bound, so it never skips a triangle the ray hits;
only equal-t ties may pick another triangle. front is bit-identical; t, u and v are the
reference's FP64 test (bvh.ts:228-252) on the winning triangle, rounded to FP32, except where
that test rejects the triangle, where they are the FP32 checker's."""
import numpy as np
import pytest

import oracle
from scenes_common import BVH_CASES, geometry_from, geometry_with_bvh, octahedron, random_rays, single_leaf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def ctx():
    from meercadraytracer_amd import native
    return native.default_context()


def _scene(ctx, g):
    from meercadraytracer_amd import native
    s = native.Scene(ctx)
    s.set_geometry(g, 1)
    return s


def _special_rays(g, n, seed):
    """Axis-parallel rays (zero direction components) and rays aimed at triangle vertices."""
    rng = np.random.Generator(np.random.PCG64(seed))
    v = g.triangles.reshape(-1, 9)
    lo = v[:, 0:3].min(0) - 1 if len(v) else np.zeros(3)
    hi = v[:, 0:3].max(0) + 1 if len(v) else np.ones(3)
    out = []
    for i in range(n):
        o = rng.uniform(lo, hi)
        if i % 2 and len(v):
            k = rng.integers(0, len(v))
            target = v[k, 0:3].astype(np.float64) + (v[k, 3:6] if i % 4 == 1 else 0)
            d = target - o
        else:
            d = np.zeros(3)
            d[i % 3] = 1 if (i // 3) % 2 else -1
            o[i % 3] = lo[i % 3] - 1 if d[i % 3] > 0 else hi[i % 3] + 1
        nd = np.linalg.norm(d)
        out.append(np.concatenate([o, d / nd if nd else d]))
    return np.array(out, np.float32)


def _refined(g, tri, rays, t32, u32, v32):
    """Expected outputs: for each hit, the reference's FP64 Moller-Trumbore of triangle tri[i]
    (bvh.ts:228-252, the operation order of mt_test in device.h) on the widened ray, rounded to
    FP32; the FP32 values where that test rejects the triangle."""
    t_out, u_out, v_out = t32.copy(), u32.copy(), v32.copy()
    tv = g.triangles.reshape(-1, 9).astype(np.float64)
    for i in np.flatnonzero(tri >= 0):
        a = tv[tri[i]]
        ox, oy, oz, dx, dy, dz = (float(x) for x in rays[i].astype(np.float64))
        e1x, e1y, e1z, e2x, e2y, e2z = a[3], a[4], a[5], a[6], a[7], a[8]
        px, py, pz = dy * e2z - dz * e2y, dz * e2x - dx * e2z, dx * e2y - dy * e2x
        det = e1x * px + e1y * py + e1z * pz
        if det == 0:
            continue
        inv = 1.0 / det
        sx, sy, sz = ox - a[0], oy - a[1], oz - a[2]
        u = (sx * px + sy * py + sz * pz) * inv
        if u < 0 or u > 1:
            continue
        qx, qy, qz = sy * e1z - sz * e1y, sz * e1x - sx * e1z, sx * e1y - sy * e1x
        v = (dx * qx + dy * qy + dz * qz) * inv
        if v < 0 or u + v > 1:
            continue
        t_out[i], u_out[i], v_out[i] = np.float32((e2x * qx + e2y * qy + e2z * qz) * inv), np.float32(u), np.float32(v)
    return t_out, u_out, v_out


def _check(ctx, g, rays, label, cull=False):
    s = _scene(ctx, g)
    info = s.build_wide()
    print(f'{label}: wide BVH {info}', flush=True)
    tri, t, u, v, front = s.intersect_closest_wide(rays, cull=cull)
    b_tri, b_t, b_u, b_v, b_f = oracle.wide_brute(g, rays, cull=cull)
    s.release()
    same = tri == b_tri
    # A tie: the GPU's triangle at the checker's FP32 distance too (the checker's t is the minimum
    # over all triangles), by the checker's arithmetic on that one triangle.
    g_t32 = b_t.copy()
    for i in np.flatnonzero(~same & (tri >= 0)):
        one = geometry_with_bvh(np.ascontiguousarray(g.triangles.reshape(-1, 9)[tri[i]]), 1,
                                np.zeros(6, np.float32), np.array([0, 1, 0], np.int32), 1)
        one.sides[:] = g.sides[tri[i]]
        g_t32[i] = oracle.wide_brute(one, rays[i:i + 1], cull=cull)[1][0]
    tie = ~same & (tri >= 0) & (b_tri >= 0) & (g_t32 == b_t)
    assert (same | tie).all(), f'{label}: {(~(same | tie)).sum()} rays differ'
    m = same
    e_t, e_u, e_v = _refined(g, tri, rays, b_t, b_u, b_v)
    assert np.array_equal(t[m].view(np.uint32), e_t[m].view(np.uint32))
    assert np.array_equal(u[m].view(np.uint32), e_u[m].view(np.uint32))
    assert np.array_equal(v[m].view(np.uint32), e_v[m].view(np.uint32))
    assert np.array_equal(front[m], b_f[m])
    return info, tri


@pytest.mark.parametrize('case', sorted(BVH_CASES))
def test_wide_equals_fp32_brute_force(ctx, case):
    tris, n = BVH_CASES[case]()
    g = geometry_from(tris, n, oracle.build_bvh)[0]
    rays = np.concatenate([random_rays(3000, 31).astype(np.float32), _special_rays(g, 1000, 32)])
    info, tri = _check(ctx, g, rays, case)
    assert info['triangle_count'] == n and info['node_count'] >= 1
    assert (tri >= 0).any()


@pytest.mark.parametrize('side', [0, 1])
def test_wide_culling(ctx, side):
    g = geometry_from(BVH_CASES['soup20k']()[0], 20000, oracle.build_bvh, sides=side)[0]
    _check(ctx, g, random_rays(4000, 33).astype(np.float32), f'cull side {side}', cull=True)


def test_wide_watertight(ctx):
    tris, n = octahedron(16)
    g = geometry_from(tris, n, oracle.build_bvh)[0]
    rng = np.random.Generator(np.random.PCG64(34))
    o = rng.uniform(-0.25, 0.25, (4000, 3))
    d = rng.normal(size=(4000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    _, tri = _check(ctx, g, np.concatenate([o, d], 1).astype(np.float32), 'octahedron')
    assert (tri >= 0).all()


def test_wide_long_leaf(ctx):
    """A root leaf of 300 triangles (beyond the 8 x 16 one node's leaf slots hold) is split in
    halves across slots and nodes."""
    tris, n, bounds, nodes = single_leaf(300)
    g = geometry_with_bvh(tris, n, bounds, nodes, 1)
    info, tri = _check(ctx, g, random_rays(3000, 35).astype(np.float32), 'single leaf')
    assert info['node_count'] >= 2 and (tri >= 0).any()


def test_wide_empty(ctx):
    g = geometry_from(np.zeros(0, np.float32), 0, oracle.build_bvh)[0]
    s = _scene(ctx, g)
    assert s.build_wide()['node_count'] == 0
    tri, t, *_ = s.intersect_closest_wide(random_rays(10, 1).astype(np.float32))
    assert (tri == -1).all() and np.isinf(t).all()
    s.release()


def test_wide_requires_build(ctx):
    from meercadraytracer_amd import native
    g = geometry_from(BVH_CASES['five']()[0], 5, oracle.build_bvh)[0]
    s = _scene(ctx, g)
    with pytest.raises(native.NativeError) as e:
        s.intersect_closest_wide(random_rays(4, 1).astype(np.float32))
    assert 'no wide BVH' in str(e.value)
    s.build_wide()
    s.set_geometry(g, 2)  # a new geometry version drops it
    with pytest.raises(native.NativeError):
        s.intersect_closest_wide(random_rays(4, 1).astype(np.float32))
    s.release()
