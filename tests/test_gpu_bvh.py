"""GPU builder and GPU ray queries against the oracle: BVH arrays and order bit-identical,
first hits exact (triangle, t, u, v, front), any-hit counts exact."""
import numpy as np
import pytest

import oracle
from helpers import probe_rays_400
from scenes_common import BVH_CASES, geometry_from, random_rays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def ctx():
    from meercadraytracer_amd import native
    return native.default_context()


@pytest.mark.parametrize('case', sorted(BVH_CASES))
def test_builder_bit_exact(ctx, case):
    tris, n = BVH_CASES[case]()
    print(f'builder case {case} n={n}', flush=True)
    gb, gn, go, gc = ctx.build_bvh(tris, n)
    ob, on, oo, oc = oracle.build_bvh(tris, n)
    assert gc == oc
    assert np.array_equal(gn, on)
    assert np.array_equal(go, oo)
    assert gb.view(np.uint32).tolist() == ob.view(np.uint32).tolist()


def test_builder_empty(ctx):
    b, nd, o, c = ctx.build_bvh(np.zeros(0, np.float32), 0)
    assert c == 0


def test_builder_scene_sizes(ctx):
    from meercadraytracer_amd import scenes
    from meercadraytracer_amd.compile import add_mesh, compile_shading
    scene, camera, _ = scenes.plain_cpu_scene(3)
    shading, meshes, index = compile_shading(scene)
    parts = [add_mesh(m, index[id(m.material)], m.material) for m in meshes]
    P = np.concatenate([p['p'] for p in parts])
    tri = np.empty((len(P), 9), np.float32)
    tri[:, 0:3] = P[:, 0]
    tri[:, 3:6] = P[:, 1] - P[:, 0]
    tri[:, 6:9] = P[:, 2] - P[:, 0]
    tri = np.ascontiguousarray(tri.reshape(-1))
    g = ctx.build_bvh(tri, len(P))
    o = oracle.build_bvh(tri, len(P))
    assert g[3] == o[3] and np.array_equal(g[1], o[1]) and np.array_equal(g[2], o[2])
    assert np.array_equal(g[0].view(np.uint32), o[0].view(np.uint32))


def _scene(ctx, g):
    from meercadraytracer_amd import native
    s = native.Scene(ctx)
    s.set_geometry(g, 1)
    return s


@pytest.mark.parametrize('case', ['lcg3000', 'coincident500', 'soup20k', 'clustered', 'single', 'quad'])
@pytest.mark.parametrize('cull', [True, False])
def test_closest_exact(ctx, case, cull):
    tris, n = BVH_CASES[case]()
    for sides in (0, 1, 2):
        g, _ = geometry_from(tris, n, oracle.build_bvh, sides=sides)
        s = _scene(ctx, g)
        rays = np.concatenate([probe_rays_400(), random_rays(3000, 1)])
        if case in ('coincident500', 'single', 'quad'):
            rng = np.random.Generator(np.random.PCG64(2))
            o = np.concatenate([rng.uniform(-0.5, 1.5, (2000, 2)), rng.choice([-3.0, 3.0], (2000, 1))], 1)
            d = np.zeros((2000, 3))
            d[:, 2] = -np.sign(o[:, 2])
            d[:, :2] = rng.uniform(-0.2, 0.2, (2000, 2))
            d /= np.linalg.norm(d, axis=1, keepdims=True)
            rays = np.concatenate([rays, np.concatenate([o, d], 1)])
        gt, gtt, gu, gv, gf = s.intersect_closest(rays, cull=cull)
        ot, ott, ou, ov, of, _ = oracle.intersect_closest(g, rays, cull=cull)
        assert np.array_equal(gt, ot)
        hit = ot >= 0
        assert np.array_equal(gtt[hit].view(np.uint64), ott[hit].view(np.uint64))
        assert np.array_equal(gu[hit].view(np.uint64), ou[hit].view(np.uint64))
        assert np.array_equal(gv[hit].view(np.uint64), ov[hit].view(np.uint64))
        assert np.array_equal(gf[hit], of[hit])
        s.release()


@pytest.mark.parametrize('case', ['lcg3000', 'coincident500', 'soup20k', 'clustered'])
def test_any_count_exact(ctx, case):
    tris, n = BVH_CASES[case]()
    g, _ = geometry_from(tris, n, oracle.build_bvh)
    s = _scene(ctx, g)
    rays = np.concatenate([probe_rays_400(), random_rays(3000, 4)])
    for tmax in (np.inf, 20.0):
        gc = s.intersect_any_count(rays, 0.0, tmax)
        oc, _ = oracle.intersect_any_count(g, rays, 0.0, tmax)
        assert np.array_equal(gc, oc)
    s.release()


def test_bvh_reference_cases_on_gpu(ctx):
    """tests/bvh.test.ts:82-148 on the GPU path (GPU-built BVH, GPU queries)."""
    from helpers import random_triangles
    raw = random_triangles(3000, 7)
    g, _ = geometry_from(raw, 3000, ctx.build_bvh)
    s = _scene(ctx, g)
    rays = probe_rays_400()
    tri, t, *_ = s.intersect_closest(rays)
    btri, bt, ball = oracle.brute_force(g.triangles, 3000, rays)
    assert np.array_equal(tri, btri)
    hit = tri >= 0
    assert hit.sum() > 100 and np.all(np.abs(t[hit] - bt[hit]) < 5e-4)
    assert np.array_equal(s.intersect_any_count(rays), ball)


def _grazing_rays(g, n, seed):
    """Rays aimed exactly at corners of BVH node boxes and at triangle vertices, some starting on a
    box face and some ending (tmax) at the corner: the cases where an FP32 slab estimate cannot
    decide and the FP64 test must."""
    rng = np.random.Generator(np.random.PCG64(seed))
    b = g.bvh.bounds.reshape(-1, 6).astype(np.float64)
    tv = g.triangles.reshape(-1, 9).astype(np.float64)[:, :3]
    idx = rng.integers(0, len(b), n)
    corner = np.where(rng.integers(0, 2, (n, 3)) == 1, b[idx, 3:], b[idx, :3])
    verts = tv[rng.integers(0, len(tv), n)]
    target = np.where(rng.integers(0, 2, (n, 1)) == 1, corner, verts)
    o = target + rng.normal(0, 5, (n, 3))
    on_face = rng.integers(0, 4, n) == 0  # start on a face of the box: one coordinate copied
    ax = rng.integers(0, 3, n)
    o[on_face, ax[on_face]] = corner[on_face, ax[on_face]]
    d = target - o
    dist = np.linalg.norm(d, axis=1)
    d /= dist[:, None]
    return np.ascontiguousarray(np.concatenate([o, d], 1)), dist


@pytest.mark.parametrize('case', ['lcg3000', 'soup20k', 'clustered', 'flat'])
def test_any_count_exact_grazing(ctx, case):
    """Any-hit traversals decide most boxes with an FP32 estimate and hand the rest to the FP64
    test; rays that graze box corners and faces must still visit exactly the reference's hits."""
    tris, n = BVH_CASES[case]()
    g, _ = geometry_from(tris, n, oracle.build_bvh)
    s = _scene(ctx, g)
    rays, dist = _grazing_rays(g, 6000, 7)
    for tmax in (np.inf, 20.0):
        gc = s.intersect_any_count(rays, 0.0, tmax)
        oc, _ = oracle.intersect_any_count(g, rays, 0.0, tmax)
        assert np.array_equal(gc, oc), (case, tmax, int((gc != oc).sum()))
    # tmax at the aimed point: the box whose corner it is sits on the boundary of the segment.
    for k in range(0, 6000, 750):
        gc = s.intersect_any_count(rays[k:k + 750], 0.0, float(dist[k]))
        oc, _ = oracle.intersect_any_count(g, rays[k:k + 750], 0.0, float(dist[k]))
        assert np.array_equal(gc, oc)
    gt, gtt, *_ = s.intersect_closest(rays, cull=False)
    ot, ott, *_ = oracle.intersect_closest(g, rays, cull=False)
    assert np.array_equal(gt, ot) and np.array_equal(gtt[ot >= 0].view(np.uint64), ott[ot >= 0].view(np.uint64))
    s.release()
