"""The precision form's checker (oracle.wide_brute: the FP32 watertight test of Woop, Benthin and
Wald 2013 over every triangle), pinned without the GPU: a hand-derived hit, the culling and
front/back convention of the FP64 intersectClosest (bvh.ts:233-250), agreement with the FP64
oracle on random scenes up to FP32 rounding, and watertightness on a closed surface."""
import numpy as np

import oracle
from scenes_common import geometry_from, octahedron, random_rays, soup


def _geom(tris, n, sides=2):
    return geometry_from(tris, n, oracle.build_bvh, sides)[0]


def test_known_answer():
    # v0 = (0,0,0), v1 = (1,0,0), v2 = (0,1,0); straight down through (0.25, 0.5).
    g = _geom(np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32), 1)
    rays = np.array([[0.25, 0.5, 2, 0, 0, -1], [0.25, 0.5, -2, 0, 0, 1], [2, 2, 1, 0, 0, -1]], np.float32)
    tri, t, u, v, front = oracle.wide_brute(g, rays, cull=False)
    assert tri.tolist() == [0, 0, -1]
    assert t[:2].tolist() == [2.0, 2.0] and u[:2].tolist() == [0.25, 0.25] and v[:2].tolist() == [0.5, 0.5]
    # The FP64 query: the same hits and sides (det > 0 is the front).
    o_tri, o_t, o_u, o_v, o_f, _ = oracle.intersect_closest(g, rays.astype(np.float64), cull=False)
    assert o_tri.tolist() == tri.tolist()
    assert o_f[:2].tolist() == front[:2].tolist() == [1, 0]


def test_culling_matches_fp64():
    """FrontSide (0) drops back faces, BackSide (1) front faces, DoubleSide (2) neither."""
    tris = soup(300, 21)
    rays = random_rays(4000, 22).astype(np.float32)
    for side in (0, 1, 2):
        g = _geom(tris, 300, side)
        tri, t, *_ = oracle.wide_brute(g, rays, cull=True)
        o_tri, o_t, *_ = oracle.intersect_closest(g, rays.astype(np.float64), cull=True)
        same = tri == o_tri
        assert same.mean() > 0.999, side
        assert (o_tri >= 0).sum() > 200


def test_agrees_with_fp64_up_to_rounding():
    tris = soup(20000, 11)
    g = _geom(tris, 20000)
    rays = random_rays(20000, 12).astype(np.float32)
    tri, t, u, v, front = oracle.wide_brute(g, rays, cull=False)
    o_tri, o_t, o_u, o_v, o_f, _ = oracle.intersect_closest(g, rays.astype(np.float64), cull=False)
    hits = (o_tri >= 0).sum()
    assert hits > 5000
    same = tri == o_tri
    assert (~same).sum() <= max(2, hits // 2000)
    m = same & (tri >= 0)
    err = np.abs(t[m].astype(np.float64) - o_t[m])
    # FP32 rounding of the vertices relative to the origin (|v - o| ~ t + the triangle's size 1):
    # a few units of 2^-24 of that, more for hits seen at grazing angles.
    assert np.percentile(err / o_t[m], 99.9) < 1e-5
    assert (err / (o_t[m] + 1.0)).max() < 1e-5
    assert np.abs(u[m] - o_u[m]).max() < 1e-3 and np.abs(v[m] - o_v[m]).max() < 1e-3
    assert np.array_equal(front[m], o_f[m])


def _plain_fp32_escapes(tris, rays):
    """Control: rays a plain FP32 Möller–Trumbore (not watertight) lets through the surface."""
    T = tris.reshape(-1, 9).astype(np.float32)
    v0, e1, e2 = T[:, 0:3], T[:, 3:6], T[:, 6:9]
    esc = 0
    for r in rays:
        o, d = r[:3], r[3:]
        p = np.cross(d[None], e2).astype(np.float32)
        det = (e1 * p).sum(1, dtype=np.float32)
        with np.errstate(divide='ignore', invalid='ignore'):
            inv = (np.float32(1) / det).astype(np.float32)
            s = (o[None] - v0).astype(np.float32)
            u = ((s * p).sum(1, dtype=np.float32) * inv).astype(np.float32)
            q = np.cross(s, e1).astype(np.float32)
            v = ((d[None] * q).sum(1, dtype=np.float32) * inv).astype(np.float32)
            t = ((e2 * q).sum(1, dtype=np.float32) * inv).astype(np.float32)
        esc += not ((det != 0) & (u >= 0) & (u <= 1) & (v >= 0) & (u + v <= 1) & (t > 0)).any()
    return esc


def test_watertight_closed_surface():
    """Rays from inside the closed octahedron always hit it, including rays aimed exactly at its
    edges and vertices, where a plain FP32 Möller–Trumbore lets some through."""
    tris, n = octahedron(16)
    g = _geom(tris, n)
    rng = np.random.Generator(np.random.PCG64(5))
    o = rng.uniform(-0.25, 0.25, (3000, 3))
    d = rng.normal(size=(3000, 3))
    # aim at grid points (vertices and edge points of the mesh)
    grid = rng.integers(-16, 17, (3000, 2))
    k = rng.integers(0, 3, 3000)
    target = np.zeros((3000, 3))
    for i in range(3000):
        a, b = grid[i]
        rest = 16 - abs(a) - abs(b)
        if rest < 0:
            a, b = a // 2, b // 2
            rest = 16 - abs(a) - abs(b)
        p = [a, b, rest * (1 if i % 2 else -1)]
        target[i] = np.roll(p, k[i]) / 16.0
    d2 = target - o
    dirs = np.concatenate([d, d2])
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    rays = np.concatenate([np.concatenate([o, o]), dirs], 1).astype(np.float32)
    tri, *_ = oracle.wide_brute(g, rays, cull=False)
    assert (tri >= 0).all(), f'{(tri < 0).sum()} rays escaped'
    assert _plain_fp32_escapes(g.triangles, rays[3000:]) > 0
