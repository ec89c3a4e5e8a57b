"""Geometry fixtures shared by the CPU (oracle) and GPU tests."""
import numpy as np

from helpers import random_triangles, reorder
from meercadraytracer_amd.types import CompiledBvh, CompiledGeometry


def geometry_from(tris, count, build, sides=2):
    """CompiledGeometry of raw triangles (reordered by the BVH), with the given builder."""
    bounds, nodes, order, nc = build(tris, count)
    data = reorder(tris, order) if count else np.zeros(0, np.float32)
    return geometry_with_bvh(data, count, bounds, nodes, nc, sides), order


def geometry_with_bvh(data, count, bounds, nodes, node_count, sides=2):
    """CompiledGeometry of triangles already in BVH order under the given BVH arrays."""
    return CompiledGeometry(
        triangleCount=count, triangles=data, normals=np.zeros(count * 9, np.float32),
        uvs=np.zeros(count * 6, np.float32), tangents=np.zeros(count * 4, np.float32),
        colors=np.zeros(0, np.float32), materials=np.zeros(count, np.uint32),
        flags=np.full(count, 1, np.uint8), sides=np.full(count, sides, np.uint8),
        bvh=CompiledBvh(bounds, nodes, node_count), epsilon=1e-5)


def coincident(count=500):  # tests/bvh.test.ts:135-142
    same = np.zeros(9 * count, np.float32)
    for i in range(count):
        same[i * 9:i * 9 + 9] = [0, 0, 0, 1, 0, 0, 0, 1, 0]
    return same


def soup(count, seed, spread=10.0, size=1.0):
    """Random triangles with a seeded numpy generator (a second family next to the LCG set)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.uniform(-spread, spread, (count, 3))
    v = c[:, None, :] + rng.uniform(-size, size, (count, 3, 3))
    t = np.empty((count, 9), np.float64)
    t[:, 0:3] = v[:, 0]
    t[:, 3:6] = v[:, 1] - v[:, 0]
    t[:, 6:9] = v[:, 2] - v[:, 0]
    return np.ascontiguousarray(t.astype(np.float32).reshape(-1))


def random_rays(n, seed, spread=30.0, target=8.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    o = rng.uniform(-spread, spread, (n, 3))
    t = rng.uniform(-target, target, (n, 3))
    d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.ascontiguousarray(np.concatenate([o, d], 1))


def octahedron(n=16):
    """Closed surface |x|+|y|+|z| = 1, each face cut into n*n triangles on a grid of step 1/n.
    Coordinates and edge vectors are exact in float32, so triangles sharing an edge rebuild it
    bit for bit and a watertight test lets no ray from inside escape."""
    out = []
    for sx in (-1, 1):
        for sy in (-1, 1):
            for sz in (-1, 1):
                a, b, c = np.array([sx, 0, 0.]), np.array([0, sy, 0.]), np.array([0, 0, sz * 1.])
                if sx * sy * sz < 0:
                    b, c = c, b

                def p(i, j):
                    return a + (b - a) * i / n + (c - a) * j / n
                for i in range(n):
                    for j in range(n - i):
                        out.append((p(i, j), p(i + 1, j), p(i, j + 1)))
                        if i + j < n - 1:
                            out.append((p(i + 1, j), p(i + 1, j + 1), p(i, j + 1)))
    v = np.array(out)
    t = np.empty((len(v), 9))
    t[:, 0:3] = v[:, 0]
    t[:, 3:6] = v[:, 1] - v[:, 0]
    t[:, 6:9] = v[:, 2] - v[:, 0]
    return np.ascontiguousarray(t.astype(np.float32).reshape(-1)), len(v)


def single_leaf(count=100, seed=4):
    """A soup under a hand-made BVH of one root leaf holding every triangle, as buildBvh leaves at
    its depth limit (bvh.ts:82-85): a leaf of more than 16 triangles."""
    tris = soup(count, seed)
    v = tris.reshape(-1, 3, 3).astype(np.float64)
    p = np.stack([v[:, 0], v[:, 0] + v[:, 1], v[:, 0] + v[:, 2]], 1)
    bounds = np.concatenate([p.min((0, 1)), p.max((0, 1))]).astype(np.float32)
    nodes = np.array([0, count, 0], np.int32)
    return tris, count, bounds, nodes


BVH_CASES = {
    'lcg3000': lambda: (random_triangles(3000, 7), 3000),
    'coincident500': lambda: (coincident(500), 500),
    'single': lambda: (np.array([-1, -1, 0, 2, 0, 0, 0, 2, 0], np.float32), 1),
    'quad': lambda: (np.array([-1, -1, 0, 2, 0, 0, 2, 2, 0, -1, -1, 0, 2, 2, 0, 0, 2, 0], np.float32), 2),
    'five': lambda: (soup(5, 3), 5),
    'soup20k': lambda: (soup(20000, 11), 20000),
    'clustered': lambda: (soup(4000, 5, spread=0.01, size=0.5), 4000),
    'flat': lambda: (np.ascontiguousarray(np.where(np.arange(9 * 3000) % 3 == 2, 0.0, soup(3000, 9)).astype(np.float32)), 3000),
}
