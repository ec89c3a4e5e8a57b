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


def _data_texture(T, rng, w, h, kind='u8', srgb=False, wrap=None, repeat=None, lo=0.0, hi=1.0):
    if kind == 'u8':
        tex = T.DataTexture(rng.integers(0, 256, (h * w * 4,), dtype=np.uint8), w, h)
    else:
        tex = T.DataTexture(rng.uniform(lo, hi, (h * w * 4,)).astype(np.float32), w, h, T.FloatType)
    if srgb:
        tex.colorSpace = T.SRGBColorSpace
    if wrap is not None:
        tex.wrapS = tex.wrapT = wrap
    if repeat is not None:
        tex.repeat = list(repeat)
        tex.updateMatrix()
    return tex


def textured_scene(background='environment', seed=0):
    """A scene that takes every texture and environment branch of the tracer (tracer.ts:119-406):
    a physical floor with colour (sRGB, repeated), normal, roughness, metalness and ao maps; a
    transparent textured glass sphere (coloured, textured shadows); an alpha-mapped cut-out panel
    (cut-outs in camera and shadow rays); an emissive-mapped lambert box; lambert, phong and basic
    boxes with equirect and cube envMaps under the Multiply, Mix and Add combines; a rotated HDR
    scene environment (the physical materials' reflections, misses after a bounce and the
    environment-diffuse ray); directional, point and ambient light. background: 'environment'
    (the scene environment), 'cube', 'screen' (a screen-space texture) or 'color'."""
    import math
    from meercadraytracer_amd import abi
    from meercadraytracer_amd import three as T
    rng = np.random.Generator(np.random.PCG64(seed))
    s = T.Scene()
    env = _data_texture(T, rng, 16, 8, 'f32', lo=0.2, hi=2.5)
    env.mapping = T.EquirectangularReflectionMapping
    s.environment = env
    s.environmentIntensity = 0.7
    s.environmentRotation = (0.0, 0.6, 0.0)
    faces = [{'data': rng.integers(0, 256, (4 * 4 * 4,), dtype=np.uint8), 'width': 4, 'height': 4} for _ in range(6)]
    cube = T.CubeTexture(faces)
    cube.colorSpace = T.SRGBColorSpace
    if background == 'environment':
        s.background = env
        s.backgroundIntensity = 1.3
    elif background == 'cube':
        s.background = cube
    elif background == 'screen':
        s.background = _data_texture(T, rng, 5, 3, 'u8', srgb=True)
        s.backgroundIntensity = 0.8
    else:
        s.background = T.Color(0.1, 0.2, 0.3)

    floor = T.Mesh(T.PlaneGeometry(10, 10, 3, 3), T.MeshStandardMaterial(
        color=0xddccbb, roughness=0.7, metalness=0.6,
        map=_data_texture(T, rng, 8, 8, 'u8', srgb=True, wrap=abi.RepeatWrapping, repeat=(3, 2)),
        normalMap=_data_texture(T, rng, 4, 4, 'u8', wrap=abi.MirroredRepeatWrapping, repeat=(2, 2)),
        normalScale=(0.7, -0.5),
        roughnessMap=_data_texture(T, rng, 4, 4, 'u8'), metalnessMap=_data_texture(T, rng, 4, 4, 'u8'),
        aoMap=_data_texture(T, rng, 4, 4, 'u8'), aoMapIntensity=0.8))
    floor.rotation = [-math.pi / 2, 0.0, 0.0]
    floor.receiveShadow = True
    floor.castShadow = True

    sphere_geo = T.IcosphereGeometry(1.0, 2)
    p = sphere_geo.position / np.linalg.norm(sphere_geo.position, axis=1, keepdims=True)
    sphere_geo.uv = np.stack([np.arctan2(p[:, 2], p[:, 0]) / (2 * math.pi) + 0.5,
                              np.arcsin(np.clip(p[:, 1], -1, 1)) / math.pi + 0.5], 1).astype(np.float32)
    glass = T.Mesh(sphere_geo, T.MeshPhysicalMaterial(
        color=0xffffff, roughness=0.1, transmission=0.8, thickness=1.0, transparent=True, opacity=0.7,
        map=_data_texture(T, rng, 4, 4, 'f32', wrap=abi.RepeatWrapping), attenuationColor=0xaaffcc,
        attenuationDistance=2.0))
    glass.position = [-1.5, 1.0, 0.5]
    glass.castShadow = True
    glass.receiveShadow = True

    stripes = np.zeros((2, 4, 4), np.uint8)
    stripes[:, :, 3] = 255
    stripes[:, 1::2, 1] = 255
    panel = T.Mesh(T.BoxGeometry(2.0, 2.0, 0.1), T.MeshStandardMaterial(
        color=0x88aaff, roughness=0.4, alphaMap=T.DataTexture(stripes.reshape(-1), 4, 2), alphaTest=0.5,
        side=abi.DoubleSide))
    panel.position = [1.2, 1.2, -1.0]
    panel.rotation = [0.0, 0.5, 0.0]
    panel.castShadow = True

    glow = T.Mesh(T.BoxGeometry(0.6, 0.6, 0.6), T.MeshLambertMaterial(
        color=0x666666, emissive=0xffaa33, emissiveIntensity=1.5, emissiveMap=_data_texture(T, rng, 2, 2, 'u8')))
    glow.position = [0.3, 0.3, 1.6]
    glow.castShadow = True

    eq_env = _data_texture(T, rng, 8, 4, 'f32', lo=0.0, hi=1.5)
    eq_env.mapping = T.EquirectangularReflectionMapping
    boxes = []
    for i, (cls, combine, refl, envmap) in enumerate([(T.MeshLambertMaterial, abi.MultiplyOperation, 0.7, eq_env),
                                                     (T.MeshPhongMaterial, abi.MixOperation, 0.6, cube),
                                                     (T.MeshBasicMaterial, abi.AddOperation, 0.3, eq_env)]):
        b = T.Mesh(T.BoxGeometry(0.7, 0.7, 0.7), cls(color=[0xcc8844, 0x44cc88, 0x8844cc][i], envMap=envmap,
                                                   combine=combine, reflectivity=refl))
        b.position = [-2.2 + 1.1 * i, 0.35, 2.4]
        b.rotation = [0.3 * i, 0.7, 0.0]
        b.castShadow = True
        b.receiveShadow = True
        boxes.append(b)

    sun = T.DirectionalLight(0xfff4e0, 2.0)
    sun.position = [3.0, 6.0, 2.0]
    sun.castShadow = True
    bulb = T.PointLight(0x88aaff, 6.0, 0.0, 2.0)
    bulb.position = [-1.0, 2.5, 2.5]
    bulb.castShadow = True
    s.add(floor, glass, panel, glow, *boxes, sun, bulb, T.AmbientLight(0x202020, 1.0))
    s.updateMatrixWorld()
    camera = T.PerspectiveCamera(55, 1.5, 0.1, 100)
    camera.position = [0.5, 3.5, 7.5]
    camera.lookAt(0, 0.5, 0)
    camera.updateMatrixWorld()
    return s, camera
