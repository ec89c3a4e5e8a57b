"""Seeded scene generators for the benchmark and parity workloads.

  plain_cpu_scene(seed)   8 icospheres (4-way subdivision 3: 1,280 triangles each, centres
                          U[-1,1]^3, radii U[0.2,0.4]) and a 2-triangle 20x20 ground plane at
                          y = -1.2; MeshStandardMaterial with albedo U[0.2,0.9]^3, roughness
                          U[0.1,0.9], metalness Bernoulli(0.3) in {0,1}; one directional and one
                          point light; perspective camera, fov 50, at (0,1,4) looking at the origin.
                          10,242 triangles. Rendered at 256x256, 16 spp, 4 bounces, GI off, ACES.
  plain_gpu_scene(seed)   A 724x724-quad heightfield (1,048,352 triangles) over 40x40 with fBm
                          displacement of amplitude 0.1, plus 64 icospheres (subdivision 4: 5,120
                          triangles each) at seeded positions above it, 1,376,032 triangles in all;
                          the same material distributions with 20% glass (MeshPhysicalMaterial,
                          transmission 1, ior 1.5, thickness = diameter); a sun and 4 point lights;
                          perspective camera, fov 50. Rendered at 1920x1080, 64 spp, GI on, 8 bounces.
  textured_gpu_scene(seed) plain_gpu_scene(seed) with textures and an environment: the terrain
                          gets a repeated 512x512 sRGB colour map, a 256x256 normal map and a
                          roughness map from fBm noise; non-glass spheres get a 64x32 sRGB map
                          over spherical uvs; a 256x128 HDR sky is the scene environment and the
                          background. Rendered like plain_gpu.
Every value is drawn from numpy's PCG64 with the given seed, so a seed names the scene.
"""
import math

import numpy as np

from . import abi
from . import three as T
from .compile import compile_camera, compile_scene


def _standard_material(rng):
    albedo = rng.uniform(0.2, 0.9, 3)
    return T.MeshStandardMaterial(color=T.Color(*albedo), roughness=float(rng.uniform(0.1, 0.9)),
                                  metalness=float(rng.random() < 0.3))


def _glass_material(rng, diameter):
    albedo = rng.uniform(0.2, 0.9, 3)
    return T.MeshPhysicalMaterial(color=T.Color(*albedo), roughness=float(rng.uniform(0.1, 0.9)),
                                  metalness=0.0, transmission=1.0, ior=1.5, thickness=float(diameter))


def plain_cpu_scene(seed=0, width=256, height=256):
    rng = np.random.Generator(np.random.PCG64(seed))
    scene = T.Scene()
    for _ in range(8):
        centre = rng.uniform(-1, 1, 3)
        radius = float(rng.uniform(0.2, 0.4))
        m = T.Mesh(T.IcosphereGeometry(radius, 3), _standard_material(rng))
        m.position = [float(c) for c in centre]
        scene.add(m)
    ground = T.Mesh(T.PlaneGeometry(20, 20), _standard_material(rng))
    ground.rotation = [-math.pi / 2, 0.0, 0.0]
    ground.position = [0.0, -1.2, 0.0]
    scene.add(ground)
    sun = T.DirectionalLight(0xffffff, float(rng.uniform(1.5, 3.0)))
    sun.position = [float(v) for v in rng.uniform([-5, 4, -5], [5, 8, 5])]
    point = T.PointLight(0xffffff, float(rng.uniform(5, 15)), 0, 2)
    point.position = [float(v) for v in rng.uniform([-2, 1.5, -2], [2, 3, 2])]
    scene.add(sun, point)
    camera = T.PerspectiveCamera(50, width / height, 0.1, 100)
    camera.position = [0.0, 1.0, 4.0]
    camera.lookAt(0, 0, 0)
    camera.updateMatrixWorld()
    options = dict(samples=16, maxBounces=4, globalIllumination=False,
                   toneMapping=abi.ACESFilmicToneMapping, outputColorSpace='srgb', seed=0)
    return scene, camera, options


def fbm_heights(rng, rows, cols, amplitude=0.1, octaves=5, lacunarity=2.0, gain=0.5, base_cells=4):
    """Value-noise fBm on a rows x cols grid: `octaves` layers of bilinearly interpolated random
    lattices, each `lacunarity` times finer and `gain` times weaker; peak amplitude `amplitude`."""
    h = np.zeros((rows, cols))
    total = 0.0
    amp = 1.0
    cells = base_cells
    ys = np.linspace(0, 1, rows)
    xs = np.linspace(0, 1, cols)
    for _ in range(octaves):
        lattice = rng.uniform(-1, 1, (cells + 1, cells + 1))
        gy, gx = ys * cells, xs * cells
        iy = np.minimum(gy.astype(int), cells - 1)
        ix = np.minimum(gx.astype(int), cells - 1)
        fy = (gy - iy)[:, None]
        fx = (gx - ix)[None, :]
        a = lattice[iy][:, ix]
        b = lattice[iy][:, ix + 1]
        c = lattice[iy + 1][:, ix]
        d = lattice[iy + 1][:, ix + 1]
        h += amp * ((a * (1 - fx) + b * fx) * (1 - fy) + (c * (1 - fx) + d * fx) * fy)
        total += amp
        amp *= gain
        cells = int(cells * lacunarity)
    return h * (amplitude / total)


def plain_gpu_scene(seed=0, width=1920, height=1080, quads=724, spheres=64, subdivisions=4):
    rng = np.random.Generator(np.random.PCG64(seed))
    scene = T.Scene()
    size = 40.0
    heights = fbm_heights(rng, quads + 1, quads + 1)
    terrain = T.Mesh(T.HeightfieldGeometry(heights, size, size), _standard_material(rng))
    scene.add(terrain)
    for _ in range(spheres):
        radius = float(rng.uniform(0.25, 0.6))
        pos = [float(rng.uniform(-8, 8)), float(rng.uniform(0.2, 2.5)) + radius, float(rng.uniform(-8, 4))]
        mat = _glass_material(rng, 2 * radius) if rng.random() < 0.2 else _standard_material(rng)
        m = T.Mesh(T.IcosphereGeometry(radius, subdivisions), mat)
        m.position = pos
        scene.add(m)
    sun = T.DirectionalLight(0xfff4e0, 3.0)
    sun.position = [4.0, 10.0, 3.0]
    scene.add(sun)
    for _ in range(4):
        p = T.PointLight(0xffffff, float(rng.uniform(5, 20)), 0, 2)
        p.position = [float(rng.uniform(-8, 8)), float(rng.uniform(2, 5)), float(rng.uniform(-8, 4))]
        scene.add(p)
    camera = T.PerspectiveCamera(50, width / height, 0.1, 200)
    camera.position = [0.0, 3.5, 13.0]
    camera.lookAt(0, 0.6, -2)
    camera.updateMatrixWorld()
    options = dict(samples=64, maxBounces=8, globalIllumination=True,
                   toneMapping=abi.ACESFilmicToneMapping, outputColorSpace='srgb', seed=0)
    return scene, camera, options


def _noise_rgba(rng, size, lo, hi, srgb):
    """A seeded RGBA8 texture of fBm noise: each channel a separate field mapped to [lo, hi]."""
    chans = [fbm_heights(rng, size, size, amplitude=1.0) for _ in range(3)]
    data = np.empty((size, size, 4), np.uint8)
    for c, f in enumerate(chans):
        f = (f - f.min()) / max(f.max() - f.min(), 1e-12)
        data[..., c] = np.clip(np.round((lo + (hi - lo) * f) * 255), 0, 255).astype(np.uint8)
    data[..., 3] = 255
    tex = T.DataTexture(data.reshape(-1), size, size)
    if srgb:
        tex.colorSpace = T.SRGBColorSpace
    return tex


def _sky(rng, width=256, height=128):
    """An HDR equirect sky: a vertical gradient, a warm sun disc and seeded cloud noise."""
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    U, V = np.meshgrid(u, v)
    clouds = fbm_heights(rng, height, width, amplitude=1.0)
    clouds = (clouds - clouds.min()) / max(clouds.max() - clouds.min(), 1e-12)
    sky = np.stack([0.35 + 0.5 * V, 0.5 + 0.4 * V, 0.9 + 0.3 * V], -1) * (0.6 + 0.4 * clouds[..., None])
    sun = np.exp(-(((U - 0.3) / 0.02) ** 2 + ((V - 0.75) / 0.03) ** 2)) * 40.0
    sky = sky + sun[..., None] * np.array([1.0, 0.9, 0.7])
    data = np.concatenate([sky, np.ones((height, width, 1))], -1).astype(np.float32)
    tex = T.DataTexture(data.reshape(-1), width, height, T.FloatType)
    tex.mapping = T.EquirectangularReflectionMapping
    return tex


def textured_gpu_scene(seed=0, width=1920, height=1080, quads=724, spheres=64, subdivisions=4):
    scene, camera, options = plain_gpu_scene(seed, width, height, quads, spheres, subdivisions)
    rng = np.random.Generator(np.random.PCG64(seed + 1_000_003))
    terrain = scene.children[0]
    colour = _noise_rgba(rng, 512, 0.25, 0.95, True)
    colour.wrapS = colour.wrapT = abi.RepeatWrapping
    colour.repeat = [16.0, 16.0]
    normal = _noise_rgba(rng, 256, 0.3, 0.7, False)
    normal.wrapS = normal.wrapT = abi.RepeatWrapping
    normal.repeat = [32.0, 32.0]
    rough = _noise_rgba(rng, 256, 0.4, 1.0, False)
    rough.wrapS = rough.wrapT = abi.MirroredRepeatWrapping
    rough.repeat = [8.0, 8.0]
    terrain.material.map, terrain.material.normalMap, terrain.material.roughnessMap = colour, normal, rough
    stripes = _noise_rgba(rng, 64, 0.2, 1.0, True)
    stripes.wrapS = abi.RepeatWrapping
    for m in scene.children[1:]:
        if isinstance(m, T.Mesh) and getattr(m.material, 'transmission', 0) == 0:
            g = m.geometry
            p = g.position / np.linalg.norm(g.position, axis=1, keepdims=True)
            g.uv = np.stack([np.arctan2(p[:, 2], p[:, 0]) / (2 * math.pi) + 0.5,
                             np.arcsin(np.clip(p[:, 1], -1, 1)) / math.pi + 0.5], 1).astype(np.float32)
            m.material.map = stripes
    sky = _sky(rng)
    scene.environment = sky
    scene.background = sky
    return scene, camera, options


def sized_gpu_scene(seed, triangles):
    """The plain_gpu generator scaled to about `triangles` triangles, in the same proportions as
    the 1.38M-triangle scene: about 76% heightfield and 24% icospheres (subdivision 4, or 3 below
    400k triangles)."""
    sub = 4 if triangles >= 400_000 else 3
    per = 20 * 4 ** sub
    spheres = max(1, int(round(0.24 * triangles / per)))
    quads = max(8, int(round(math.sqrt(max(triangles - spheres * per, 128) / 2))))
    scene, _, options = plain_gpu_scene(seed, 64, 64, quads, spheres, sub)
    return scene, options


def orbit_camera(width, height, angle=0.0, radius=15.0):
    """The plain_gpu camera moved round the scene centre (0, 0.6, -2) by `angle` radians."""
    camera = T.PerspectiveCamera(50, width / height, 0.1, 200)
    camera.position = [radius * math.sin(angle), 3.5, -2 + radius * math.cos(angle)]
    camera.lookAt(0, 0.6, -2)
    camera.updateMatrixWorld()
    return camera


def compiled(kind, seed=0, build_bvh=None, **kw):
    """(CompiledScene, CompiledCamera, options) of a named workload."""
    maker = {'plain_cpu': plain_cpu_scene, 'plain_gpu': plain_gpu_scene, 'textured_gpu': textured_gpu_scene}[kind]
    scene, camera, options = maker(seed, **kw)
    return compile_scene(scene, build_bvh), compile_camera(camera), options
