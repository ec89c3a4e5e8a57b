"""Scene compilation: the three.js-shaped scene graph to the plain-data format.

Restates src/compile.ts (meshes, materials with their maps, lights, backgrounds, environments,
fog) and the TextureTable of src/textures.ts, vectorised with numpy in the reference's FP64 order
of operations:
  collect_objects   compile.ts:64-80
  compile_material  compile.ts:94-147
  compile_light     compile.ts:149-177
  add_mesh          compile.ts:212-306 (world transform, normal matrix, mirrored winding, flat
                    normals)
  compile_geometry  compile.ts:327-403 (triangle soup v0/e1/e2, BVH build, reorder by `order`,
                    per-triangle tangents, epsilon)
  compile_camera    compile.ts:536-557
The BVH builder is a parameter: the default is the native GPU builder (mcr_build_bvh).
"""
import math
import weakref

import numpy as np

from . import abi
from .types import (CompiledBvh, CompiledCamera, CompiledEnvironment, CompiledFog, CompiledGeometry, CompiledLight,
                    CompiledTexture,
                    CompiledMaterial, CompiledScene, ShadingPayload)


def js_hypot(*vals):
    """Math.hypot as V8 evaluates it (largest-magnitude scaling, Kahan-compensated sum of
    squares), elementwise over numpy arrays or scalars."""
    arrs = [np.abs(np.asarray(v, dtype=np.float64)) for v in vals]
    m = arrs[0]
    for a in arrs[1:]:
        m = np.maximum(m, a)
    safe = np.where(m == 0, 1.0, m)
    s = np.zeros_like(safe)
    comp = np.zeros_like(safe)
    for a in arrs:
        n = a / safe
        summand = n * n - comp
        pre = s + summand
        comp = (pre - s) - summand
        s = pre
    out = np.sqrt(s) * safe
    out = np.where(m == 0, 0.0, out)
    out = np.where(np.isinf(m), np.inf, out)
    return out if out.ndim else float(out)


def _normalize3(v):  # compile.ts:52-55
    length = js_hypot(*v)
    return (v[0] / length, v[1] / length, v[2] / length) if length > 0 else (0.0, 0.0, 1.0)


def _rgb(color, scale=1.0):  # compile.ts:43-45
    if color is None:
        return (scale, scale, scale)
    return (color.r * scale, color.g * scale, color.b * scale)


def collect_objects(scene):  # compile.ts:64-80
    meshes, lights = [], []

    def visit(o):
        if not o.visible:
            return
        from .three import Light, Mesh
        if isinstance(o, Light):
            lights.append(o)
        elif isinstance(o, Mesh):
            meshes.append(o)
        for c in o.children:
            visit(c)
    visit(scene)
    return meshes, lights


def _volume_attenuation(m):  # compile.ts:86-92
    distance = getattr(m, 'attenuationDistance', math.inf)
    colour = getattr(m, 'attenuationColor', None)
    if colour is None or not (distance > 0) or distance == math.inf:
        return (0.0, 0.0, 0.0)

    def coeff(c):
        return -math.log(min(max(c, 1e-6), 1)) / distance
    return (coeff(colour.r), coeff(colour.g), coeff(colour.b))


IDENTITY3 = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)


def _half_to_float(h):  # color.ts:28-35
    h = np.asarray(h, dtype=np.uint32)
    sign = np.where(h & 0x8000, -1.0, 1.0)
    e = (h >> 10) & 0x1f
    f = (h & 0x3ff).astype(np.float64)
    sub = sign * 2.0 ** -14 * (f / 1024)
    inf = np.where(f != 0, np.nan, sign * np.inf)
    norm = sign * 2.0 ** (e.astype(np.float64) - 15) * (1 + f / 1024)
    return np.where(e == 0, sub, np.where(e == 0x1f, inf, norm))


def _read_pixels(image, type_):
    """textures.ts:63-105 for RGBA data: uint8 stays 8-bit; float and half float become float32."""
    data = np.asarray(image['data'])
    w, h = image['width'], image['height']
    if type_ == 1016:  # HalfFloatType
        data = _half_to_float(data).astype(np.float32)
    elif data.dtype != np.uint8:
        data = data.astype(np.float32)
    if data.size != w * h * 4:
        return None
    return np.ascontiguousarray(data.reshape(-1))


_pixel_cache = {}  # (id(texture), face) -> (weak reference to the texture, version, pixels)


def _cached_pixels(texture, face, read):
    """textures.ts:107-113: one decoded pixel array per texture (or cube face) and version, so a
    scene recompiled unchanged hands out the same arrays and keeps its texture ids."""
    key = (id(texture), face)
    version = getattr(texture, 'version', 0)
    hit = _pixel_cache.get(key)
    if hit is not None and hit[0]() is texture and hit[1] == version:
        return hit[2]
    pixels = read()
    _pixel_cache[key] = (weakref.ref(texture, lambda _r, k=key: _pixel_cache.pop(k, None)), version, pixels)
    return pixels


def _euler_matrix_transposed(euler):  # textures.ts:204-217 (XYZ order)
    if euler is None or tuple(euler) == (0, 0, 0):
        return IDENTITY3
    x, y, z = euler

    def mul(a, b):
        return [a[r] * b[c * 3] + a[3 + r] * b[c * 3 + 1] + a[6 + r] * b[c * 3 + 2] for c in range(3) for r in range(3)]
    rx = [1, 0, 0, 0, math.cos(x), math.sin(x), 0, -math.sin(x), math.cos(x)]
    ry = [math.cos(y), 0, -math.sin(y), 0, 1, 0, math.sin(y), 0, math.cos(y)]
    rz = [math.cos(z), math.sin(z), 0, -math.sin(z), math.cos(z), 0, 0, 0, 1]
    m = mul(mul(rx, ry), rz)
    return (m[0], m[3], m[6], m[1], m[4], m[7], m[2], m[5], m[8])


class TextureTable:  # textures.ts:134-202
    def __init__(self):
        self.textures, self.environments = [], []
        self._index, self._env_index = {}, {}

    def add(self, texture):
        if texture is None or not getattr(texture, 'isTexture', False) or texture.isCubeTexture:
            return -1
        if id(texture) in self._index:
            return self._index[id(texture)]
        if texture.matrixAutoUpdate is not False:
            texture.updateMatrix()
        pixels = _cached_pixels(texture, None, lambda: _read_pixels(texture.image, texture.type))
        index = -1
        if pixels is not None:
            index = len(self.textures)
            self.textures.append(CompiledTexture(
                texture.image['width'], texture.image['height'], pixels,
                srgb=pixels.dtype == np.uint8 and texture.colorSpace == 'srgb', flipY=texture.flipY is True,
                wrapS=texture.wrapS, wrapT=texture.wrapT, matrix=tuple(texture.matrix)))
        self._index[id(texture)] = index
        return index

    def add_environment(self, texture, intensity, rotation):
        if texture is None or not getattr(texture, 'isTexture', False):
            return -1
        matrix = _euler_matrix_transposed(rotation)
        key = (id(texture), texture.version, intensity, matrix)
        if key in self._env_index:
            return self._env_index[key]
        env = None
        if texture.isCubeTexture or texture.mapping in (301, 302):
            faces = []
            for f, image in enumerate(texture.image):
                pixels = _cached_pixels(texture, f, lambda image=image: _read_pixels(image, texture.type))
                if pixels is None:
                    break
                faces.append(len(self.textures))
                self.textures.append(CompiledTexture(image['width'], image['height'], pixels,
                                                     srgb=pixels.dtype == np.uint8 and texture.colorSpace == 'srgb',
                                                     wrapS=texture.wrapS, wrapT=texture.wrapT))
            if len(faces) == 6:
                env = CompiledEnvironment('cube', faces=tuple(faces), flipX=True, rotation=matrix, intensity=intensity)
        elif texture.mapping in (303, 304):
            index = self.add(texture)
            if index >= 0:
                env = CompiledEnvironment('equirect', texture=index, rotation=matrix, intensity=intensity)
        index = -1
        if env is not None:
            index = len(self.environments)
            self.environments.append(env)
        self._env_index[key] = index
        return index


def compile_material(m, table=None):  # compile.ts:94-147
    if m.isMeshStandardMaterial:
        model = 'physical'
    elif m.isMeshPhongMaterial:
        model = 'phong'
    elif m.isMeshLambertMaterial:
        model = 'lambert'
    elif m.isMeshToonMaterial:
        model = 'toon'
    elif m.isMeshBasicMaterial:
        model = 'basic'
    else:
        model = 'lambert'
    physical = m.isMeshPhysicalMaterial
    emissive = getattr(m, 'emissive', None)
    spec = getattr(m, 'specular', None)
    return CompiledMaterial(
        model=model, side=m.side, color=_rgb(m.color), opacity=m.opacity,
        transparent=m.transparent is True, alphaTest=m.alphaTest,
        vertexColors=m.vertexColors is True, flatShading=m.flatShading is True,
        toneMapped=m.toneMapped is not False,
        emissive=_rgb(emissive, getattr(m, 'emissiveIntensity', 1.0)) if emissive is not None else (0.0, 0.0, 0.0),
        roughness=getattr(m, 'roughness', 1.0), metalness=getattr(m, 'metalness', 0.0),
        physical=physical, ior=getattr(m, 'ior', 1.5),
        specularIntensity=getattr(m, 'specularIntensity', 1.0),
        specularColor=_rgb(m.specularColor) if physical else (1.0, 1.0, 1.0),
        transmission=getattr(m, 'transmission', 0.0) if physical else 0.0,
        thickness=getattr(m, 'thickness', 0.0) if physical else 0.0,
        attenuation=_volume_attenuation(m) if physical else (0.0, 0.0, 0.0),
        shininess=getattr(m, 'shininess', 30.0),
        specular=_rgb(spec) if spec is not None else (0.0666, 0.0666, 0.0666),
        envMapIntensity=getattr(m, 'envMapIntensity', 1.0),
        reflectivity=getattr(m, 'reflectivity', 1.0), combine=getattr(m, 'combine', 0),
        **_material_maps(m, table if table is not None else TextureTable()))


def _material_maps(m, table):  # the texture fields of compile.ts:108-145
    env = getattr(m, 'envMap', None)
    ns = getattr(m, 'normalScale', None)
    return dict(map=table.add(getattr(m, 'map', None)), alphaMap=table.add(getattr(m, 'alphaMap', None)),
                emissiveMap=table.add(getattr(m, 'emissiveMap', None)),
                normalMap=table.add(getattr(m, 'normalMap', None)), aoMap=table.add(getattr(m, 'aoMap', None)),
                roughnessMap=table.add(getattr(m, 'roughnessMap', None)),
                metalnessMap=table.add(getattr(m, 'metalnessMap', None)),
                normalScale=(float(ns[0]), float(ns[1])) if ns is not None else (1.0, 1.0),
                aoMapIntensity=getattr(m, 'aoMapIntensity', 1.0),
                envMap=table.add_environment(env, getattr(m, 'envMapIntensity', 1.0),
                                             getattr(m, 'envMapRotation', None)) if env is not None else -1)


def compile_light(light):  # compile.ts:149-177
    color = _rgb(light.color, light.intensity)
    pos = tuple(light.matrixWorld[12:15])
    if light.kind == 'ambient':
        return CompiledLight('ambient', color=color)
    if light.kind == 'hemisphere':
        return CompiledLight('hemisphere', color=color, groundColor=_rgb(light.groundColor, light.intensity),
                             direction=_normalize3(pos))
    if light.kind in ('directional', 'spot'):
        tgt = light.target
        if tgt.parent is None:
            tgt.updateMatrixWorld()
        t = tgt.matrixWorld[12:15]
        direction = _normalize3((pos[0] - t[0], pos[1] - t[1], pos[2] - t[2]))
        if light.kind == 'directional':
            return CompiledLight('directional', color=color, direction=direction,
                                 castShadow=light.castShadow is True)
        return CompiledLight('spot', color=color, position=pos, direction=direction,
                             distance=light.distance, decay=light.decay,
                             coneCos=math.cos(light.angle),
                             penumbraCos=math.cos(light.angle * (1 - light.penumbra)),
                             castShadow=light.castShadow is True)
    if light.kind == 'point':
        return CompiledLight('point', color=color, position=pos, distance=light.distance,
                             decay=light.decay, castShadow=light.castShadow is True)
    return None


def _normal_matrix(m):  # compile.ts:192-205
    a, b, c, d, e, f, g, h, i = m[0], m[4], m[8], m[1], m[5], m[9], m[2], m[6], m[10]
    A = e * i - f * h
    B = -(d * i - f * g)
    Cc = d * h - e * g
    det = a * A + b * B + c * Cc
    s = 1 / det if det != 0 else 0
    n = [A * s, (-(b * i - c * h)) * s, (b * f - c * e) * s,
         B * s, (a * i - c * g) * s, (-(a * f - c * d)) * s,
         Cc * s, (-(a * h - b * g)) * s, (a * e - b * d) * s]
    return n, det


def add_mesh(mesh, mat_index, mat):
    """compile.ts:212-306 for one mesh with one material: world-space positions (FP64), normals,
    uvs and colours per triangle. Returns None when the mesh contributes nothing."""
    geo = mesh.geometry
    if geo is None or geo.position is None or not mat.visible or mat.wireframe:
        return None
    world = mesh.matrixWorld
    nm, det = _normal_matrix(world)
    pos = geo.position.astype(np.float64)
    idx = geo.index if geo.index is not None else np.arange(len(pos))
    ntri = len(idx) // 3
    idx = idx[:ntri * 3].reshape(ntri, 3).copy()
    if det < 0:
        idx[:, [1, 2]] = idx[:, [2, 1]]
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    wx = world[0] * x + world[4] * y + world[8] * z + world[12]
    wy = world[1] * x + world[5] * y + world[9] * z + world[13]
    wz = world[2] * x + world[6] * y + world[10] * z + world[14]
    wp = np.stack([wx, wy, wz], 1)
    p = wp[idx]  # (ntri, 3, 3) FP64
    if geo.normal is not None and not mat.flatShading:
        nv = geo.normal.astype(np.float64)
        nx = nm[0] * nv[:, 0] + nm[3] * nv[:, 1] + nm[6] * nv[:, 2]
        ny = nm[1] * nv[:, 0] + nm[4] * nv[:, 1] + nm[7] * nv[:, 2]
        nz = nm[2] * nv[:, 0] + nm[5] * nv[:, 1] + nm[8] * nv[:, 2]
        length = js_hypot(nx, ny, nz)
        length = np.where((length == 0) | np.isnan(length), 1.0, length)
        n = np.stack([nx / length, ny / length, nz / length], 1)[idx]
    else:
        ex, ey, ez = (p[:, 1, k] - p[:, 0, k] for k in range(3))
        fx, fy, fz = (p[:, 2, k] - p[:, 0, k] for k in range(3))
        cx, cy, cz = ey * fz - ez * fy, ez * fx - ex * fz, ex * fy - ey * fx
        length = js_hypot(cx, cy, cz)
        ok = length > 0
        safe = np.where(ok, length, 1.0)
        face = np.stack([np.where(ok, cx / safe, 0.0), np.where(ok, cy / safe, 0.0),
                         np.where(ok, cz / safe, 1.0)], 1)
        n = np.repeat(face[:, None, :], 3, axis=1)
    if geo.uv is not None:
        uv = geo.uv.astype(np.float64)[idx]
    else:
        uv = np.zeros((ntri, 3, 2))
    colors = None
    if mat.vertexColors and geo.color is not None:
        c = geo.color.astype(np.float64)
        if geo.color_size == 3:
            c = np.concatenate([c, np.ones((len(c), 1))], 1)
        colors = c[idx]
    opaque = mat.transparent is not True and not (mat.alphaTest > 0) and not (getattr(mat, 'transmission', 0) > 0)
    flags = ((abi.TRI_CAST_SHADOW if mesh.castShadow else 0) |
             (abi.TRI_RECEIVE_SHADOW if mesh.receiveShadow else 0) | (abi.TRI_OPAQUE if opaque else 0))
    return dict(p=p, n=n, uv=uv, colors=colors, material=mat_index, flags=flags, side=mat.side)


def _write_tangents(tri, uvs):  # compile.ts:406-432, vectorised
    t = tri.astype(np.float64)
    u = uvs.astype(np.float64)
    e1x, e1y, e1z, e2x, e2y, e2z = (t[:, k] for k in range(3, 9))
    du1, dv1 = u[:, 2] - u[:, 0], u[:, 3] - u[:, 1]
    du2, dv2 = u[:, 4] - u[:, 0], u[:, 5] - u[:, 1]
    nx, ny, nz = e1y * e2z - e1z * e2y, e1z * e2x - e1x * e2z, e1x * e2y - e1y * e2x
    d = du1 * dv2 - du2 * dv1
    big = np.abs(d) > 1e-20
    r = 1 / np.where(big, d, 1.0)
    tx = (e1x * dv2 - e2x * dv1) * r
    ty = (e1y * dv2 - e2y * dv1) * r
    tz = (e1z * dv2 - e2z * dv1) * r
    bx = (e2x * du1 - e1x * du2) * r
    by = (e2y * du1 - e1y * du2) * r
    bz = (e2z * du1 - e1z * du2) * r
    cx, cy, cz = ny * tz - nz * ty, nz * tx - nx * tz, nx * ty - ny * tx
    sign = np.where(big, np.where(cx * bx + cy * by + cz * bz < 0, -1.0, 1.0), 1.0)
    alt = np.abs(nx) < 0.9 * js_hypot(nx, ny, nz)
    fx = np.where(alt, 0.0, nz)
    fy = np.where(alt, -nz, 0.0)
    fz = np.where(alt, ny, -nx)
    tx, ty, tz = np.where(big, tx, fx), np.where(big, ty, fy), np.where(big, tz, fz)
    length = js_hypot(tx, ty, tz)
    length = np.where((length == 0) | np.isnan(length), 1.0, length)
    return np.stack([tx / length, ty / length, tz / length, sign], 1).astype(np.float32)


def compile_geometry(parts, build_bvh=None):
    """compile.ts:327-403 over the add_mesh results. build_bvh(triangles_f32, count) returns
    (bounds, nodes, order, node_count); the default is the native GPU builder."""
    if build_bvh is None:
        from .native import build_bvh
    parts = [p for p in parts if p is not None and len(p['p'])]
    if parts:
        P = np.concatenate([q['p'] for q in parts])
        N = np.concatenate([q['n'] for q in parts])
        UV = np.concatenate([q['uv'] for q in parts])
        any_colors = any(q['colors'] is not None for q in parts)
        COL = None
        if any_colors:
            COL = np.concatenate([q['colors'] if q['colors'] is not None else np.ones((len(q['p']), 3, 4))
                                  for q in parts])
        MAT = np.concatenate([np.full(len(q['p']), q['material'], np.uint32) for q in parts])
        FLG = np.concatenate([np.full(len(q['p']), q['flags'], np.uint8) for q in parts])
        SID = np.concatenate([np.full(len(q['p']), q['side'], np.uint8) for q in parts])
    else:
        P = np.zeros((0, 3, 3))
        N = np.zeros((0, 3, 3))
        UV = np.zeros((0, 3, 2))
        COL, any_colors = None, False
        MAT = np.zeros(0, np.uint32)
        FLG = np.zeros(0, np.uint8)
        SID = np.zeros(0, np.uint8)
    count = len(P)
    max_coord = float(np.abs(P).max()) if count else 0.0
    tri = np.empty((count, 9), np.float32)
    tri[:, 0:3] = P[:, 0]
    tri[:, 3:6] = P[:, 1] - P[:, 0]
    tri[:, 6:9] = P[:, 2] - P[:, 0]
    tri = np.ascontiguousarray(tri.reshape(-1))
    bounds, nodes, order, node_count = build_bvh(tri, count)
    tri9 = tri.reshape(count, 9)[order]
    normals = N.reshape(count, 9).astype(np.float32)[order]
    uvs = UV.reshape(count, 6).astype(np.float32)[order]
    colors = COL.reshape(count, 12).astype(np.float32)[order] if any_colors else np.zeros(0, np.float32)
    t64 = tri9.astype(np.float64)
    if count:
        vx = np.concatenate([t64[:, 0], t64[:, 0] + t64[:, 3], t64[:, 0] + t64[:, 6]])
        vy = np.concatenate([t64[:, 1], t64[:, 1] + t64[:, 4], t64[:, 1] + t64[:, 7]])
        vz = np.concatenate([t64[:, 2], t64[:, 2] + t64[:, 5], t64[:, 2] + t64[:, 8]])
        diagonal = js_hypot(vx.max() - vx.min(), vy.max() - vy.min(), vz.max() - vz.min())
    else:
        diagonal = 1.0
    tangents = _write_tangents(tri9, uvs) if count else np.zeros((0, 4), np.float32)
    epsilon = max(diagonal * 1e-5, max_coord * 2e-6, 1e-7)
    return CompiledGeometry(
        triangleCount=count, triangles=np.ascontiguousarray(tri9.reshape(-1)),
        normals=np.ascontiguousarray(normals.reshape(-1)), uvs=np.ascontiguousarray(uvs.reshape(-1)),
        tangents=np.ascontiguousarray(tangents.reshape(-1)), colors=np.ascontiguousarray(colors.reshape(-1)),
        materials=np.ascontiguousarray(MAT[order]), flags=np.ascontiguousarray(FLG[order]),
        sides=np.ascontiguousarray(SID[order]),
        bvh=CompiledBvh(np.ascontiguousarray(bounds), np.ascontiguousarray(nodes), int(node_count)),
        epsilon=float(epsilon))


def compile_shading(scene):  # compile.ts:476-510
    scene.updateMatrixWorld()
    meshes, lights = collect_objects(scene)
    materials, index = [], {}
    table = TextureTable()
    for mesh in meshes:
        m = mesh.material
        if m is None or id(m) in index:
            continue
        index[id(m)] = len(materials)
        materials.append(compile_material(m, table))
    background = {'kind': 'clear'}  # compile.ts:434-449
    bg = scene.background
    if bg is not None and getattr(bg, 'isTexture', False):
        intensity = getattr(scene, 'backgroundIntensity', 1.0)
        if bg.isCubeTexture or bg.mapping in (301, 302, 303, 304):
            e = table.add_environment(bg, intensity, getattr(scene, 'backgroundRotation', None))
            if e >= 0:
                background = {'kind': 'environment', 'environment': e, 'intensity': intensity}
        else:
            t = table.add(bg)
            if t >= 0:
                background = {'kind': 'screen', 'texture': t, 'intensity': intensity}
    elif bg is not None:
        background = {'kind': 'color', 'color': _rgb(bg)}
    environment = -1
    if scene.environment is not None:
        environment = table.add_environment(scene.environment, getattr(scene, 'environmentIntensity', 1.0),
                                            getattr(scene, 'environmentRotation', None))
    fog = None
    if scene.fog is not None:
        f = scene.fog
        if f.isFogExp2:
            fog = CompiledFog('exp2', _rgb(f.color), 0.0, 0.0, f.density)
        else:
            fog = CompiledFog('linear', _rgb(f.color), f.near, f.far, 0.0)
    compiled_lights = [l for l in (compile_light(x) for x in lights) if l is not None]
    shading = ShadingPayload(materials=materials, lights=compiled_lights, background=background, fog=fog,
                             environment=environment, textures=table.textures, environments=table.environments)
    return shading, meshes, index


def compile_scene(scene, build_bvh=None):  # compile.ts:513-527
    shading, meshes, index = compile_shading(scene)
    parts = [add_mesh(m, index.get(id(m.material), 0), m.material) for m in meshes]
    return CompiledScene(compile_geometry(parts, build_bvh), shading)


def mesh_descriptor(mesh, mat_index, mat):
    """The plain arrays addMesh reads (compile.ts:212-251) for mcr_compile_geometry, or None when
    the mesh contributes nothing (as add_mesh)."""
    geo = mesh.geometry
    if geo is None or geo.position is None or not mat.visible or mat.wireframe:
        return None
    opaque = mat.transparent is not True and not (mat.alphaTest > 0) and not (getattr(mat, 'transmission', 0) > 0)
    flags = ((abi.TRI_CAST_SHADOW if mesh.castShadow else 0) |
             (abi.TRI_RECEIVE_SHADOW if mesh.receiveShadow else 0) | (abi.TRI_OPAQUE if opaque else 0))
    colors = bool(mat.vertexColors) and geo.color is not None
    return {'world': list(mesh.matrixWorld), 'positions': geo.position, 'normals': geo.normal, 'uvs': geo.uv,
            'colors': geo.color, 'index': geo.index, 'draw': None,
            'groups': [(0, None, mat_index, flags, mat.side, bool(mat.flatShading), colors)]}


def compile_scene_device(scene, ctx=None):
    """compile_scene with compileGeometry on the device (mcr_compile_geometry): flattening, the BVH
    build, the BVH-order gather and the tangents in one device pass; the same arrays as
    compile_scene with the GPU builder."""
    from . import native
    shading, meshes, index = compile_shading(scene)
    descs = [mesh_descriptor(m, index.get(id(m.material), 0), m.material) for m in meshes]
    out = native.compile_meshes(ctx, [d for d in descs if d is not None])
    geometry = CompiledGeometry(
        triangleCount=out['count'], triangles=out['triangles'], normals=out['normals'], uvs=out['uvs'],
        tangents=out['tangents'], colors=out['colors'], materials=out['materials'], flags=out['flags'],
        sides=out['sides'], bvh=CompiledBvh(out['bounds'], out['nodes'], out['node_count']),
        epsilon=float(out['epsilon']))
    return CompiledScene(geometry, shading)


def compile_scene_resident(scene, gpu_scene, version):
    """compileGeometry on the device straight into a resident scene (mcr_scene_compile_geometry),
    then the shading; returns the ShadingPayload and the compile info. Renders match a scene given
    compile_scene_device's arrays."""
    shading, meshes, index = compile_shading(scene)
    descs = [mesh_descriptor(m, index.get(id(m.material), 0), m.material) for m in meshes]
    info = gpu_scene.compile_meshes([d for d in descs if d is not None], version)
    gpu_scene.set_shading(shading, version)
    return shading, info


def compile_camera(camera, aspect=None):  # compile.ts:536-557
    if camera.parent is None:
        camera.updateMatrixWorld()
    world = list(camera.matrixWorld)
    proj = camera.projection()
    inv = [float(v) for v in _invert_three(proj)]
    if aspect is not None and aspect > 0:
        camera_aspect = proj[5] / proj[0] if proj[0] != 0 else 0
        if camera_aspect > 0:
            k = aspect / camera_aspect
            for i in range(4):
                inv[i] = inv[i] * k
    return CompiledCamera(projectionInverse=inv, world=world,
                          orthographic=camera.isOrthographicCamera is True,
                          forward=_normalize3((-world[8], -world[9], -world[10])),
                          position=(world[12], world[13], world[14]))


def _invert_three(te):
    """Matrix4.invert: cofactor expansion in three.js's order."""
    n11, n21, n31, n41 = te[0], te[1], te[2], te[3]
    n12, n22, n32, n42 = te[4], te[5], te[6], te[7]
    n13, n23, n33, n43 = te[8], te[9], te[10], te[11]
    n14, n24, n34, n44 = te[12], te[13], te[14], te[15]
    t11 = n23 * n34 * n42 - n24 * n33 * n42 + n24 * n32 * n43 - n22 * n34 * n43 - n23 * n32 * n44 + n22 * n33 * n44
    t12 = n14 * n33 * n42 - n13 * n34 * n42 - n14 * n32 * n43 + n12 * n34 * n43 + n13 * n32 * n44 - n12 * n33 * n44
    t13 = n13 * n24 * n42 - n14 * n23 * n42 + n14 * n22 * n43 - n12 * n24 * n43 - n13 * n22 * n44 + n12 * n23 * n44
    t14 = n14 * n23 * n32 - n13 * n24 * n32 - n14 * n22 * n33 + n12 * n24 * n33 + n13 * n22 * n34 - n12 * n23 * n34
    det = n11 * t11 + n21 * t12 + n31 * t13 + n41 * t14
    if det == 0:
        return [0.0] * 16
    d = 1 / det
    out = [0.0] * 16
    out[0] = t11 * d
    out[1] = (n24 * n33 * n41 - n23 * n34 * n41 - n24 * n31 * n43 + n21 * n34 * n43 + n23 * n31 * n44 - n21 * n33 * n44) * d
    out[2] = (n22 * n34 * n41 - n24 * n32 * n41 + n24 * n31 * n42 - n21 * n34 * n42 - n22 * n31 * n44 + n21 * n32 * n44) * d
    out[3] = (n23 * n32 * n41 - n22 * n33 * n41 - n23 * n31 * n42 + n21 * n33 * n42 + n22 * n31 * n43 - n21 * n32 * n43) * d
    out[4] = t12 * d
    out[5] = (n13 * n34 * n41 - n14 * n33 * n41 + n14 * n31 * n43 - n11 * n34 * n43 - n13 * n31 * n44 + n11 * n33 * n44) * d
    out[6] = (n14 * n32 * n41 - n12 * n34 * n41 - n14 * n31 * n42 + n11 * n34 * n42 + n12 * n31 * n44 - n11 * n32 * n44) * d
    out[7] = (n12 * n33 * n41 - n13 * n32 * n41 + n13 * n31 * n42 - n11 * n33 * n42 - n12 * n31 * n43 + n11 * n32 * n43) * d
    out[8] = t13 * d
    out[9] = (n14 * n23 * n41 - n13 * n24 * n41 - n14 * n21 * n43 + n11 * n24 * n43 + n13 * n21 * n44 - n11 * n23 * n44) * d
    out[10] = (n12 * n24 * n41 - n14 * n22 * n41 + n14 * n21 * n42 - n11 * n24 * n42 - n12 * n21 * n44 + n11 * n22 * n44) * d
    out[11] = (n13 * n22 * n41 - n12 * n23 * n41 - n13 * n21 * n42 + n11 * n23 * n42 + n12 * n21 * n43 - n11 * n22 * n43) * d
    out[12] = t14 * d
    out[13] = (n13 * n24 * n31 - n14 * n23 * n31 + n14 * n21 * n33 - n11 * n24 * n33 - n13 * n21 * n34 + n11 * n23 * n34) * d
    out[14] = (n14 * n22 * n31 - n12 * n24 * n31 - n14 * n21 * n32 + n11 * n24 * n32 + n12 * n21 * n34 - n11 * n22 * n34) * d
    out[15] = (n12 * n23 * n31 - n13 * n22 * n31 + n13 * n21 * n32 - n11 * n23 * n32 - n12 * n21 * n33 + n11 * n22 * n33) * d
    return out
