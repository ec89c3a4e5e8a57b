"""Plain-data scene format of the reference (src/types.ts:43-211), as Python objects.

CompiledGeometry holds numpy arrays in BVH leaf order; materials, lights, camera and settings are
small objects. `to_c` methods produce the include/mcr.h structs (meercadraytracer_amd.abi) that
the native library and the oracle consume; the returned holder keeps every array alive.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import abi


def _ptr(a):
    return None if a is None or a.size == 0 else a.ctypes.data


@dataclass
class CompiledBvh:  # types.ts:158-167
    bounds: np.ndarray   # float32 [nodeCount*6]
    nodes: np.ndarray    # int32 [nodeCount*3]
    nodeCount: int


@dataclass
class CompiledGeometry:  # types.ts:137-156
    triangleCount: int
    triangles: np.ndarray
    normals: np.ndarray
    uvs: np.ndarray
    tangents: np.ndarray
    colors: np.ndarray
    materials: np.ndarray
    flags: np.ndarray
    sides: np.ndarray
    bvh: CompiledBvh
    epsilon: float

    def to_c(self):
        arrays = [self.triangles, self.normals, self.uvs, self.tangents, self.colors,
                  self.materials, self.flags, self.sides, self.bvh.bounds, self.bvh.nodes]
        for a in arrays:
            assert a.flags['C_CONTIGUOUS']
        g = abi.Geometry()
        g.triangle_count = self.triangleCount
        g.node_count = self.bvh.nodeCount
        g.triangles = _ptr(self.triangles)
        g.normals = _ptr(self.normals)
        g.uvs = _ptr(self.uvs)
        g.tangents = _ptr(self.tangents)
        g.colors = _ptr(self.colors)
        g.materials = _ptr(self.materials)
        g.flags = _ptr(self.flags)
        g.sides = _ptr(self.sides)
        g.bvh_bounds = _ptr(self.bvh.bounds)
        g.bvh_nodes = _ptr(self.bvh.nodes)
        g.epsilon = float(self.epsilon)
        g._keep = arrays
        return g


@dataclass
class CompiledTexture:  # types.ts:8-21
    width: int
    height: int
    data: np.ndarray          # RGBA texels, first row first: uint8 or float32
    srgb: bool = False
    flipY: bool = False
    wrapS: int = abi.ClampToEdgeWrapping
    wrapT: int = abi.ClampToEdgeWrapping
    matrix: tuple = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)

    def fill(self, t):
        t.width, t.height = self.width, self.height
        t.data = self.data.ctypes.data
        t.is_float = int(self.data.dtype == np.float32)
        t.srgb, t.flip_y = int(self.srgb), int(self.flipY)
        t.wrap_s, t.wrap_t = self.wrapS, self.wrapT
        t.matrix[:] = [float(x) for x in self.matrix]


@dataclass
class TextureRef:  # protocol.ts:15-16: a CompiledTexture without its pixels, which travel once by id
    id: int
    width: int
    height: int
    srgb: bool = False
    flipY: bool = False
    wrapS: int = abi.ClampToEdgeWrapping
    wrapT: int = abi.ClampToEdgeWrapping
    matrix: tuple = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)

    @classmethod
    def of(cls, texture, id):
        return cls(id, texture.width, texture.height, texture.srgb, texture.flipY, texture.wrapS, texture.wrapT,
                   tuple(texture.matrix))

    def with_data(self, data):  # protocol.ts:82-87
        return CompiledTexture(self.width, self.height, data, self.srgb, self.flipY, self.wrapS, self.wrapT,
                               self.matrix)


@dataclass
class CompiledEnvironment:  # types.ts:23-39
    kind: str                 # 'equirect' | 'cube'
    texture: int = -1
    faces: tuple = ()
    flipX: bool = True
    rotation: tuple = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)
    intensity: float = 1.0

    def fill(self, e):
        e.kind = abi.ENV_KINDS[self.kind]
        e.texture = self.texture
        e.faces[:] = list(self.faces) if self.faces else [-1] * 6
        e.flip_x = int(self.flipX)
        e.rotation[:] = [float(x) for x in self.rotation]
        e.intensity = self.intensity


@dataclass
class CompiledMaterial:  # types.ts:43-89, defaults as compile.ts:94-147 fills them
    model: str = 'physical'
    side: int = abi.FrontSide
    color: tuple = (1.0, 1.0, 1.0)
    opacity: float = 1.0
    transparent: bool = False
    alphaTest: float = 0.0
    vertexColors: bool = False
    flatShading: bool = False
    toneMapped: bool = True
    map: int = -1
    alphaMap: int = -1
    emissiveMap: int = -1
    normalMap: int = -1
    aoMap: int = -1
    roughnessMap: int = -1
    metalnessMap: int = -1
    emissive: tuple = (0.0, 0.0, 0.0)
    normalScale: tuple = (1.0, 1.0)
    aoMapIntensity: float = 1.0
    roughness: float = 1.0
    metalness: float = 0.0
    physical: bool = False
    ior: float = 1.5
    specularIntensity: float = 1.0
    specularColor: tuple = (1.0, 1.0, 1.0)
    transmission: float = 0.0
    thickness: float = 0.0
    attenuation: tuple = (0.0, 0.0, 0.0)
    shininess: float = 30.0
    specular: tuple = (0.0666, 0.0666, 0.0666)
    envMap: int = -1
    envMapIntensity: float = 1.0
    reflectivity: float = 1.0
    combine: int = 0

    def fill(self, m):
        m.model = abi.MODELS[self.model]
        m.side = self.side
        m.color[:] = self.color
        m.opacity = self.opacity
        m.transparent = int(self.transparent)
        m.vertex_colors = int(self.vertexColors)
        m.alpha_test = self.alphaTest
        m.flat_shading = int(self.flatShading)
        m.tone_mapped = int(self.toneMapped)
        m.map, m.alpha_map, m.emissive_map = self.map, self.alphaMap, self.emissiveMap
        m.normal_map, m.ao_map, m.roughness_map = self.normalMap, self.aoMap, self.roughnessMap
        m.metalness_map, m.env_map = self.metalnessMap, self.envMap
        m.emissive[:] = self.emissive
        m.normal_scale[:] = self.normalScale
        m.ao_map_intensity = self.aoMapIntensity
        m.roughness, m.metalness = self.roughness, self.metalness
        m.physical, m.combine = int(self.physical), self.combine
        m.ior, m.specular_intensity = self.ior, self.specularIntensity
        m.specular_color[:] = self.specularColor
        m.transmission, m.thickness = self.transmission, self.thickness
        m.attenuation[:] = self.attenuation
        m.shininess = self.shininess
        m.specular[:] = self.specular
        m.env_map_intensity, m.reflectivity = self.envMapIntensity, self.reflectivity


@dataclass
class CompiledLight:  # types.ts:91-113
    type: str
    color: tuple = (1.0, 1.0, 1.0)
    groundColor: tuple = (0.0, 0.0, 0.0)
    direction: tuple = (0.0, 0.0, 1.0)
    position: tuple = (0.0, 0.0, 0.0)
    distance: float = 0.0
    decay: float = 2.0
    coneCos: float = 0.0
    penumbraCos: float = 0.0
    castShadow: bool = False

    def fill(self, l):
        l.type = abi.LIGHTS[self.type]
        l.cast_shadow = int(self.castShadow)
        l.color[:] = self.color
        l.ground_color[:] = self.groundColor
        l.direction[:] = self.direction
        l.position[:] = self.position
        l.distance, l.decay = self.distance, self.decay
        l.cone_cos, l.penumbra_cos = self.coneCos, self.penumbraCos


@dataclass
class CompiledFog:  # types.ts:115-121
    kind: str
    color: tuple
    near: float = 0.0
    far: float = 0.0
    density: float = 0.0


@dataclass
class ShadingPayload:  # protocol.ts:19-27, with the textures message's arrays (protocol.ts:31-35)
    materials: List[CompiledMaterial]
    lights: List[CompiledLight]
    background: dict = field(default_factory=lambda: {'kind': 'clear'})
    environment: int = -1
    fog: Optional[CompiledFog] = None
    textures: list = field(default_factory=list)
    environments: list = field(default_factory=list)

    def to_c(self):
        mats = (abi.Material * max(1, len(self.materials)))()
        for i, m in enumerate(self.materials):
            m.fill(mats[i])
        lights = (abi.Light * max(1, len(self.lights)))()
        for i, l in enumerate(self.lights):
            l.fill(lights[i])
        s = abi.Shading()
        s.material_count = len(self.materials)
        s.light_count = len(self.lights)
        s.materials = C.cast(mats, C.c_void_p)
        s.lights = C.cast(lights, C.c_void_p)
        s.texture_count = len(self.textures)
        s.environment_count = len(self.environments)
        texs = (abi.Texture * max(1, len(self.textures)))()
        for i, t in enumerate(self.textures):
            t.fill(texs[i])
        envs = (abi.Environment * max(1, len(self.environments)))()
        for i, e in enumerate(self.environments):
            e.fill(envs[i])
        s.textures = C.cast(texs, C.c_void_p)
        s.environments = C.cast(envs, C.c_void_p)
        s.background_kind = abi.BACKGROUNDS[self.background['kind']]
        s.background_texture, s.background_environment, s.background_intensity = -1, -1, 1.0
        if self.background['kind'] == 'color':
            s.background_color[:] = self.background['color']
        elif self.background['kind'] == 'screen':
            s.background_texture = self.background['texture']
            s.background_intensity = self.background.get('intensity', 1.0)
        elif self.background['kind'] == 'environment':
            s.background_environment = self.background['environment']
            s.background_intensity = self.background.get('intensity', 1.0)
        s.environment = self.environment
        if self.fog is not None:
            s.fog_kind = abi.FOGS[self.fog.kind]
            s.fog_color[:] = self.fog.color
            s.fog_near, s.fog_far, s.fog_density = self.fog.near, self.fog.far, self.fog.density
        s._keep = (mats, lights, texs, envs, [t.data for t in self.textures])
        return s


@dataclass
class CompiledScene:  # types.ts:169-179
    geometry: CompiledGeometry
    shading: ShadingPayload


@dataclass
class CompiledCamera:  # types.ts:181-190
    projectionInverse: list
    world: list
    orthographic: bool
    forward: tuple
    position: tuple

    def to_c(self):
        c = abi.Camera()
        c.projection_inverse[:] = [float(v) for v in self.projectionInverse]
        c.world[:] = [float(v) for v in self.world]
        c.forward[:] = self.forward
        c.position[:] = self.position
        c.orthographic = int(self.orthographic)
        return c


@dataclass
class TraceSettings:  # types.ts:193-204
    maxBounces: int = 4
    shadows: str = 'all'
    globalIllumination: bool = False
    toneMapping: int = abi.NoToneMapping
    toneMappingExposure: float = 1.0
    clearColor: tuple = (0.0, 0.0, 0.0)
    clearAlpha: float = 0.0
    seed: float = 0

    def to_c(self):
        s = abi.Settings()
        s.max_bounces = int(self.maxBounces)
        s.shadows = abi.SHADOWS[self.shadows]
        s.global_illumination = int(bool(self.globalIllumination))
        s.tone_mapping = int(self.toneMapping)
        s.tone_mapping_exposure = float(self.toneMappingExposure)
        s.clear_color[:] = self.clearColor
        s.clear_alpha = float(self.clearAlpha)
        s.seed = abi.to_int32(self.seed)
        return s

    def key(self):
        """The settings' JSON key, as protocol.ts:80 compares them."""
        import json
        return json.dumps([self.maxBounces, self.shadows, self.globalIllumination, self.toneMapping,
                           self.toneMappingExposure, list(self.clearColor), self.clearAlpha,
                           self.seed])


@dataclass
class Tile:  # types.ts:206-211
    x: int
    y: int
    width: int
    height: int

    def to_c(self):
        return abi.Tile(self.x, self.y, self.width, self.height)
