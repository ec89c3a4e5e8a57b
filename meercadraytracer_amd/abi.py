"""ctypes mirror of include/mcr.h: the plain-data structs of the C-ABI.

The structs restate the reference's plain-data scene format (src/types.ts:43-211) in C layout.
Field order and types follow include/mcr.h exactly; ctypes applies the same alignment rules as
the C compiler, and tests/test_abi.py checks every size against the native library.
"""
import ctypes as C

MCR_OK = 0
MCR_E_INVALID = -1
MCR_E_UNSUPPORTED = -2
MCR_E_DEVICE = -3
MCR_E_NO_SCENE = -4
MCR_E_OVERFLOW = -5
MCR_E_CANCELLED = -6

# constants.ts:5-20, types.ts:41,91-135,195
FrontSide, BackSide, DoubleSide = 0, 1, 2
MultiplyOperation, MixOperation, AddOperation = 0, 1, 2
(NoToneMapping, LinearToneMapping, ReinhardToneMapping, CineonToneMapping,
 ACESFilmicToneMapping, CustomToneMapping, AgXToneMapping, NeutralToneMapping) = range(8)
TRI_OPAQUE, TRI_CAST_SHADOW, TRI_RECEIVE_SHADOW = 1, 2, 4
MODELS = {'physical': 0, 'phong': 1, 'lambert': 2, 'toon': 3, 'basic': 4}
LIGHTS = {'ambient': 0, 'hemisphere': 1, 'directional': 2, 'point': 3, 'spot': 4}
SHADOWS = {'all': 0, 'castShadow': 1, 'none': 2}
FOGS = {None: 0, 'linear': 1, 'exp2': 2}
RepeatWrapping, ClampToEdgeWrapping, MirroredRepeatWrapping = 1000, 1001, 1002  # constants.ts:29-31
MultiplyOperation, MixOperation, AddOperation = 0, 1, 2                           # constants.ts:9-11
ENV_KINDS = {'equirect': 0, 'cube': 1}
BACKGROUNDS = {'clear': 0, 'color': 1, 'screen': 2, 'environment': 3}

d3 = C.c_double * 3


class Geometry(C.Structure):
    _fields_ = [
        ('triangle_count', C.c_uint32), ('node_count', C.c_uint32),
        ('triangles', C.c_void_p), ('normals', C.c_void_p), ('uvs', C.c_void_p),
        ('tangents', C.c_void_p), ('colors', C.c_void_p), ('materials', C.c_void_p),
        ('flags', C.c_void_p), ('sides', C.c_void_p), ('bvh_bounds', C.c_void_p),
        ('bvh_nodes', C.c_void_p), ('epsilon', C.c_double),
    ]


class Material(C.Structure):
    _fields_ = [
        ('model', C.c_int32), ('side', C.c_int32),
        ('color', d3), ('opacity', C.c_double),
        ('transparent', C.c_int32), ('vertex_colors', C.c_int32),
        ('alpha_test', C.c_double),
        ('flat_shading', C.c_int32), ('tone_mapped', C.c_int32),
        ('map', C.c_int32), ('alpha_map', C.c_int32), ('emissive_map', C.c_int32),
        ('normal_map', C.c_int32), ('ao_map', C.c_int32), ('roughness_map', C.c_int32),
        ('metalness_map', C.c_int32), ('env_map', C.c_int32),
        ('emissive', d3), ('normal_scale', C.c_double * 2), ('ao_map_intensity', C.c_double),
        ('roughness', C.c_double), ('metalness', C.c_double),
        ('physical', C.c_int32), ('combine', C.c_int32),
        ('ior', C.c_double), ('specular_intensity', C.c_double), ('specular_color', d3),
        ('transmission', C.c_double), ('thickness', C.c_double), ('attenuation', d3),
        ('shininess', C.c_double), ('specular', d3),
        ('env_map_intensity', C.c_double), ('reflectivity', C.c_double),
    ]


class Light(C.Structure):
    _fields_ = [
        ('type', C.c_int32), ('cast_shadow', C.c_int32),
        ('color', d3), ('ground_color', d3), ('direction', d3), ('position', d3),
        ('distance', C.c_double), ('decay', C.c_double), ('cone_cos', C.c_double),
        ('penumbra_cos', C.c_double),
    ]


class Texture(C.Structure):
    _fields_ = [('width', C.c_uint32), ('height', C.c_uint32), ('data', C.c_void_p), ('is_float', C.c_int32),
                ('srgb', C.c_int32), ('flip_y', C.c_int32), ('wrap_s', C.c_int32), ('wrap_t', C.c_int32),
                ('pad0', C.c_int32), ('matrix', C.c_double * 9)]


class Environment(C.Structure):
    _fields_ = [('kind', C.c_int32), ('texture', C.c_int32), ('faces', C.c_int32 * 6), ('flip_x', C.c_int32),
                ('pad0', C.c_int32), ('rotation', C.c_double * 9), ('intensity', C.c_double)]


class Shading(C.Structure):
    _fields_ = [
        ('material_count', C.c_uint32), ('light_count', C.c_uint32),
        ('materials', C.c_void_p), ('lights', C.c_void_p),
        ('texture_count', C.c_uint32), ('environment_count', C.c_uint32),
        ('background_kind', C.c_int32), ('environment', C.c_int32),
        ('background_color', d3),
        ('fog_kind', C.c_int32), ('pad0', C.c_int32),
        ('fog_color', d3), ('fog_near', C.c_double), ('fog_far', C.c_double),
        ('fog_density', C.c_double),
        ('textures', C.c_void_p), ('environments', C.c_void_p),
        ('background_texture', C.c_int32), ('background_environment', C.c_int32),
        ('background_intensity', C.c_double),
    ]


class Camera(C.Structure):
    _fields_ = [
        ('projection_inverse', C.c_double * 16), ('world', C.c_double * 16),
        ('forward', d3), ('position', d3), ('orthographic', C.c_int32), ('pad0', C.c_int32),
    ]


class Settings(C.Structure):
    _fields_ = [
        ('max_bounces', C.c_int32), ('shadows', C.c_int32),
        ('global_illumination', C.c_int32), ('tone_mapping', C.c_int32),
        ('tone_mapping_exposure', C.c_double), ('clear_color', d3), ('clear_alpha', C.c_double),
        ('seed', C.c_uint32), ('pad0', C.c_uint32),
    ]


class Tile(C.Structure):
    _fields_ = [('x', C.c_int32), ('y', C.c_int32), ('width', C.c_int32), ('height', C.c_int32)]


class TileJob(C.Structure):
    _fields_ = [
        ('tile', Tile), ('width', C.c_uint32), ('height', C.c_uint32), ('pass_', C.c_uint32),
        ('block', C.c_uint32), ('out_offset', C.c_uint64),
    ]


class View(C.Structure):
    _fields_ = [('scene', C.c_void_p), ('camera', C.c_void_p), ('settings', C.c_void_p)]


class Counters(C.Structure):
    _fields_ = [
        ('closest_queries', C.c_uint64), ('any_queries', C.c_uint64),
        ('node_visits', C.c_uint64), ('leaf_visits', C.c_uint64),
        ('triangle_tests', C.c_uint64), ('paths', C.c_uint64),
        ('closest_node_visits', C.c_uint64), ('closest_triangle_tests', C.c_uint64),
        ('any_node_visits', C.c_uint64), ('any_triangle_tests', C.c_uint64),
        ('shade_waves', C.c_uint64), ('mixed_view_waves', C.c_uint64),
        ('shade_class_lanes', C.c_uint64 * 8), ('class_mixed_waves', C.c_uint64),
    ]

    def as_dict(self):
        return {name: (list(map(int, getattr(self, name))) if name == 'shade_class_lanes' else int(getattr(self, name)))
                for name, _ in self._fields_}


class Timing(C.Structure):
    _fields_ = [
        ('closest_ms', C.c_double), ('any_ms', C.c_double), ('shade_ms', C.c_double),
        ('other_ms', C.c_double), ('total_ms', C.c_double),
        ('closest_launches', C.c_uint64), ('any_launches', C.c_uint64),
    ]


class MeshGroup(C.Structure):
    _fields_ = [('start', C.c_uint32), ('count', C.c_uint32), ('material', C.c_uint32), ('flags', C.c_uint8),
                ('side', C.c_uint8), ('flat_shading', C.c_uint8), ('vertex_colors', C.c_uint8)]


class Mesh(C.Structure):
    _fields_ = [('world', C.c_double * 16), ('positions', C.c_void_p), ('normals', C.c_void_p),
                ('uvs', C.c_void_p), ('colors', C.c_void_p), ('index', C.c_void_p), ('groups', C.c_void_p),
                ('vertex_count', C.c_uint32), ('index_count', C.c_uint32), ('color_size', C.c_uint32),
                ('group_count', C.c_uint32), ('draw_start', C.c_uint32), ('draw_count', C.c_uint32)]


class GeometryOut(C.Structure):
    _fields_ = [('triangle_count', C.c_uint32), ('node_count', C.c_uint32), ('triangles', C.c_void_p),
                ('normals', C.c_void_p), ('uvs', C.c_void_p), ('tangents', C.c_void_p), ('colors', C.c_void_p),
                ('materials', C.c_void_p), ('flags', C.c_void_p), ('sides', C.c_void_p),
                ('bvh_bounds', C.c_void_p), ('bvh_nodes', C.c_void_p), ('epsilon', C.c_double)]


INF_COUNT = 0xffffffff


class CompileInfo(C.Structure):  # mcr_compile_info
    _fields_ = [('triangle_count', C.c_uint32), ('node_count', C.c_uint32), ('host_layout', C.c_uint32),
                ('epsilon', C.c_double)]


class WideInfo(C.Structure):
    _fields_ = [('node_count', C.c_uint32), ('triangle_count', C.c_uint32), ('depth', C.c_uint32),
                ('pad0', C.c_uint32), ('bytes', C.c_uint64), ('build_ms', C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != 'pad0'}


class ServerConfig(C.Structure):
    _fields_ = [('hbm_budget', C.c_uint64), ('slots', C.c_uint32), ('window', C.c_uint32),
                ('max_requests', C.c_uint32), ('use_graph', C.c_int32), ('max_lights', C.c_uint32),
                ('lanes', C.c_uint32), ('preview_slots', C.c_uint32), ('preview_cus', C.c_uint32),
                ('preview_passes', C.c_int32)]


class Request(C.Structure):
    _fields_ = [('scene', C.c_void_p), ('camera', Camera), ('settings', Settings), ('width', C.c_uint32),
                ('height', C.c_uint32), ('samples', C.c_uint32), ('tile_size', C.c_uint32),
                ('preview_block', C.c_uint32), ('srgb', C.c_int32), ('geometry_version', C.c_uint32),
                ('shading_version', C.c_uint32)]


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_CANCELLED, REQ_FAILED = range(5)
MCR_E_CANCELLED, MCR_E_TIMEOUT, MCR_E_INTERNAL = -6, -7, -8
MCR_ISECT_FP64, MCR_ISECT_WIDE_FP32 = 0, 1
MCR_NOALLOC_SCRATCH_BYTES, MCR_NOALLOC_MAX_RAYS = 58737664, 0x80000000  # mcr_intersect_closest_noalloc


class RequestStatus(C.Structure):
    _fields_ = [('state', C.c_int32), ('samples', C.c_uint32), ('max_samples', C.c_uint32),
                ('preview_done', C.c_int32), ('submit_ms', C.c_double), ('admit_ms', C.c_double),
                ('first_preview_ms', C.c_double), ('done_ms', C.c_double), ('paths', C.c_uint64),
                ('preview_iterations', C.c_uint32), ('pad0', C.c_uint32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServerStats(C.Structure):
    _fields_ = [('closest_queries', C.c_uint64), ('any_queries', C.c_uint64), ('paths', C.c_uint64),
                ('iterations', C.c_uint64), ('shade_waves', C.c_uint64), ('mixed_view_waves', C.c_uint64),
                ('bytes_in_use', C.c_uint64), ('requests_live', C.c_uint64),
                ('requests_queued', C.c_uint64), ('device_ms', C.c_double), ('uptime_ms', C.c_double),
                ('host_ms', C.c_double), ('plan_ms', C.c_double), ('deliver_ms', C.c_double),
                ('pass_busy_ms', C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def to_int32(x):
    """ECMAScript ToInt32, as random.ts:7 applies it to settings.seed through `^`."""
    import math
    if x != x or x in (float('inf'), float('-inf')):
        return 0
    v = int(math.trunc(x)) % (1 << 32)
    return v
