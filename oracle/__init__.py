"""ctypes binding of the CPU oracle (oracle.cpp). TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this package, as
the checker or the timed CPU baseline; the product (meercadraytracer_amd) never does. `load()`
returns the portable-math build (bit-identical to the device's arithmetic); `load(libm=True)`
returns the glibc-libm build.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from meercadraytracer_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
_libs = {}


def build():
    subprocess.run(['make', '-s', '-C', HERE], check=True)


def _lib_path(libm):
    return os.path.join(HERE, 'liboracle_libm.so' if libm else 'liboracle.so')


def load(libm=False):
    if libm in _libs:
        return _libs[libm]
    path = _lib_path(libm)
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    P = C.c_void_p
    sig = {
        'orc_build_bvh': (C.c_int, [P, C.c_uint32, P, P, P, P]),
        'orc_intersect_closest': (None, [P, P, C.c_uint32, C.c_double, C.c_double, C.c_int, P, P, P, P, P, P]),
        'orc_intersect_any_count': (None, [P, P, C.c_uint32, C.c_double, C.c_double, P, P]),
        'orc_brute_force': (None, [P, C.c_uint32, P, C.c_uint32, P, P, P]),
        'orc_wide_brute': (None, [P, P, C.c_uint32, P, C.c_uint32, C.c_float, C.c_float, C.c_int, P, P, P, P, P,
                                  C.c_uint32]),
        'orc_sample_texture': (None, [P, C.c_double, C.c_double, P]),
        'orc_sample_environment': (None, [P, P, C.c_double, C.c_double, C.c_double, P]),
        'orc_atan2': (C.c_double, [C.c_double, C.c_double]),
        'orc_asin': (C.c_double, [C.c_double]),
        'orc_hash': (C.c_uint32, [C.c_uint32]),
        'orc_seed': (C.c_uint32, [C.c_int32] * 4),
        'orc_random_stream': (None, [C.c_uint32, C.c_uint32, P]),
        'orc_dfg': (None, [C.c_double, C.c_double, P]),
        'orc_fresnel_weight': (C.c_double, [C.c_double]),
        'orc_visibility_ggx': (C.c_double, [C.c_double] * 3),
        'orc_distribution_ggx': (C.c_double, [C.c_double] * 2),
        'orc_blinn_phong': (C.c_double, [C.c_double] * 2),
        'orc_smoothstep': (C.c_double, [C.c_double] * 3),
        'orc_distance_attenuation': (C.c_double, [C.c_double] * 3),
        'orc_half_to_float': (C.c_double, [C.c_uint32]),
        'orc_linear_to_srgb': (C.c_double, [C.c_double]),
        'orc_srgb_to_linear': (C.c_double, [C.c_double]),
        'orc_tone_map': (None, [P, C.c_int, C.c_double]),
        'orc_math': (C.c_double, [C.c_int, C.c_double, C.c_double]),
        'orc_camera_ray': (None, [P, C.c_double, C.c_double, P]),
        'orc_render_tile': (C.c_int, [P, P, P, P, abi.Tile, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, P, P]),
        'orc_encode_tile': (None, [P, C.c_uint32, abi.Tile, P, P, C.c_uint32, C.c_int]),
        'orc_make_tiles': (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32, P]),
        'orc_raytrace': (C.c_int, [P, P, P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, P, P]),
        'orc_render_items_mt': (C.c_double, [P, P, P, P, P, C.c_uint32, C.c_uint32, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _libs[libm] = lib
    return lib


def _p(a):
    return a.ctypes.data if a is not None else None


def build_bvh(triangles, count, libm=False):
    """bvh.ts:16 buildBvh -> (bounds f32, nodes i32, order u32, nodeCount)."""
    lib = load(libm)
    tri = np.ascontiguousarray(triangles, dtype=np.float32)
    m = max(1, 2 * count - 1)
    bounds = np.zeros(m * 6, np.float32)
    nodes = np.zeros(m * 3, np.int32)
    order = np.zeros(max(count, 1), np.uint32)
    nc = C.c_uint32(0)
    lib.orc_build_bvh(_p(tri), count, _p(bounds), _p(nodes), _p(order), C.byref(nc))
    n = nc.value
    return bounds[:n * 6].copy(), nodes[:n * 3].copy(), order[:count].copy(), n


def intersect_closest(geometry, rays, t_min=0.0, t_max=np.inf, cull=True, libm=False):
    lib = load(libm)
    g = geometry.to_c()
    rays = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 6)
    n = len(rays)
    tri = np.empty(n, np.int32)
    t = np.empty(n)
    u = np.empty(n)
    v = np.empty(n)
    front = np.empty(n, np.uint8)
    cnt = abi.Counters()
    lib.orc_intersect_closest(C.byref(g), _p(rays), n, t_min, t_max, int(cull), _p(tri), _p(t), _p(u), _p(v),
                              _p(front), C.byref(cnt))
    return tri, t, u, v, front, cnt.as_dict()


def intersect_any_count(geometry, rays, t_min=0.0, t_max=np.inf, libm=False):
    lib = load(libm)
    g = geometry.to_c()
    rays = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 6)
    count = np.empty(len(rays), np.uint32)
    cnt = abi.Counters()
    lib.orc_intersect_any_count(C.byref(g), _p(rays), len(rays), t_min, t_max, _p(count), C.byref(cnt))
    return count, cnt.as_dict()


def brute_force(triangles, count, rays):
    lib = load()
    tri_in = np.ascontiguousarray(triangles, dtype=np.float32)
    rays = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 6)
    n = len(rays)
    tri = np.empty(n, np.int32)
    t = np.empty(n)
    all_ = np.empty(n, np.uint32)
    lib.orc_brute_force(_p(tri_in), count, _p(rays), n, _p(tri), _p(t), _p(all_))
    return tri, t, all_


def wide_brute(geometry, rays, t_min=0.0, t_max=np.inf, cull=True, threads=None):
    """Precision-form checker: FP32 watertight closest hit over every triangle (orc_wide_brute).
    rays: (n, 6) float32. Returns tri (BVH order), t, u, v (float32), front."""
    lib = load()
    tris = np.ascontiguousarray(geometry.triangles, dtype=np.float32)
    sides = np.ascontiguousarray(geometry.sides, dtype=np.uint8)
    rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
    n = len(rays)
    tri = np.empty(n, np.int32)
    t = np.empty(n, np.float32)
    u = np.empty(n, np.float32)
    v = np.empty(n, np.float32)
    front = np.empty(n, np.uint8)
    lib.orc_wide_brute(_p(tris), _p(sides), len(sides), _p(rays), n, t_min, t_max, int(cull), _p(tri), _p(t),
                       _p(u), _p(v), _p(front), threads or min(16, os.cpu_count() or 1))
    return tri, t, u, v, front


def sample_texture(texture, u, v, libm=False):
    """textures.ts:247 sampleTexture on a types.CompiledTexture -> RGBA (4 floats)."""
    t = abi.Texture()
    texture.fill(t)
    out = np.zeros(4)
    load(libm).orc_sample_texture(C.byref(t), u, v, _p(out))
    return out


def sample_environment(env, textures, d, libm=False):
    """textures.ts:287 sampleEnvironment -> RGB."""
    e = abi.Environment()
    env.fill(e)
    ts = (abi.Texture * max(1, len(textures)))()
    for i, t in enumerate(textures):
        t.fill(ts[i])
    out = np.zeros(3)
    load(libm).orc_sample_environment(C.byref(e), ts, d[0], d[1], d[2], _p(out))
    return out


def render_tile(scene, camera, settings, tile, width, height, pass_, block=1, libm=False):
    """tracer.ts:62 renderTile -> float32 [tile.h*tile.w*4] premultiplied RGBA, plus counters."""
    lib = load(libm)
    g = scene.geometry.to_c()
    s = scene.shading.to_c()
    cam = camera.to_c()
    st = settings.to_c()
    out = np.zeros(tile.width * tile.height * 4, np.float32)
    cnt = abi.Counters()
    rc = lib.orc_render_tile(C.byref(g), C.byref(s), C.byref(cam), C.byref(st), tile.to_c(), width, height,
                             pass_, block, _p(out), C.byref(cnt))
    if rc != 0:
        raise RuntimeError(f'oracle render_tile failed: {rc}')
    return out, cnt.as_dict()


def raytrace(scene, camera, settings, width, height, samples=16, srgb=True, libm=False):
    """raytrace.ts:69-82 raytraceCompiled -> uint8 [h*w*4]."""
    lib = load(libm)
    g = scene.geometry.to_c()
    s = scene.shading.to_c()
    cam = camera.to_c()
    st = settings.to_c()
    out = np.zeros(width * height * 4, np.uint8)
    cnt = abi.Counters()
    rc = lib.orc_raytrace(C.byref(g), C.byref(s), C.byref(cam), C.byref(st), width, height, samples,
                          int(srgb), _p(out), C.byref(cnt))
    if rc != 0:
        raise RuntimeError(f'oracle raytrace failed: {rc}')
    return out, cnt.as_dict()


def encode_tile(image, width, tile, data, sum_, count, srgb):
    load().orc_encode_tile(_p(image), width, tile.to_c(), _p(data), _p(sum_), count, int(srgb))


def make_tiles(width, height, size):
    lib = load()
    n = lib.orc_make_tiles(width, height, size, None)
    out = (abi.Tile * max(1, n))()
    lib.orc_make_tiles(width, height, size, C.cast(out, C.c_void_p))
    return [(t.x, t.y, t.width, t.height) for t in out[:n]]


def render_items_mt(scene, camera, settings, jobs, threads, out=None, libm=False):
    """CPU baseline: tile jobs traced by a pool of threads. Returns (seconds, counters)."""
    lib = load(libm)
    g = scene.geometry.to_c()
    s = scene.shading.to_c()
    cam = camera.to_c()
    st = settings.to_c()
    arr = (abi.TileJob * len(jobs))(*jobs)
    cnt = abi.Counters()
    secs = lib.orc_render_items_mt(C.byref(g), C.byref(s), C.byref(cam), C.byref(st), C.cast(arr, C.c_void_p),
                                   len(jobs), threads, _p(out) if out is not None else None, C.byref(cnt))
    return secs, cnt.as_dict()


def math_fn(fn, x, y=0.0, libm=False):
    return load(libm).orc_math(fn, x, y)
