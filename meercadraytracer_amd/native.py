"""ctypes binding of the native library libmcr.so (include/mcr.h).

This is the product path: every entry point runs on the GPU through the C-ABI. There is no CPU
fallback: importing works anywhere, but the first call raises NativeError when the library is not
built or no GPU is visible.
"""
import ctypes as C
import os
import threading

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libmcr.so')
DEFAULT_SLOTS = 1 << 23  # the library's default path-slot count (csrc/context.h DEFAULT_CAPACITY)


class NativeError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f'{message} (code {code})')
        self.code = code
        self.message = message


_lib = None
_lock = threading.Lock()


def lib():
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise NativeError(abi.MCR_E_DEVICE, f'{LIB_PATH} is not built: run __graft_entry__.build()')
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        sig = {
            'mcr_create': (C.c_int, [C.c_int, C.POINTER(P)]),
            'mcr_destroy': (None, [P]),
            'mcr_last_error': (C.c_char_p, []),
            'mcr_version': (C.c_char_p, []),
            'mcr_build_bvh': (C.c_int, [P, P, C.c_uint32, P, P, P, C.POINTER(C.c_uint32)]),
            'mcr_build_bvh_device': (C.c_int, [P, P, C.c_uint32, P, P, P, C.POINTER(C.c_uint32)]),
            'mcr_scene_create': (C.c_int, [P, C.POINTER(P)]),
            'mcr_scene_set_geometry': (C.c_int, [P, P, C.c_uint32]),
            'mcr_scene_set_shading': (C.c_int, [P, P, C.c_uint32]),
            'mcr_scene_release': (None, [P]),
            'mcr_scene_bytes': (C.c_uint64, [P]),
            'mcr_render_tile': (C.c_int, [P, P, C.c_uint32, C.c_uint32, P, P, abi.Tile, C.c_uint32, C.c_uint32,
                                          C.c_uint32, C.c_uint32, P, P]),
            'mcr_render_tiles': (C.c_int, [P, P, P, P, P, C.c_uint32, P, C.c_int, P]),
            'mcr_render_batch': (C.c_int, [P, P, C.c_uint32, P, P, C.c_uint32, P, C.c_int, P]),
            'mcr_set_graph': (C.c_int, [P, C.c_int]),
            'mcr_set_shade_sort': (C.c_int, [P, C.c_int]),
            'mcr_set_intersection': (C.c_int, [P, C.c_int]),
            'mcr_server_create': (C.c_int, [P, P, C.POINTER(P)]),
            'mcr_server_destroy': (None, [P]),
            'mcr_server_add_scene': (C.c_int, [P, P]),
            'mcr_server_remove_scene': (C.c_int, [P, P]),
            'mcr_server_submit': (C.c_int, [P, P, C.POINTER(C.c_uint64)]),
            'mcr_server_cancel': (C.c_int, [P, C.c_uint64]),
            'mcr_server_status': (C.c_int, [P, C.c_uint64, P]),
            'mcr_server_wait': (C.c_int, [P, C.c_uint64, C.c_double]),
            'mcr_server_read': (C.c_int, [P, C.c_uint64, P]),
            'mcr_server_release': (C.c_int, [P, C.c_uint64]),
            'mcr_server_stats_get': (C.c_int, [P, P]),
            'mcr_encode_tile': (C.c_int, [P, P, C.c_uint32, abi.Tile, P, P, C.c_uint32, C.c_int, C.c_int]),
            'mcr_encode_tile64': (C.c_int, [P, P, C.c_uint32, abi.Tile, P, P, C.c_uint32, C.c_int]),
            'mcr_intersect_closest': (C.c_int, [P, P, P, C.c_uint32, C.c_double, C.c_double, C.c_int, P, P, P, P,
                                                P, C.c_int, P]),
            'mcr_intersect_closest_noalloc': (C.c_int, [P, P, C.c_uint32, C.c_double, C.c_double, C.c_int, P, P, P,
                                                        P, P, P]),
            'mcr_scene_build_wide': (C.c_int, [P, C.POINTER(abi.WideInfo)]),
            'mcr_scene_wide_info': (C.c_int, [P, C.POINTER(abi.WideInfo)]),
            'mcr_intersect_closest_wide': (C.c_int, [P, P, P, C.c_uint32, C.c_float, C.c_float, C.c_int, P, P, P,
                                                     P, P, C.c_int, P]),
            'mcr_compile_count': (C.c_int, [P, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_int)]),
            'mcr_compile_geometry': (C.c_int, [P, P, C.c_uint32, C.POINTER(abi.GeometryOut)]),
            'mcr_scene_compile_geometry': (C.c_int, [P, P, C.c_uint32, C.c_uint32, C.POINTER(abi.CompileInfo)]),
            'mcr_intersect_any_count': (C.c_int, [P, P, P, C.c_uint32, C.c_double, C.c_double, P, C.c_int, P]),
            'mcr_device_alloc': (C.c_int, [P, C.c_uint64, C.POINTER(P)]),
            'mcr_device_free': (C.c_int, [P, P]),
            'mcr_copy': (C.c_int, [P, P, P, C.c_uint64, C.c_int]),
            'mcr_memset': (C.c_int, [P, P, C.c_int, C.c_uint64]),
            'mcr_last_timing': (C.c_int, [P, C.POINTER(abi.Timing)]),
            'mcr_set_timing': (C.c_int, [P, C.c_int]),
            'mcr_set_wavefront_capacity': (C.c_int, [P, C.c_uint32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
        return L


def check(rc):
    if rc != abi.MCR_OK:
        msg = lib().mcr_last_error().decode(errors='replace')
        raise NativeError(rc, msg)


def _p(a):
    return None if a is None else a.ctypes.data


class DeviceBuffer:
    """Device memory owned through the C-ABI (no torch needed)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        ptr = C.c_void_p()
        check(lib().mcr_device_alloc(ctx.handle, self.nbytes, C.byref(ptr)))
        self.ptr = ptr.value

    def upload(self, array):
        a = np.ascontiguousarray(array)
        assert a.nbytes <= self.nbytes
        check(lib().mcr_copy(self.ctx.handle, self.ptr, a.ctypes.data, a.nbytes, 0))

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        check(lib().mcr_copy(self.ctx.handle, out.ctypes.data, self.ptr, out.nbytes, 1))
        return out

    def zero(self):
        check(lib().mcr_memset(self.ctx.handle, self.ptr, 0, self.nbytes))

    def free(self):
        if self.ptr:
            lib().mcr_device_free(self.ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One GPU (mcr_create)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().mcr_create(device, C.byref(h)))
        self.handle = h.value
        self.device = device

    def close(self):
        if self.handle:
            lib().mcr_destroy(self.handle)
            self.handle = None

    def set_timing(self, on=True):
        check(lib().mcr_set_timing(self.handle, int(on)))

    def set_graph(self, on=True):
        check(lib().mcr_set_graph(self.handle, int(on)))

    def set_shade_sort(self, on=True):
        """Material-sorted shading (bin live paths by shade class before shade); same results."""
        check(lib().mcr_set_shade_sort(self.handle, int(on)))

    def set_intersection(self, mode):
        """abi.MCR_ISECT_FP64 (default) or abi.MCR_ISECT_WIDE_FP32 for renders on this context."""
        check(lib().mcr_set_intersection(self.handle, int(mode)))

    def set_capacity(self, slots=None):
        """Path slots of this context's wavefront; None restores the library default."""
        slots = DEFAULT_SLOTS if slots is None else slots
        check(lib().mcr_set_wavefront_capacity(self.handle, int(slots)))

    def last_timing(self):
        t = abi.Timing()
        check(lib().mcr_last_timing(self.handle, C.byref(t)))
        return {f: getattr(t, f) for f, _ in t._fields_}

    def build_bvh(self, triangles, count):
        """bvh.ts:16 buildBvh on the GPU -> (bounds, nodes, order, nodeCount)."""
        tri = np.ascontiguousarray(triangles, dtype=np.float32)
        m = max(1, 2 * count - 1)
        bounds = np.zeros(m * 6, np.float32)
        nodes = np.zeros(m * 3, np.int32)
        order = np.zeros(max(count, 1), np.uint32)
        nc = C.c_uint32(0)
        check(lib().mcr_build_bvh(self.handle, _p(tri), count, _p(bounds), _p(nodes), _p(order), C.byref(nc)))
        n = nc.value
        return bounds[:n * 6].copy(), nodes[:n * 3].copy(), order[:count].copy(), n


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def build_bvh(triangles, count):
    return default_context().build_bvh(triangles, count)


class Scene:
    """A scene resident in HBM (mcr_scene): geometry and shading uploaded once per version."""

    def __init__(self, ctx=None):
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        check(lib().mcr_scene_create(self.ctx.handle, C.byref(h)))
        self.handle = h.value
        self.geometry_version = None
        self.shading_version = None

    def set_geometry(self, geometry, version):
        g = geometry.to_c()
        check(lib().mcr_scene_set_geometry(self.handle, C.byref(g), version))
        self.geometry_version = version

    def compile_meshes(self, meshes, version):
        """compileGeometry straight into this scene (mcr_scene_compile_geometry): mesh descriptors as
        compile_meshes takes them; the arrays never leave the device. Returns triangle and node
        counts, the epsilon and whether the scene fell back to the host layout."""
        arr, keep = _mesh_array(meshes)
        info = abi.CompileInfo()
        check(lib().mcr_scene_compile_geometry(self.handle, C.cast(arr, C.c_void_p), len(meshes), version,
                                               C.byref(info)))
        self.geometry_version = version
        return {'triangle_count': info.triangle_count, 'node_count': info.node_count,
                'host_layout': bool(info.host_layout), 'epsilon': info.epsilon}

    def set_shading(self, shading, version):
        s = shading.to_c()
        check(lib().mcr_scene_set_shading(self.handle, C.byref(s), version))
        self.shading_version = version

    def bytes(self):
        return int(lib().mcr_scene_bytes(self.handle))

    def release(self):
        if self.handle:
            lib().mcr_scene_release(self.handle)
            self.handle = None

    def render_tile(self, camera, settings, tile, width, height, pass_, block=1, counters=None):
        """tracer.ts:62 renderTile on the GPU -> float32 premultiplied RGBA of the tile."""
        out = np.zeros(tile.width * tile.height * 4, np.float32)
        cam = camera.to_c()
        st = settings.to_c()
        cnt = counters if counters is not None else abi.Counters()
        check(lib().mcr_render_tile(self.ctx.handle, self.handle, self.geometry_version, self.shading_version,
                                    C.byref(cam), C.byref(st), tile.to_c(), width, height, pass_, block,
                                    _p(out), C.byref(cnt)))
        return out

    def render_jobs(self, camera, settings, jobs, out=None, out_device=False, counters=None):
        """Batch of mcr_tile_job (abi.TileJob) in one device run. Returns the host output array
        (or None when `out` is a device pointer)."""
        cam = camera.to_c()
        st = settings.to_c()
        arr = (abi.TileJob * max(1, len(jobs)))(*jobs)
        cnt = counters if counters is not None else abi.Counters()
        if out is None:
            total = max((j.out_offset + j.tile.width * j.tile.height * 4 for j in jobs), default=0)
            out = np.zeros(total, np.float32)
        ptr = out if out_device else _p(out)
        check(lib().mcr_render_tiles(self.ctx.handle, self.handle, C.byref(cam), C.byref(st),
                                     C.cast(arr, C.c_void_p), len(jobs), ptr, int(out_device), C.byref(cnt)))
        return None if out_device else out

    def intersect_closest(self, rays, t_min=0.0, t_max=np.inf, cull=True, counters=None):
        rays = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 6)
        n = len(rays)
        tri = np.empty(n, np.int32)
        t = np.empty(n)
        u = np.empty(n)
        v = np.empty(n)
        front = np.empty(n, np.uint8)
        cnt = counters if counters is not None else abi.Counters()
        check(lib().mcr_intersect_closest(self.ctx.handle, self.handle, _p(rays), n, t_min, t_max, int(cull),
                                          _p(tri), _p(t), _p(u), _p(v), _p(front), 0, C.byref(cnt)))
        return tri, t, u, v, front

    def build_wide(self):
        """Builds the precision form's 8-wide BVH (mcr_scene_build_wide); returns its info."""
        info = abi.WideInfo()
        check(lib().mcr_scene_build_wide(self.handle, C.byref(info)))
        return info.as_dict()

    def intersect_closest_wide(self, rays, t_min=0.0, t_max=np.inf, cull=True, counters=None, device=None):
        """FP32 closest hits over the wide BVH. rays: (n, 6) float32 (host), or device=(ptr, n)
        with output device pointers in device[2:] (tri, t, u, v, front; None to skip)."""
        cnt = counters if counters is not None else abi.Counters()
        if device is not None:
            ptr, n, *outs = device
            outs = (list(outs) + [None] * 5)[:5]
            check(lib().mcr_intersect_closest_wide(self.ctx.handle, self.handle, ptr, n, t_min, t_max, int(cull),
                                                   *outs, 1, C.byref(cnt)))
            return None
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        n = len(rays)
        tri = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        u = np.empty(n, np.float32)
        v = np.empty(n, np.float32)
        front = np.empty(n, np.uint8)
        check(lib().mcr_intersect_closest_wide(self.ctx.handle, self.handle, _p(rays), n, t_min, t_max, int(cull),
                                               _p(tri), _p(t), _p(u), _p(v), _p(front), 0, C.byref(cnt)))
        return tri, t, u, v, front

    def intersect_any_count(self, rays, t_min=0.0, t_max=np.inf, counters=None):
        rays = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 6)
        count = np.empty(len(rays), np.uint32)
        cnt = counters if counters is not None else abi.Counters()
        check(lib().mcr_intersect_any_count(self.ctx.handle, self.handle, _p(rays), len(rays), t_min, t_max,
                                            _p(count), 0, C.byref(cnt)))
        return count

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def _mesh_array(meshes):
    """The mcr_mesh array of mesh descriptors (see compile_meshes), with the numpy arrays it points
    into (keep them alive while the array is used)."""
    keep = []
    arr = (abi.Mesh * max(1, len(meshes)))()
    for i, m in enumerate(meshes):
        d = arr[i]
        d.world[:] = [float(x) for x in m['world']]
        for name, dt, width in (('positions', np.float32, 3), ('normals', np.float32, 3), ('uvs', np.float32, 2)):
            a = m.get(name)
            if a is not None:
                a = np.ascontiguousarray(a, dtype=dt).reshape(-1, width)
                keep.append(a)
                setattr(d, name, a.ctypes.data)
                d.vertex_count = len(a)
        c = m.get('colors')
        if c is not None:
            c = np.ascontiguousarray(c, dtype=np.float32)
            keep.append(c)
            d.colors, d.color_size = c.ctypes.data, c.shape[1]
        idx = m.get('index')
        if idx is not None:
            idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
            keep.append(idx)
            d.index, d.index_count = idx.ctypes.data, len(idx)
        draw = m.get('draw') or (0, None)
        d.draw_start = draw[0]
        d.draw_count = abi.INF_COUNT if draw[1] is None else draw[1]
        groups = (abi.MeshGroup * max(1, len(m['groups'])))()
        for k, (start, count, material, flags, side, flat, colors) in enumerate(m['groups']):
            groups[k] = abi.MeshGroup(start, abi.INF_COUNT if count is None else count, material, flags, side,
                                      int(bool(flat)), int(bool(colors)))
        keep.append(groups)
        d.groups, d.group_count = C.cast(groups, C.c_void_p), len(m['groups'])
    return arr, keep


def compile_meshes(ctx, meshes):
    """compileGeometry on the device (mcr_compile_geometry). meshes: list of dicts with 'world'
    (16 floats, column-major), 'positions' (v, 3) float32 and optional 'normals', 'uvs', 'colors'
    ((v, 3|4) float32), 'index' (uint32), 'draw' ((start, count) or None) and 'groups': list of
    (start, count, material, flags, side, flat_shading, vertex_colors), count None = Infinity.
    Returns the arrays of a CompiledGeometry as a dict."""
    ctx = ctx or default_context()
    arr, keep = _mesh_array(meshes)
    n = C.c_uint32()
    any_colors = C.c_int()
    check(lib().mcr_compile_count(C.cast(arr, C.c_void_p), len(meshes), C.byref(n), C.byref(any_colors)))
    n = n.value
    nodes = max(1, 2 * n - 1)
    out = {'triangles': np.empty(n * 9, np.float32), 'normals': np.empty(n * 9, np.float32),
           'uvs': np.empty(n * 6, np.float32), 'tangents': np.empty(n * 4, np.float32),
           'colors': np.empty(n * 12 if any_colors.value else 0, np.float32),
           'materials': np.empty(n, np.uint32), 'flags': np.empty(n, np.uint8), 'sides': np.empty(n, np.uint8),
           'bounds': np.empty(nodes * 6, np.float32), 'nodes': np.empty(nodes * 3, np.int32)}
    g = abi.GeometryOut()
    g.triangle_count = n
    for f in ('triangles', 'normals', 'uvs', 'tangents', 'colors', 'materials', 'flags', 'sides'):
        setattr(g, f, out[f].ctypes.data if out[f].size else None)
    g.bvh_bounds, g.bvh_nodes = out['bounds'].ctypes.data, out['nodes'].ctypes.data
    check(lib().mcr_compile_geometry(ctx.handle, C.cast(arr, C.c_void_p), len(meshes), C.byref(g)))
    out['bounds'] = out['bounds'][:g.node_count * 6]
    out['nodes'] = out['nodes'][:g.node_count * 3]
    out['count'], out['node_count'], out['epsilon'] = n, g.node_count, g.epsilon
    return out


def encode_tile64_device(ctx, image_ptr, width, tile, data_ptr, sum_ptr, count, srgb):
    """writeTile with the running sum in FP64 (renders beyond 1024 samples)."""
    check(lib().mcr_encode_tile64(ctx.handle, image_ptr, width, tile.to_c(), data_ptr, sum_ptr, count, int(srgb)))


def render_batch(ctx, views, jobs, job_views, out=None, out_device=False, counters=None):
    """Tile jobs of many views (scene, camera, settings) in one wavefront run (mcr_render_batch).
    views: [(Scene, CompiledCamera, TraceSettings)]; jobs: [abi.TileJob]; job_views: view per job."""
    ctx = ctx or default_context()
    keep = []
    varr = (abi.View * max(1, len(views)))()
    for i, (scene, camera, settings) in enumerate(views):
        cam, st = camera.to_c(), settings.to_c()
        keep += [cam, st]
        varr[i] = abi.View(scene.handle, C.cast(C.pointer(cam), C.c_void_p), C.cast(C.pointer(st), C.c_void_p))
    jarr = (abi.TileJob * max(1, len(jobs)))(*jobs)
    vidx = np.ascontiguousarray(job_views, dtype=np.uint32)
    cnt = counters if counters is not None else abi.Counters()
    if out is None:
        total = max((j.out_offset + j.tile.width * j.tile.height * 4 for j in jobs), default=0)
        out = np.zeros(total, np.float32)
    ptr = out if out_device else _p(out)
    check(lib().mcr_render_batch(ctx.handle, C.cast(varr, C.c_void_p), len(views), C.cast(jarr, C.c_void_p),
                                 _p(vidx), len(jobs), ptr, int(out_device), C.byref(cnt)))
    return None if out_device else out


def encode_tile_device(ctx, image_ptr, width, tile, data_ptr, sum_ptr, count, srgb):
    """progressive.ts:288 writeTile on device buffers."""
    check(lib().mcr_encode_tile(ctx.handle, image_ptr, width, tile.to_c(), data_ptr, sum_ptr, count, int(srgb), 1))


class Server:
    """The served form (mcr_server_*): a request loop over scenes resident in HBM. Each request is
    one progressive render (camera, settings, size, samples, tile size, preview block)."""

    def __init__(self, ctx=None, hbm_budget=0, slots=0, window=0, max_requests=0, use_graph=True, max_lights=0,
                 lanes=0, preview_slots=0, preview_cus=0, preview_passes=False):
        self.ctx = ctx or default_context()
        cfg = abi.ServerConfig(int(hbm_budget), int(slots), int(window), int(max_requests), int(bool(use_graph)),
                               int(max_lights), int(lanes), int(preview_slots), int(preview_cus),
                               1 if preview_passes else 0)
        h = C.c_void_p()
        check(lib().mcr_server_create(self.ctx.handle, C.byref(cfg), C.byref(h)))
        self.handle = h.value
        self.sizes = {}

    def add_scene(self, scene):
        check(lib().mcr_server_add_scene(self.handle, scene.handle))

    def remove_scene(self, scene):
        check(lib().mcr_server_remove_scene(self.handle, scene.handle))

    def submit(self, scene, camera, settings, width, height, samples, tile_size=64, preview_block=4, srgb=True,
               geometry_version=0, shading_version=0):
        r = abi.Request(scene.handle, camera.to_c(), settings.to_c(), width, height, samples, tile_size,
                        preview_block, int(bool(srgb)), int(geometry_version), int(shading_version))
        rid = C.c_uint64()
        check(lib().mcr_server_submit(self.handle, C.byref(r), C.byref(rid)))
        self.sizes[rid.value] = (width, height)
        return rid.value

    def cancel(self, rid):
        check(lib().mcr_server_cancel(self.handle, rid))

    def status(self, rid):
        st = abi.RequestStatus()
        check(lib().mcr_server_status(self.handle, rid, C.byref(st)))
        return st.as_dict()

    def wait(self, rid, timeout_ms=-1.0):
        """Returns the final state (done / cancelled); raises on failure or timeout."""
        rc = lib().mcr_server_wait(self.handle, rid, float(timeout_ms))
        if rc == abi.MCR_E_CANCELLED:
            return abi.REQ_CANCELLED
        check(rc)
        return abi.REQ_DONE

    def read(self, rid):
        w, h = self.sizes[rid]
        out = np.zeros(w * h * 4, np.uint8)
        check(lib().mcr_server_read(self.handle, rid, _p(out)))
        return out

    def release(self, rid):
        check(lib().mcr_server_release(self.handle, rid))
        self.sizes.pop(rid, None)

    def stats(self):
        st = abi.ServerStats()
        check(lib().mcr_server_stats_get(self.handle, C.byref(st)))
        return st.as_dict()

    def close(self):
        if self.handle:
            lib().mcr_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
