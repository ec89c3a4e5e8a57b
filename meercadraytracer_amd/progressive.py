"""Progressive rendering (src/progressive.ts): tiles, scene sync and the tile scheduler.

  make_tiles            progressive.ts:160-170  row-major tiles, stable sort by distance of the
                        tile centre to the image centre (Math.hypot as V8 evaluates it)
  geometry_key          progressive.ts:53-73    the key whose change triggers a geometry rebuild
  SceneSync             progressive.ts:80-157   compile what changed; send each executor what it
                        lacks, once per version
  ProgressiveRenderer   progressive.ts:177-282  work list = block preview of every tile, then
                        pass-major tiles; generation-based cancellation; failing executors give
                        their tiles back and are dropped, with one fall back to an in-process GPU
                        executor; frames throttled by frameInterval
  write_tile            progressive.ts:288-313  the FP32 running sum and the 8-bit straight-alpha
                        image live in HBM and every tile is accumulated there (mcr_encode_tile)
Executors take up to `concurrency` tiles at once (render_batch), so a GPU executor traces a whole
batch of tiles in one wavefront run; results are written in work-list order.
"""
import dataclasses
import json
import math
import threading
import time
import warnings
import weakref

import numpy as np

from . import native
from .compile import add_mesh, compile_camera, compile_geometry, compile_shading, js_hypot
from .executors import LocalExecutor
from .types import TextureRef, Tile


def make_tiles(width, height, size):  # progressive.ts:160-170
    tiles = []
    for y in range(0, height, size):
        for x in range(0, width, size):
            tiles.append(Tile(x, y, min(size, width - x), min(size, height - y)))
    cx, cy = width / 2, height / 2
    return sorted(tiles, key=lambda t: js_hypot(t.x + t.width / 2 - cx, t.y + t.height / 2 - cy))


def geometry_key(meshes):  # progressive.ts:53-73
    parts = []
    for m in meshes:
        g = m.geometry
        parts.append(f'{id(m)}:{id(g)}:{getattr(g, "version", 0)}:{int(m.castShadow)}{int(m.receiveShadow)}:' +
                     ','.join(repr(v) for v in m.matrixWorld))
        mat = m.material
        if mat is not None:
            parts.append(f'{id(mat)}:{mat.side}:{int(mat.transparent)}:{int(mat.alphaTest > 0)}:'
                         f'{int(getattr(mat, "transmission", 0) > 0)}:{int(mat.wireframe)}:{int(mat.visible)}:'
                         f'{int(mat.vertexColors)}:{int(mat.flatShading)}')
    return '|'.join(parts)


_texture_ids = {}  # id(pixel array) -> (weak reference, texture id)
_next_texture_id = [0]


def texture_id(data):  # progressive.ts:42-48: one id per pixel array, for the array's lifetime
    entry = _texture_ids.get(id(data))
    if entry is not None and entry[0]() is data:
        return entry[1]
    _next_texture_id[0] += 1
    key = id(data)
    _texture_ids[key] = (weakref.ref(data, lambda _r, k=key: _texture_ids.pop(k, None)), _next_texture_id[0])
    return _next_texture_id[0]


def _shading_key(shading):
    def enc(o):
        return o.__dict__ if hasattr(o, '__dict__') else str(o)
    return json.dumps(shading, default=enc, sort_keys=True)


class SceneSync:  # progressive.ts:80-157
    def __init__(self, build_bvh=None):
        self.geometry = None
        self.geometry_version = 0
        self.shading_version = 0
        self.shading = None
        self._geometry_key = ''
        self._shading_key = ''
        self._sent = {}
        self._textures = []          # the current shading's CompiledTextures, with pixels
        self._texture_id_list = []   # their ids, as the payload's refs name them
        self.build_bvh = build_bvh

    def update(self, scene, camera=None):
        shading, meshes, index = compile_shading(scene)
        key = geometry_key(meshes)
        if self.geometry is None or key != self._geometry_key:
            parts = [add_mesh(m, index.get(id(m.material), 0), m.material) for m in meshes]
            self.geometry = compile_geometry(parts, self.build_bvh)
            self._geometry_key = key
            self.geometry_version += 1
        payload = self._to_payload(shading)
        skey = _shading_key(payload)
        if skey != self._shading_key:
            self.shading = payload
            self._shading_key = skey
            self._textures = shading.textures
            self.shading_version += 1

    def _to_payload(self, shading):  # progressive.ts:111-124: texture refs in place of the pixels
        self._texture_id_list = [texture_id(t.data) for t in shading.textures]
        refs = [TextureRef.of(t, i) for t, i in zip(shading.textures, self._texture_id_list)]
        return dataclasses.replace(shading, textures=refs)

    def sync(self, executor):
        state = self._sent.setdefault(id(executor), {'geometry': 0, 'shading': 0})
        if state['geometry'] != self.geometry_version and self.geometry is not None:
            executor.send({'type': 'geometry', 'version': self.geometry_version, 'geometry': self.geometry})
            state['geometry'] = self.geometry_version
        if state['shading'] != self.shading_version and self.shading is not None:
            items = []
            for t, tid in zip(self._textures, self._texture_id_list):
                if tid not in executor.textures:
                    items.append({'id': tid, 'data': t.data})
                    executor.textures.add(tid)
            if items:
                executor.send({'type': 'textures', 'items': items})
            executor.send({'type': 'shading', 'version': self.shading_version, 'shading': self.shading})
            state['shading'] = self.shading_version

    def forget(self, executor):
        self._sent.pop(id(executor), None)


class _DeviceImage:
    """The frame's 8-bit image and FP32 running sum in HBM (progressive.ts:200-201)."""

    def __init__(self, ctx, width, height):
        self.ctx, self.width, self.height = ctx, width, height
        n = width * height
        self.image = native.DeviceBuffer(ctx, n * 4)
        self.sum = native.DeviceBuffer(ctx, n * 16)
        self.image.zero()
        self.sum.zero()
        self.staging = None

    def write_tile(self, tile, data, accumulate, count, srgb):  # progressive.ts:288-313
        need = data.nbytes
        if self.staging is None or self.staging.nbytes < need:
            if self.staging is not None:
                self.staging.free()
            self.staging = native.DeviceBuffer(self.ctx, max(need, 1 << 16))
        self.staging.upload(data)
        native.encode_tile_device(self.ctx, self.image.ptr, self.width, tile, self.staging.ptr,
                                  self.sum.ptr if accumulate else None, count, srgb)

    def download(self):
        return self.image.download(np.uint8, self.width * self.height * 4)

    def free(self):
        for b in (self.image, self.sum, self.staging):
            if b is not None:
                b.free()


class ProgressiveRenderer:  # progressive.ts:177-282
    """image_factory(ctx, width, height) and local_factory(ctx) default to the HBM image and the
    in-process GPU executor; tests of the scheduling logic swap in host stand-ins."""

    def __init__(self, executors, ctx=None, build_bvh=None, image_factory=None, local_factory=None):
        self.image_factory = image_factory or _DeviceImage
        self.local_factory = local_factory or LocalExecutor
        self.executors = list(executors)
        self.generation = 0
        self.job = 0
        self.fell_back = False
        self.ctx = ctx
        self.sync = SceneSync(build_bvh)
        self._lock = threading.Lock()

    def stop(self):
        with self._lock:
            self.generation += 1

    def render(self, scene, camera, options, on_frame):
        with self._lock:
            self.generation += 1
            generation = self.generation
        width, height = options['width'], options['height']
        if width <= 0 or height <= 0 or not self.executors:
            return False
        self.sync.update(scene, camera)
        compiled_camera = compile_camera(camera, width / height if options.get('fitAspect') else None)
        ctx = self.ctx if self.ctx is not None or self.image_factory is not _DeviceImage else native.default_context()
        img = self.image_factory(ctx, width, height)
        try:
            return self._run(generation, compiled_camera, options, on_frame, img)
        finally:
            img.free()

    def _run(self, generation, camera, options, on_frame, img):
        width, height = options['width'], options['height']
        tiles = make_tiles(width, height, options['tileSize'])
        tile_samples = [0] * len(tiles)
        srgb = options.get('outputColorSpace', 'srgb') == 'srgb'
        max_samples = max(1, math.floor(options['samples']))
        work = []
        if options.get('previewBlock', 1) > 1:
            work += [{'tile': i, 'pass': 0, 'block': math.floor(options['previewBlock'])} for i in range(len(tiles))]
        for p in range(max_samples):
            work += [{'tile': i, 'pass': p, 'block': 1} for i in range(len(tiles))]
        state = {'next': 0, 'last': 0.0}
        live = lambda: generation == self.generation

        def min_samples():
            return min(tile_samples) if tile_samples else 0

        def emit(force):
            now = time.time() * 1000
            if not force and now - state['last'] < options.get('frameInterval', 100):
                return
            state['last'] = now
            s = min_samples()
            on_frame({'image': img.download(), 'width': width, 'height': height, 'samples': s,
                      'maxSamples': max_samples, 'done': s >= max_samples})

        while live() and state['next'] < len(work) and self.executors:
            for executor in list(self.executors):
                if not live() or state['next'] >= len(work):
                    break
                batch = work[state['next']:state['next'] + max(1, executor.concurrency)]
                state['next'] += len(batch)
                self.sync.sync(executor)
                messages = []
                for item in batch:
                    self.job += 1
                    messages.append({'type': 'tile', 'job': self.job, 'geometryVersion': self.sync.geometry_version,
                                     'shadingVersion': self.sync.shading_version, 'camera': camera,
                                     'settings': options['settings'], 'tile': tiles[item['tile']], 'width': width,
                                     'height': height, 'pass': item['pass'], 'block': item['block']})
                results = executor.render_batch(messages)
                if not live():
                    return False
                failed = [item for item, r in zip(batch, results) if isinstance(r, Exception)]
                for item, r in zip(batch, results):
                    if isinstance(r, Exception):
                        continue
                    tile = tiles[item['tile']]
                    if item['block'] > 1:
                        if tile_samples[item['tile']] == 0:
                            img.write_tile(tile, r, False, 1, srgb)
                    else:
                        tile_samples[item['tile']] += 1
                        img.write_tile(tile, r, True, tile_samples[item['tile']], srgb)
                    emit(False)
                if failed:
                    # Give the tiles back and drop this executor; the others carry on.
                    work[state['next']:state['next']] = failed
                    error = next(r for r in results if isinstance(r, Exception))
                    self.executors = [e for e in self.executors if e is not executor]
                    self.sync.forget(executor)
                    if not self.executors:
                        if isinstance(executor, LocalExecutor) or self.fell_back:
                            raise error
                        self.fell_back = True
                        warnings.warn(f'meercad-raytracer: the ray tracing workers failed ({error}); '
                                      'tracing in-process instead')
                        self.executors.append(self.local_factory(self.ctx))
        if not live():
            return False
        emit(True)
        return min_samples() >= max_samples

    def dispose(self):
        self.stop()
        for e in self.executors:
            e.dispose()
        self.executors = []
