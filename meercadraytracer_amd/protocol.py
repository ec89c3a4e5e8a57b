"""The renderer's side of a worker (src/protocol.ts:29-127), backed by a scene resident in HBM.

Messages keep the reference's shapes (protocol.ts:29-49), as dicts:
  {'type': 'geometry', 'version', 'geometry': CompiledGeometry}
  {'type': 'textures', 'items': [{'id', 'data'}]}
  {'type': 'shading', 'version', 'shading': ShadingPayload}
  {'type': 'tile', 'job', 'geometryVersion', 'shadingVersion', 'camera', 'settings', 'tile',
   'width', 'height', 'pass', 'block'}
Responses: {'type': 'tile', 'job', 'data': float32 RGBA} or {'type': 'error', 'job', 'message'}.
Geometry and shading are uploaded to the GPU once per version (mcr_scene_set_*), as the reference
keeps the latest of each (protocol.ts:64-74); a tile naming another version gets the reference's
error text (protocol.ts:77-79). The shading payload names its textures by id (TextureRef); their
pixels arrive in textures messages and are joined to the refs when a tile first needs the shading
(protocol.ts:81-87), so a missing texture is reported with the tile, in the reference's words.
"""
import dataclasses

from . import abi, native

NO_SCENE = 'the tile refers to a scene this worker has not received'


class TileHandler:
    """createTileHandler (protocol.ts:55-110): a callable over RequestMessages."""

    def __init__(self, ctx=None):
        self.ctx = ctx
        self.scene = None
        self.geometry_version = None
        self.shading_version = None
        self.shading = None
        self.pixels = {}
        self._uploaded = None  # shading version on the device

    def _gpu(self):
        if self.scene is None:
            self.scene = native.Scene(self.ctx)
        return self.scene

    def __call__(self, message):
        kind = message.get('type')
        if kind == 'geometry':
            self._gpu().set_geometry(message['geometry'], message['version'])
            self.geometry_version = message['version']
            return None
        if kind == 'textures':
            for item in message['items']:
                self.pixels[item['id']] = item['data']
            return None
        if kind == 'shading':
            self.shading = message['shading']
            self.shading_version = message['version']
            return None
        if kind == 'tile':
            return self.tile(message)
        return None

    def tile(self, m):
        job = m['job']
        try:
            if (self.geometry_version is None or self.geometry_version != m['geometryVersion'] or
                    self.shading_version is None or self.shading_version != m['shadingVersion']):
                raise RuntimeError(NO_SCENE)
            self._upload_shading()
            data = self.scene.render_tile(m['camera'], m['settings'], m['tile'], m['width'], m['height'],
                                          m['pass'], m.get('block', 1))
            return {'type': 'tile', 'job': job, 'data': data}
        except (RuntimeError, native.NativeError) as e:
            return {'type': 'error', 'job': job, 'message': getattr(e, 'message', None) or str(e)}

    def _upload_shading(self):
        """Joins the payload's texture refs to their pixels and uploads the shading, once per
        version (protocol.ts:81-87)."""
        if self._uploaded == self.shading_version:
            return
        textures = []
        for ref in self.shading.textures:
            data = self.pixels.get(ref.id)
            if data is None:
                raise RuntimeError(f'texture {ref.id} was not sent')
            textures.append(ref.with_data(data))
        self._gpu().set_shading(dataclasses.replace(self.shading, textures=textures), self.shading_version)
        self._uploaded = self.shading_version

    def tiles(self, messages):
        """Several tile messages of one camera and settings in one device run (the batched form
        of the GPU worker). Returns one response per message, in order."""
        if not messages:
            return []
        m0 = messages[0]
        ok = (self.geometry_version is not None and self.shading_version is not None and
              all(m['geometryVersion'] == self.geometry_version and m['shadingVersion'] == self.shading_version and
                  m['camera'] is m0['camera'] and m['settings'] is m0['settings'] for m in messages))
        if not ok:
            return [self.tile(m) for m in messages]
        try:
            self._upload_shading()
        except (RuntimeError, native.NativeError) as e:
            msg = getattr(e, 'message', None) or str(e)
            return [{'type': 'error', 'job': m['job'], 'message': msg} for m in messages]
        jobs, off = [], 0
        for m in messages:
            t = m['tile']
            jobs.append(abi.TileJob(t.to_c(), m['width'], m['height'], m['pass'], m.get('block', 1), off))
            off += t.width * t.height * 4
        try:
            out = self.scene.render_jobs(m0['camera'], m0['settings'], jobs)
        except native.NativeError as e:
            return [{'type': 'error', 'job': m['job'], 'message': e.message} for m in messages]
        res = []
        for m, j in zip(messages, jobs):
            n = m['tile'].width * m['tile'].height * 4
            res.append({'type': 'tile', 'job': m['job'], 'data': out[j.out_offset:j.out_offset + n]})
        return res

    def release(self):
        if self.scene is not None:
            self.scene.release()
            self.scene = None
        self._uploaded = None


def create_tile_handler(ctx=None):
    return TileHandler(ctx)
