"""Stand-ins for exercising the scheduler on the CPU: an executor that renders tiles with the
oracle, an image that encodes with the oracle, and a fake WorkerLike (tests/progressive.test.ts:
14-36): messages are copied both ways and answered by a handler."""
import copy
import dataclasses

import numpy as np

import oracle
from meercadraytracer_amd.executors import ExecutorError


class OracleHandler:
    """createTileHandler (protocol.ts:55-110) over the oracle: texture refs are joined to the
    pixels of earlier textures messages when a tile is rendered."""

    def __init__(self):
        self.geometry = self.shading = None
        self.gv = self.sv = None
        self.pixels = {}

    def __call__(self, m):
        if m['type'] == 'geometry':
            self.geometry, self.gv = m['geometry'], m['version']
        elif m['type'] == 'textures':
            for item in m['items']:
                self.pixels[item['id']] = item['data']
        elif m['type'] == 'shading':
            self.shading, self.sv = m['shading'], m['version']
        elif m['type'] == 'tile':
            if self.gv != m['geometryVersion'] or self.sv != m['shadingVersion']:
                return {'type': 'error', 'job': m['job'], 'message': 'the tile refers to a scene this worker has not received'}
            missing = [r.id for r in self.shading.textures if r.id not in self.pixels]
            if missing:
                return {'type': 'error', 'job': m['job'], 'message': f'texture {missing[0]} was not sent'}
            from meercadraytracer_amd.types import CompiledScene
            shading = dataclasses.replace(self.shading, textures=[r.with_data(self.pixels[r.id])
                                                                  for r in self.shading.textures])
            data, _ = oracle.render_tile(CompiledScene(self.geometry, shading), m['camera'], m['settings'],
                                         m['tile'], m['width'], m['height'], m['pass'], m['block'])
            return {'type': 'tile', 'job': m['job'], 'data': data}
        return None


class OracleExecutor:
    def __init__(self, concurrency=1, fail_after=None):
        self.textures = set()
        self.concurrency = concurrency
        self.handle = OracleHandler()
        self.fail_after = fail_after
        self.rendered = 0
        self.disposed = False

    def send(self, m):
        self.handle(m)

    def render_batch(self, messages):
        out = []
        for m in messages:
            if self.fail_after is not None and self.rendered >= self.fail_after:
                out.append(ExecutorError('worker failed'))
                continue
            self.rendered += 1
            r = self.handle(m)
            out.append(r['data'] if r['type'] == 'tile' else ExecutorError(r['message']))
        return out

    def dispose(self):
        self.disposed = True


class OracleImage:
    def __init__(self, ctx, width, height):
        self.width, self.height = width, height
        self.image = np.zeros(width * height * 4, np.uint8)
        self.sum = np.zeros(width * height * 4, np.float32)

    def write_tile(self, tile, data, accumulate, count, srgb):
        oracle.encode_tile(self.image, self.width, tile, np.ascontiguousarray(data, np.float32),
                           self.sum if accumulate else None, count, srgb)

    def download(self):
        return self.image.copy()

    def free(self):
        pass


class FakeWorker:
    """A stand-in Web Worker: every message is deep-copied (structured clone) both ways and
    answered by the handler; responses are delivered when flush() runs (the event loop turn)."""

    def __init__(self, handler, fail=False):
        self.received = []
        self.terminated = False
        self.handle = handler
        self.listeners = {}
        self.queue = []
        self.fail = fail

    def post_message(self, message):
        m = copy.deepcopy(message)
        self.received.append(m)
        self.queue.append(m)

    def add_event_listener(self, kind, fn):
        self.listeners[kind] = fn

    def flush(self):
        if self.fail:
            self.queue.clear()
            if 'error' in self.listeners:
                self.listeners['error']({'message': 'script failed to load'})
            return
        while self.queue:
            r = self.handle(self.queue.pop(0))
            if r is not None and 'message' in self.listeners:
                self.listeners['message']({'data': copy.deepcopy(r)})

    def terminate(self):
        self.terminated = True
