"""Tile executors (src/executors.ts:4-94): things that render tile messages.

TileExecutor contract (executors.ts:4-14): send(message) processes scene messages in order;
render(message) returns the tile's premultiplied RGBA or raises; `textures` holds the texture ids
the executor already has; `concurrency` is how many tiles it takes at once; dispose() releases it.
render_batch(messages) renders up to `concurrency` tiles together and returns one result (array or
exception) per message, in order.

  LocalExecutor    the in-process GPU handler (the reference's page-side executor, concurrency 1)
  GpuExecutor      the same handler with a deep queue: each batch of tiles is one device run
  WorkerExecutor   wraps a WorkerLike (post_message / add_event_listener / terminate), with the
                   reference's pending-job map and error fan-out (executors.ts:54-94)
"""
from .protocol import create_tile_handler


class ExecutorError(RuntimeError):
    pass


class LocalExecutor:
    def __init__(self, ctx=None, concurrency=1):
        self.textures = set()
        self.concurrency = concurrency
        self.handle = create_tile_handler(ctx)

    def send(self, message):
        self.handle(message)

    def render(self, message):
        response = self.handle(message)
        if response and response['type'] == 'tile':
            return response['data']
        raise ExecutorError(response['message'] if response else 'no response')

    def render_batch(self, messages):
        out = []
        for r in self.handle.tiles(messages):
            out.append(r['data'] if r['type'] == 'tile' else ExecutorError(r['message']))
        return out

    def dispose(self):
        self.handle.release()


class GpuExecutor(LocalExecutor):
    """The GPU as one executor with many tiles in flight: a batch is one wavefront run."""

    def __init__(self, ctx=None, concurrency=256):
        super().__init__(ctx, concurrency)


class WorkerExecutor:
    """executors.ts:54-94 over a WorkerLike."""

    def __init__(self, worker, concurrency=2):
        self.worker = worker
        self.textures = set()
        self.concurrency = concurrency
        self.pending = {}
        self.failed = None
        worker.add_event_listener('message', self._on_message)
        worker.add_event_listener('error', self._on_error)

    def _on_message(self, event):
        response = event['data']
        job = self.pending.pop(response['job'], None)
        if job is None:
            return
        job['result'] = response['data'] if response['type'] == 'tile' else ExecutorError(response['message'])

    def _on_error(self, event):
        self.failed = ExecutorError(f"ray tracing worker failed: {event.get('message') or 'unknown error'}")
        for job in self.pending.values():
            job['result'] = self.failed
        self.pending.clear()

    def send(self, message):
        self.worker.post_message(message)

    def render_batch(self, messages):
        if self.failed:
            return [self.failed for _ in messages]
        slots = []
        for m in messages:
            slot = {'result': None}
            self.pending[m['job']] = slot
            slots.append(slot)
            self.worker.post_message(m)
        self.worker.flush()
        return [s['result'] if s['result'] is not None else ExecutorError('no response') for s in slots]

    def render(self, message):
        r = self.render_batch([message])[0]
        if isinstance(r, Exception):
            raise r
        return r

    def dispose(self):
        for job in self.pending.values():
            job['result'] = ExecutorError('the renderer was disposed')
        self.pending.clear()
        self.worker.terminate()
