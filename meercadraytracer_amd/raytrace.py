"""Headless API: raytrace / raytraceCompiled / traceSettings (src/raytrace.ts:9-82), on the GPU.

raytrace_compiled keeps the image, the running FP32 sum and each pass's tile in HBM: every pass is
one mcr_render_tiles call over the full-image tile (raytrace.ts:89-95) followed by the device
writeTile (mcr_encode_tile), and only the final 8-bit image is copied back.
"""
import math

import numpy as np

from . import abi, native
from .compile import compile_camera, compile_scene
from .types import TraceSettings


WIDE_SUM_SAMPLES = 1024  # renders with more passes keep their running sum in FP64


def trace_settings(options=None):  # raytrace.ts:45-56
    o = options or {}
    return TraceSettings(
        maxBounces=max(0, math.floor(o.get('maxBounces', 4))),
        shadows=o.get('shadows', 'all'),
        globalIllumination=o.get('globalIllumination', False),
        toneMapping=o.get('toneMapping', abi.NoToneMapping),
        toneMappingExposure=o.get('toneMappingExposure', 1.0),
        clearColor=tuple(o.get('clearColor', (0.0, 0.0, 0.0))),
        clearAlpha=o.get('clearAlpha', 0.0),
        seed=o.get('seed', 0),
    )


class RaytraceResult(dict):
    """{width, height, data}: data is RGBA uint8, top row first, straight alpha."""

    @property
    def data(self):
        return self['data']


def raytrace(scene, camera, width, height, options=None, ctx=None):  # raytrace.ts:63-66
    compiled = compile_scene(scene)
    return raytrace_compiled(compiled, compile_camera(camera), width, height, options, ctx=ctx)


def raytrace_compiled(scene, camera, width, height, options=None, ctx=None, counters=None,
                      pass_batch_bytes=4 << 30):
    """raytrace.ts:69-82: `samples` passes of one full-image tile, each accumulated into the FP32
    sum and encoded as writeTile does. Passes are traced in batches (one device run streams all
    the batch's samples through the wavefront); each pass keeps its own output so the passes are
    still summed in order, as the reference sums them. Beyond WIDE_SUM_SAMPLES passes the running
    sum is kept in FP64 (mcr_encode_tile64)."""
    o = options or {}
    settings = trace_settings(o)
    samples = max(1, math.floor(o.get('samples', 16)))
    srgb = o.get('outputColorSpace', 'srgb') != 'srgb-linear'
    ctx = ctx or native.default_context()
    gpu = native.Scene(ctx)
    gpu.set_geometry(scene.geometry, 1)
    gpu.set_shading(scene.shading, 1)
    npx = width * height
    batch = max(1, min(samples, pass_batch_bytes // (npx * 16)))
    data = native.DeviceBuffer(ctx, batch * npx * 16)
    wide = samples > WIDE_SUM_SAMPLES
    acc = native.DeviceBuffer(ctx, npx * (32 if wide else 16))
    image = native.DeviceBuffer(ctx, npx * 4)
    acc.zero()
    image.zero()
    from .types import Tile
    t = Tile(0, 0, width, height)
    cnt = counters if counters is not None else abi.Counters()
    for p0 in range(0, samples, batch):
        passes = range(p0, min(samples, p0 + batch))
        render_passes(gpu, camera, settings, width, height, passes, data.ptr, cnt)
        for i, p in enumerate(passes):
            if wide:
                native.encode_tile64_device(ctx, image.ptr, width, t, data.ptr + i * npx * 16, acc.ptr, p + 1, srgb)
            else:
                native.encode_tile_device(ctx, image.ptr, width, t, data.ptr + i * npx * 16, acc.ptr, p + 1, srgb)
    out = image.download(np.uint8, npx * 4)
    for b in (data, acc, image):
        b.free()
    gpu.release()
    return RaytraceResult(width=width, height=height, data=out)


def render_passes(gpu, camera, settings, width, height, passes, out_ptr, counters=None):
    """Full-image passes in one device run; pass i of the list lands at out_ptr + i * frame."""
    tile = abi.Tile(0, 0, width, height)
    npx = width * height
    jobs = [abi.TileJob(tile, width, height, p, 1, i * npx * 4) for i, p in enumerate(passes)]
    gpu.render_jobs(camera, settings, jobs, out=out_ptr, out_device=True, counters=counters)
