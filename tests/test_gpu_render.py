"""GPU tile rendering against the portable-math oracle: the same per-pixel seeds and the same
FP64 arithmetic, so every float of every tile is identical."""
import numpy as np
import pytest

import oracle
from meercadraytracer_amd import abi, scenes
from meercadraytracer_amd.raytrace import trace_settings
from meercadraytracer_amd.types import Tile

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def c0():
    return scenes.compiled('plain_cpu', 0, build_bvh=oracle.build_bvh, width=96, height=64)


def _gpu_scene(sc):
    from meercadraytracer_amd import native
    s = native.Scene()
    s.set_geometry(sc.geometry, 1)
    s.set_shading(sc.shading, 1)
    return s


@pytest.mark.parametrize('opts', [
    {},
    {'globalIllumination': True, 'maxBounces': 6},
    {'shadows': 'castShadow'},
    {'shadows': 'none', 'toneMapping': abi.ReinhardToneMapping},
    {'toneMapping': abi.AgXToneMapping, 'seed': 7},
    {'toneMapping': abi.NeutralToneMapping, 'clearAlpha': 1, 'clearColor': (0, 0, 1)},
])
def test_tile_bit_exact_plain_cpu_scene(c0, opts):
    sc, cam, base = c0
    o = dict(base)
    o.update(opts)
    st = trace_settings(o)
    s = _gpu_scene(sc)
    for tile, p, block in [(Tile(0, 0, 96, 64), 0, 1), (Tile(16, 8, 40, 33), 3, 1), (Tile(0, 0, 96, 64), 0, 4),
                           (Tile(90, 60, 6, 4), 1, 4)]:
        print(f'tile {tile} pass {p} block {block} opts {opts}', flush=True)
        g = s.render_tile(cam, st, tile, 96, 64, p, block)
        ref, _ = oracle.render_tile(sc, cam, st, tile, 96, 64, p, block)
        bad = np.flatnonzero(g.view(np.uint32) != ref.view(np.uint32))
        assert bad.size == 0, f'{bad.size} floats differ, first at {bad[:5]}: {g[bad[:5]]} vs {ref[bad[:5]]}'
    s.release()


def test_glass_scene_bit_exact():
    sc, cam, o = scenes.compiled('plain_gpu', 1, build_bvh=oracle.build_bvh, width=64, height=36, quads=40,
                                 spheres=24, subdivisions=2)
    st = trace_settings(o)
    s = _gpu_scene(sc)
    for p in (0, 5):
        g = s.render_tile(cam, st, Tile(0, 0, 64, 36), 64, 36, p, 1)
        ref, _ = oracle.render_tile(sc, cam, st, Tile(0, 0, 64, 36), 64, 36, p, 1)
        assert np.array_equal(g.view(np.uint32), ref.view(np.uint32))
    s.release()


def test_batched_jobs_equal_single_tiles(c0):
    sc, cam, base = c0
    st = trace_settings(base)
    s = _gpu_scene(sc)
    tiles = [Tile(0, 0, 32, 32), Tile(32, 0, 32, 32), Tile(64, 32, 32, 32), Tile(10, 10, 7, 9)]
    jobs, off = [], 0
    for i, t in enumerate(tiles):
        jobs.append(abi.TileJob(t.to_c(), 96, 64, i, 1 + (i % 2) * 3, off))
        off += t.width * t.height * 4
    out = s.render_jobs(cam, st, jobs)
    for j, t in zip(jobs, tiles):
        one = s.render_tile(cam, st, t, 96, 64, j.pass_, j.block)
        assert np.array_equal(out[j.out_offset:j.out_offset + t.width * t.height * 4], one)
    s.release()


def test_slot_refill_small_capacity(c0):
    """A 256-slot wavefront streams 18 jobs (6,912 samples) through refilled slots: every tile is
    still bit-identical to the oracle's."""
    from meercadraytracer_amd import native
    sc, cam, base = c0
    st = trace_settings(dict(base, globalIllumination=True, maxBounces=5))
    ctx = native.default_context()
    s = _gpu_scene(sc)
    jobs, tiles, off = [], [], 0
    for p in range(6):
        for t in (Tile(0, 0, 48, 32), Tile(48, 0, 48, 32), Tile(20, 32, 40, 32)):
            jobs.append(abi.TileJob(t.to_c(), 96, 64, p, 1 if p % 3 else 2, off))
            tiles.append(t)
            off += t.width * t.height * 4
    ctx.set_capacity(256)
    try:
        out = s.render_jobs(cam, st, jobs)
    finally:
        ctx.set_capacity()
    for j, t in zip(jobs, tiles):
        ref, _ = oracle.render_tile(sc, cam, st, t, 96, 64, j.pass_, j.block)
        got = out[j.out_offset:j.out_offset + t.width * t.height * 4]
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    s.release()


def test_default_capacity_full_frame_matches_smaller_capacity():
    """Five 1920x1080 passes (10.4M samples) fill the default 2^23 path slots; the frames are
    bit-identical to a run on 2^21 slots (the capacity changes memory use, not results), and a
    16x16 block of the last pass matches the oracle."""
    from meercadraytracer_amd import native
    W, H = 1920, 1080
    sc, cam, base = scenes.compiled('plain_cpu', 0, build_bvh=oracle.build_bvh, width=W, height=H)
    st = trace_settings(dict(base, globalIllumination=True, maxBounces=4))
    ctx = native.default_context()
    s = _gpu_scene(sc)
    npx = W * H
    jobs = [abi.TileJob(Tile(0, 0, W, H).to_c(), W, H, p, 1, p * npx * 4) for p in range(5)]
    print(f'{len(jobs)} passes of {W}x{H}, default slots', flush=True)
    ctx.set_capacity()
    big = s.render_jobs(cam, st, jobs)
    ctx.set_capacity(1 << 21)
    try:
        small = s.render_jobs(cam, st, jobs)
    finally:
        ctx.set_capacity()
    assert np.array_equal(big.view(np.uint32), small.view(np.uint32))
    t = Tile(952, 532, 16, 16)
    ref, _ = oracle.render_tile(sc, cam, st, t, W, H, 4, 1)
    frame = big[4 * npx * 4:5 * npx * 4].reshape(H, W, 4)
    got = np.ascontiguousarray(frame[t.y:t.y + t.height, t.x:t.x + t.width]).reshape(-1)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    s.release()


def test_raytrace_bytes_equal_oracle(c0):
    from meercadraytracer_amd.raytrace import raytrace_compiled
    sc, cam, base = c0
    o = dict(base, samples=3)
    res = raytrace_compiled(sc, cam, 96, 64, o)
    ref, _ = oracle.raytrace(sc, cam, trace_settings(o), 96, 64, samples=3, srgb=True)
    assert np.array_equal(res['data'], ref)
