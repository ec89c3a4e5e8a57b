"""Textures and environments on the GPU product path against the oracle, every float of every
tile identical: colour, normal, roughness, metalness, ao, emissive and alpha maps, textured
see-through shadows, alpha-map cut-outs in camera and shadow rays, envMap combines, the scene
environment (reflections, misses after a bounce, the environment-diffuse ray) and the screen,
equirect, cube and colour backgrounds (tracer.ts:119-406, 544-594, 659-691; textures.ts:229-317)."""
import numpy as np
import pytest

import oracle
from meercadraytracer_amd import abi, native, scenes
from meercadraytracer_amd.compile import compile_camera, compile_scene
from meercadraytracer_amd.raytrace import trace_settings
from meercadraytracer_amd.types import Tile
from scenes_common import textured_scene

pytestmark = pytest.mark.gpu

W, H = 96, 64


def _compiled(background, seed=0):
    s, cam = textured_scene(background, seed)
    return compile_scene(s, oracle.build_bvh), compile_camera(cam)


def _resident(sc):
    s = native.Scene()
    s.set_geometry(sc.geometry, 1)
    s.set_shading(sc.shading, 1)
    return s


def _same(got, ref, label):
    bad = np.flatnonzero(got.view(np.uint32) != ref.view(np.uint32))
    assert bad.size == 0, f'{label}: {bad.size} floats differ, first at {bad[:5]}: {got[bad[:5]]} vs {ref[bad[:5]]}'


@pytest.mark.parametrize('background', ['environment', 'cube', 'screen', 'color'])
@pytest.mark.parametrize('opts', [
    {},
    {'globalIllumination': True, 'maxBounces': 6},
    {'shadows': 'castShadow', 'seed': 5},
    {'shadows': 'none', 'toneMapping': abi.ACESFilmicToneMapping},
])
def test_textured_tiles_bit_exact(background, opts):
    sc, cam = _compiled(background)
    st = trace_settings(opts)
    s = _resident(sc)
    try:
        for tile, p, block in [(Tile(0, 0, W, H), 0, 1), (Tile(0, 0, W, H), 3, 1), (Tile(10, 6, 50, 41), 1, 4)]:
            print(f'{background} {opts} tile {tile} pass {p} block {block}', flush=True)
            g = s.render_tile(cam, st, tile, W, H, p, block)
            ref, _ = oracle.render_tile(sc, cam, st, tile, W, H, p, block)
            _same(g, ref, f'{background} {opts} pass {p}')
    finally:
        s.release()


@pytest.mark.parametrize('capacity,graph', [(1 << 21, True), (256, True), (1 << 21, False)])
def test_textured_and_plain_views_in_one_batch(capacity, graph):
    """A textured view and an untextured one share waves: shadow kernels with texture lookups run
    over both, replayed from a graph or launched directly, from a full or a 256-slot wavefront."""
    tex_sc, tex_cam = _compiled('screen', seed=2)
    plain_sc, plain_cam, o = scenes.compiled('plain_cpu', 0, build_bvh=oracle.build_bvh, width=W, height=H)
    views = [(tex_sc, tex_cam, trace_settings({'seed': 4})), (plain_sc, plain_cam, trace_settings(o)),
             (tex_sc, tex_cam, trace_settings({'globalIllumination': True, 'maxBounces': 3}))]
    res = [_resident(tex_sc), _resident(plain_sc)]
    vlist = [(res[0], views[0][1], views[0][2]), (res[1], views[1][1], views[1][2]), (res[0], views[2][1], views[2][2])]
    jobs, jv, meta, off = [], [], [], 0
    for k in range(3):
        for v in range(3):
            t = Tile(8 * k, 4 * v, W - 16 * k, H - 8 * v)
            block = 1 if (k + v) % 3 else 4
            jobs.append(abi.TileJob(t.to_c(), W, H, k, block, off))
            jv.append(v)
            meta.append((v, t, k, block))
            off += t.width * t.height * 4
    ctx = native.default_context()
    print(f'capacity {capacity} graph {graph}', flush=True)
    ctx.set_capacity(capacity)
    ctx.set_graph(graph)
    try:
        out = native.render_batch(ctx, vlist, jobs, jv)
    finally:
        ctx.set_capacity()
        ctx.set_graph(True)
        for r in res:
            r.release()
    for j, (v, t, p, block) in zip(jobs, meta):
        sc, cam, st = views[v]
        ref, _ = oracle.render_tile(sc, cam, st, t, W, H, p, block)
        _same(out[j.out_offset:j.out_offset + t.width * t.height * 4], ref, f'view {v} tile {t} pass {p}')


def test_textured_render_through_wide_bvh():
    """The precision form with textures: FP32 closest and shadow hits feed the same texture
    lookups; 128 passes of a crop agree with the FP64 render within 3 sigma per pixel channel
    (at most 1% outside), and the crop means within 1%."""
    sc, cam = _compiled('environment', seed=1)
    s = _resident(sc)
    s.build_wide()
    ctx = native.default_context()
    tile = Tile(20, 12, 48, 40)
    n = 128

    def passes(mode, seed):
        st = trace_settings({'seed': seed})
        jobs = [abi.TileJob(tile.to_c(), W, H, p, 1, p * tile.width * tile.height * 4) for p in range(n)]
        ctx.set_intersection(mode)
        try:
            return s.render_jobs(cam, st, jobs).reshape(n, -1, 4)[..., :3].astype(np.float64)
        finally:
            ctx.set_intersection(abi.MCR_ISECT_FP64)
    try:
        f = passes(abi.MCR_ISECT_FP64, 0)
        g = passes(abi.MCR_ISECT_WIDE_FP32, 1)
    finally:
        s.release()
    se = np.sqrt(f.var(0, ddof=1) / n + g.var(0, ddof=1) / n)
    diff = np.abs(f.mean(0) - g.mean(0))
    outside = ((diff > 3 * se) & (diff > 1e-12)).mean()
    rel = abs(g.mean() - f.mean()) / f.mean()
    print(f'textured wide: {outside:.4%} outside 3 sigma, mean rel {rel:.2e}', flush=True)
    assert outside <= 0.01 and rel <= 0.01


def test_textured_workload_bit_exact():
    """The textured benchmark workload's generator (repeated sRGB colour, normal and roughness maps
    on a heightfield, mapped spheres, an HDR sky as environment and background, GI on, 8 bounces)
    at a small size, every float identical to the oracle."""
    sc, cam, o = scenes.compiled('textured_gpu', 0, build_bvh=oracle.build_bvh, width=96, height=54, quads=60,
                                 spheres=16, subdivisions=2)
    st = trace_settings(o)
    s = _resident(sc)
    try:
        for tile, p in [(Tile(0, 0, 96, 54), 0), (Tile(20, 10, 40, 30), 2)]:
            print(f'textured_gpu tile {tile} pass {p}', flush=True)
            g = s.render_tile(cam, st, tile, 96, 54, p, 1)
            ref, _ = oracle.render_tile(sc, cam, st, tile, 96, 54, p, 1)
            _same(g, ref, f'textured_gpu pass {p}')
    finally:
        s.release()


def test_bad_texture_data_is_refused():
    """Indices out of range are refused when the shading is set, with a message naming the part;
    a scene with environments keeps one light slot for the environment-diffuse ray (at most 63
    lights)."""
    import dataclasses
    from meercadraytracer_amd.types import CompiledLight
    sc, cam = _compiled('environment')
    s = _resident(sc)
    try:
        bad_mat = dataclasses.replace(sc.shading.materials[0], map=len(sc.shading.textures))
        cases = [
            (dataclasses.replace(sc.shading, materials=[bad_mat] + sc.shading.materials[1:]), 'material 0'),
            (dataclasses.replace(sc.shading, environments=[dataclasses.replace(sc.shading.environments[0],
                                                                                 texture=99)]
                                 + sc.shading.environments[1:]), 'environment 0'),
            (dataclasses.replace(sc.shading, environment=len(sc.shading.environments)), 'scene environment'),
            (dataclasses.replace(sc.shading, lights=[CompiledLight('point', position=(0.0, 3.0, float(i)))
                                                     for i in range(64)]), 'more than 63 lights'),
        ]
        for shading, words in cases:
            with pytest.raises(native.NativeError) as e:
                s.set_shading(shading, 7)
            assert words in str(e.value), (words, str(e.value))
        # 63 lights with environments still trace
        ok = dataclasses.replace(sc.shading, lights=[CompiledLight('point', color=(0.2, 0.2, 0.2),
                                                                   position=(float(i % 5) - 2, 3.0, float(i // 5)))
                                                     for i in range(63)])
        s.set_shading(ok, 8)
        st = trace_settings({})
        g = s.render_tile(cam, st, Tile(0, 0, 24, 16), W, H, 0, 1)
        from meercadraytracer_amd.types import CompiledScene
        ref, _ = oracle.render_tile(CompiledScene(sc.geometry, ok), cam, st, Tile(0, 0, 24, 16), W, H, 0, 1)
        _same(g, ref, '63 lights with environments')
    finally:
        s.release()
