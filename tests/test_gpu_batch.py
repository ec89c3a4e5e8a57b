"""Batched form: tile jobs of several scenes, cameras and settings in one wavefront run
(mcr_render_batch). Each job's tile must be bit-identical to the oracle's renderTile of that
job alone, whatever the jobs it shares waves with; replayed (graph) and directly launched
iterations must agree."""
import numpy as np
import pytest

import oracle
from meercadraytracer_amd import abi, native, scenes
from meercadraytracer_amd.raytrace import trace_settings
from meercadraytracer_amd.types import Tile

pytestmark = pytest.mark.gpu


def _resident(sc):
    s = native.Scene()
    s.set_geometry(sc.geometry, 1)
    s.set_shading(sc.shading, 1)
    return s


@pytest.fixture(scope='module')
def views():
    a = scenes.compiled('plain_cpu', 0, build_bvh=oracle.build_bvh, width=64, height=48)
    b = scenes.compiled('plain_gpu', 2, build_bvh=oracle.build_bvh, width=40, height=40, quads=30,
                        spheres=12, subdivisions=2)
    c = scenes.compiled('plain_cpu', 5, build_bvh=oracle.build_bvh, width=64, height=48)
    out = []
    for (sc, cam, o), extra, (w, h) in ((a, {}, (64, 48)), (b, {'seed': 3}, (40, 40)),
                                        (c, {'shadows': 'castShadow', 'toneMapping': 1}, (64, 48)),
                                        (a, {'seed': 9, 'globalIllumination': True}, (64, 48))):
        o = dict(o)
        o.update(extra)
        out.append((sc, cam, trace_settings(o), w, h))
    return out


def _jobs(views, per_view=4):
    jobs, jv, meta, off = [], [], [], 0
    rng = np.random.default_rng(11)
    for k in range(per_view):
        for v, (_, _, _, w, h) in enumerate(views):
            tw, th = int(rng.integers(5, w + 1)), int(rng.integers(5, h + 1))
            x, y = int(rng.integers(0, w - tw + 1)), int(rng.integers(0, h - th + 1))
            t = Tile(x, y, tw, th)
            block = 1 if k % 3 else 4
            jobs.append(abi.TileJob(t.to_c(), w, h, k, block, off))
            jv.append(v)
            meta.append((v, t, k, block))
            off += tw * th * 4
    return jobs, jv, meta


@pytest.mark.parametrize('capacity,graph', [(1 << 21, True), (1 << 21, False), (256, True), (256, False)])
def test_mixed_views_bit_exact(views, capacity, graph):
    ctx = native.default_context()
    res = [_resident(sc) for sc, *_ in views]
    vlist = [(res[i], cam, st) for i, (sc, cam, st, w, h) in enumerate(views)]
    jobs, jv, meta = _jobs(views)
    print(f'capacity {capacity} graph {graph} jobs {len(jobs)}', flush=True)
    ctx.set_capacity(capacity)
    ctx.set_graph(graph)
    try:
        out = native.render_batch(ctx, vlist, jobs, jv)
    finally:
        ctx.set_capacity()
        ctx.set_graph(True)
    for j, (v, t, p, block) in zip(jobs, meta):
        sc, cam, st, w, h = views[v]
        ref, _ = oracle.render_tile(sc, cam, st, t, w, h, p, block)
        got = out[j.out_offset:j.out_offset + t.width * t.height * 4]
        bad = np.flatnonzero(got.view(np.uint32) != ref.view(np.uint32))
        assert bad.size == 0, f'view {v} tile {t} pass {p}: {bad.size} floats differ'
    for r in res:
        r.release()


def test_batch_rejects_bad_view_index(views):
    ctx = native.default_context()
    sc, cam, st, w, h = views[0]
    r = _resident(sc)
    jobs = [abi.TileJob(Tile(0, 0, 8, 8).to_c(), w, h, 0, 1, 0)]
    with pytest.raises(native.NativeError) as e:
        native.render_batch(ctx, [(r, cam, st)], jobs, [1])
    assert 'names no view' in str(e.value)
    r.release()


def test_graph_replay_counts_equal_direct(views):
    """Counters (queries, node visits, triangle tests) do not depend on how iterations launch."""
    ctx = native.default_context()
    sc, cam, st, w, h = views[1]
    r = _resident(sc)
    jobs = [abi.TileJob(Tile(0, 0, w, h).to_c(), w, h, p, 1, p * w * h * 4) for p in range(6)]
    got = []
    for graph in (True, False, True):
        ctx.set_graph(graph)
        cnt = abi.Counters()
        out = r.render_jobs(cam, st, jobs, counters=cnt)
        got.append((out.copy(), cnt.closest_queries, cnt.any_queries, cnt.node_visits, cnt.triangle_tests))
    ctx.set_graph(True)
    for g in got[1:]:
        assert np.array_equal(g[0].view(np.uint32), got[0][0].view(np.uint32))
        assert g[1:] == got[0][1:]
    r.release()


def test_call_split_into_runs(views, monkeypatch):
    """A call with more samples than one run holds is traced as consecutive runs; results are the
    same as one run."""
    ctx = native.default_context()
    res = [_resident(sc) for sc, *_ in views]
    vlist = [(res[i], cam, st) for i, (sc, cam, st, w, h) in enumerate(views)]
    jobs, jv, meta = _jobs(views)
    whole = native.render_batch(ctx, vlist, jobs, jv)
    monkeypatch.setenv('MCR_RUN_SAMPLES', '4000')
    split = native.render_batch(ctx, vlist, jobs, jv)
    monkeypatch.delenv('MCR_RUN_SAMPLES')
    assert np.array_equal(whole.view(np.uint32), split.view(np.uint32))
    for r in res:
        r.release()
