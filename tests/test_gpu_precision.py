"""Precision form: a seeded probe set of 1M rays (500k primary rays through uniform random
sub-pixel positions of the 1.38M-triangle scene, 500k secondary rays in cosine directions from
their hits), stored as float32 so both sides trace identical rays, against the oracle's FP64
intersectClosest (bvh.ts:202-282). The GPU path computes the reference's FP64 arithmetic, so the
first-hit triangle ids must agree exactly and t must be equal (relative error 0, within the
1e-6 bound the precision form allows)."""
import json
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from meercadraytracer_amd import native, scenes, workloads
from meercadraytracer_amd.compile import compile_camera, compile_scene

pytestmark = pytest.mark.gpu


def _oracle_closest(geometry, rays, cull):
    chunks = np.array_split(rays, max(1, min(16, os.cpu_count() or 1)))
    with ThreadPoolExecutor(len(chunks)) as ex:
        parts = list(ex.map(lambda c: oracle.intersect_closest(geometry, c, cull=cull)[:5], chunks))
    return [np.concatenate([p[i] for p in parts]) for i in range(5)]


@pytest.fixture(scope='module')
def probe():
    """The 1M probe rays of the 1.38M-triangle scene, and the FP64 oracle's hits for cull on/off."""
    ctx = native.default_context()
    scene, camera, _ = scenes.plain_gpu_scene(0)
    sc = compile_scene(scene, ctx.build_bvh)
    cam = compile_camera(camera)
    gpu = native.Scene(ctx)
    gpu.set_geometry(sc.geometry, 1)
    rays = workloads.probe_rays(sc.geometry, cam, 1920, 1080, 1_000_000, 1234,
                                lambda r: gpu.intersect_closest(r, cull=True))
    assert rays.shape == (1_000_000, 6) and rays.dtype == np.float32
    ref = {}
    for cull in (True, False):
        print(f'oracle probe hits, cull={cull}', flush=True)
        ref[cull] = _oracle_closest(sc.geometry, rays.astype(np.float64), cull)
    yield gpu, rays, ref
    gpu.release()


def test_probe_set_1m(probe):
    gpu, rays, ref = probe
    r64 = rays.astype(np.float64)
    report = {}
    for cull in (True, False):
        print(f'probe set, cull={cull}', flush=True)
        g_tri, g_t, g_u, g_v, g_f = gpu.intersect_closest(r64, cull=cull)
        o_tri, o_t, o_u, o_v, o_f = ref[cull]
        mismatch = int(np.count_nonzero(g_tri != o_tri))
        both = (g_tri == o_tri) & (g_tri >= 0)
        rel = np.abs(g_t[both] - o_t[both]) / np.maximum(np.abs(o_t[both]), 1e-300)
        report[f'cull_{cull}'] = {'rays': len(rays), 'hits': int(both.sum()), 'id_mismatch_rate': mismatch / len(rays),
                                  'max_rel_t_error': float(rel.max()) if rel.size else 0.0}
        assert mismatch == 0
        assert np.array_equal(g_t.view(np.uint64), o_t.view(np.uint64))
        assert np.array_equal(g_u.view(np.uint64), o_u.view(np.uint64))
        assert np.array_equal(g_v.view(np.uint64), o_v.view(np.uint64))
        assert np.array_equal(g_f, o_f)
    print('PRECISION ' + json.dumps(report), flush=True)


def test_probe_set_1m_wide(probe):
    """The FP32 wide form against the FP64 oracle: the rate of rays whose first-hit triangle
    differs, and the relative t error where it agrees (largest, and 99.99th percentile). FP32
    rounding of the vertices relative to the ray origin moves t by a few parts in 10^7; where two
    triangles meet it can choose the neighbour."""
    gpu, rays, ref = probe
    info = gpu.build_wide()
    report = {'wide_bvh': info}
    for cull in (True, False):
        print(f'wide probe set, cull={cull}', flush=True)
        w_tri, w_t, w_u, w_v, w_f = gpu.intersect_closest_wide(rays, cull=cull)
        o_tri, o_t, o_u, o_v, o_f = ref[cull]
        mismatch = int(np.count_nonzero(w_tri != o_tri))
        both = (w_tri == o_tri) & (w_tri >= 0)
        rel = np.abs(w_t[both].astype(np.float64) - o_t[both]) / np.abs(o_t[both])
        report[f'cull_{cull}'] = {
            'rays': len(rays), 'hits': int(both.sum()), 'id_mismatch_rate': mismatch / len(rays),
            'max_rel_t_error': float(rel.max()), 'p9999_rel_t_error': float(np.percentile(rel, 99.99)),
            'frac_rel_t_within_1e-6': float((rel <= 1e-6).mean()),
            'front_mismatch': int(np.count_nonzero(w_f[both] != o_f[both]))}
        assert mismatch / len(rays) <= 1e-3
        assert np.percentile(rel, 99.99) <= 1e-5
    print('PRECISION_WIDE ' + json.dumps(report), flush=True)
