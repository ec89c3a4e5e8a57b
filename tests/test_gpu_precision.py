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
    yield gpu, rays, ref, sc.geometry
    gpu.release()


def test_probe_set_1m(probe):
    gpu, rays, ref, _ = probe
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


def _mt64(tris, k, rays):
    """The reference's FP64 Möller-Trumbore (bvh.ts:228-252) of triangle k[i] against ray i, without
    the range tests: (t, u, v), the barycentric weights of v1 and v2, as the reference computes them.
    Used to classify first-hit disagreements, not to decide hits."""
    t9 = tris.reshape(-1, 9)[k].astype(np.float64)
    o, d = rays[:, :3], rays[:, 3:]
    e1, e2 = t9[:, 3:6], t9[:, 6:9]
    p = np.cross(d, e2)
    det = np.einsum('ij,ij->i', e1, p)
    inv = 1.0 / np.where(det == 0, np.nan, det)
    s_ = o - t9[:, 0:3]
    u = np.einsum('ij,ij->i', s_, p) * inv
    q = np.cross(s_, e1)
    v = np.einsum('ij,ij->i', d, q) * inv
    t = np.einsum('ij,ij->i', e2, q) * inv
    return t, u, v


def _classify(geometry, rays, w_tri, o_tri, o_t, edge_tol=1e-4, tie_tol=1e-6):
    """Sorts the rays whose first-hit triangle differs between the wide form and the FP64 oracle:
    'tie' - both triangles lie at the same distance along the ray (within tie_tol relative), so
    the choice between them is the order of an exact tie; 'edge' - the ray passes within edge_tol
    (barycentric) of an edge of the triangle one side hit and the other did not, where FP32 and
    FP64 rounding of the inclusive edge tests (bvh.ts:241-246) differ; 'real' - anything else."""
    idx = np.flatnonzero(w_tri != o_tri)
    out = {'tie': 0, 'edge': 0, 'real': 0}
    if idx.size == 0:
        return out, []
    r = rays[idx].astype(np.float64)
    real = []
    for j, i in enumerate(idx):
        wk, ok = int(w_tri[i]), int(o_tri[i])
        rr = r[j:j + 1]
        near_edge = False
        tw = to = None
        if wk >= 0:
            t, u, v = _mt64(geometry.triangles, np.array([wk]), rr)
            tw = t[0]
            near_edge |= min(u[0], v[0], 1 - u[0] - v[0]) > -edge_tol and min(u[0], v[0], 1 - u[0] - v[0]) < edge_tol
        if ok >= 0:
            t, u, v = _mt64(geometry.triangles, np.array([ok]), rr)
            to = t[0]
            near_edge |= min(u[0], v[0], 1 - u[0] - v[0]) < edge_tol
        if tw is not None and to is not None and abs(tw - to) <= tie_tol * abs(to):
            out['tie'] += 1
        elif near_edge:
            out['edge'] += 1
        else:
            out['real'] += 1
            real.append(int(i))
    return out, real


def test_probe_set_1m_wide(probe):
    """The FP32 wide form against the FP64 oracle. The first hit is found in FP32 and its t, u, v
    refined with the reference's FP64 test on that triangle, so where the ids agree t is the
    oracle's rounded to FP32 (relative error <= 2^-24, inside the 1e-6 bound). Rays whose id
    differs are classified: ties between triangles at the same distance, and rays through an
    edge within FP32 rounding of it, are the only disagreements allowed."""
    gpu, rays, ref, geometry = probe
    info = gpu.build_wide()
    report = {'wide_bvh': info}
    for cull in (True, False):
        print(f'wide probe set, cull={cull}', flush=True)
        w_tri, w_t, w_u, w_v, w_f = gpu.intersect_closest_wide(rays, cull=cull)
        o_tri, o_t, o_u, o_v, o_f = ref[cull]
        mismatch = int(np.count_nonzero(w_tri != o_tri))
        both = (w_tri == o_tri) & (w_tri >= 0)
        rel = np.abs(w_t[both].astype(np.float64) - o_t[both]) / np.abs(o_t[both])
        classes, real = _classify(geometry, rays, w_tri, o_tri, o_t)
        report[f'cull_{cull}'] = {
            'rays': len(rays), 'hits': int(both.sum()), 'id_mismatch_rate': mismatch / len(rays),
            'id_mismatches': classes, 'max_rel_t_error': float(rel.max()),
            'front_mismatch': int(np.count_nonzero(w_f[both] != o_f[both]))}
        assert classes['real'] == 0, real[:10]
        assert rel.max() <= 1e-6
        assert mismatch / len(rays) <= 1e-4
    print('PRECISION_WIDE ' + json.dumps(report), flush=True)
