"""The drop-in path end to end on the GPU: the reference's protocol messages (geometry, textures,
shading, tile; src/protocol.ts:29-45) posted to the WorkerLike of meercadraytracer_amd/napi
(gpu_worker.js over the Node-API addon over libmcr.so), in Node, exactly as RaytracingRenderer's
workerFactory would receive them (src/RaytracingRenderer.ts:117-125). Every tile response must be
the oracle's renderTile float for float; a tile naming a version the worker does not hold gets
the protocol's error text (protocol.ts:77-79); a new shading version applies to later tiles."""
import dataclasses
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from meercadraytracer_amd import build, native, scenes
from meercadraytracer_amd.compile import compile_camera, compile_scene
from meercadraytracer_amd.raytrace import trace_settings
from meercadraytracer_amd.types import CompiledScene, TextureRef, Tile
from scenes_common import textured_scene

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TYPES = {np.float32: 'Float32Array', np.int32: 'Int32Array', np.uint32: 'Uint32Array', np.uint8: 'Uint8Array'}


class _Spec:
    def __init__(self, d):
        self.dir, self.n = d, 0

    def arr(self, a):
        a = np.ascontiguousarray(a)
        name = f'a{self.n}.bin'
        self.n += 1
        a.tofile(os.path.join(self.dir, name))
        return {'__array__': {'file': name, 'type': TYPES[a.dtype.type]}}

    def geometry(self, g):
        out = {k: self.arr(getattr(g, k)) for k in ('triangles', 'normals', 'uvs', 'tangents', 'colors',
                                                     'materials', 'flags', 'sides')}
        out.update(triangleCount=int(g.triangleCount), epsilon=float(g.epsilon),
                   bvh={'bounds': self.arr(g.bvh.bounds), 'nodes': self.arr(g.bvh.nodes), 'nodeCount': int(g.bvh.nodeCount)})
        return out

    def shading(self, sh):
        d = {'materials': [dataclasses.asdict(m) for m in sh.materials],
             'lights': [dataclasses.asdict(l) for l in sh.lights],
             'textures': [dataclasses.asdict(TextureRef.of(t, i)) for i, t in enumerate(sh.textures)],
             'environments': [dataclasses.asdict(e) for e in sh.environments],
             'background': sh.background, 'environment': sh.environment,
             'fog': dataclasses.asdict(sh.fog) if sh.fog is not None else None}
        return json.loads(json.dumps(d, default=float))

    def textures(self, sh):
        return [{'id': i, 'data': self.arr(t.data)} for i, t in enumerate(sh.textures)]


def _tile_msg(job, gv, sv, cam, st, tile, w, h, p, block):
    return {'type': 'tile', 'job': job, 'geometryVersion': gv, 'shadingVersion': sv,
            'camera': dataclasses.asdict(cam), 'settings': dataclasses.asdict(st),
            'tile': dataclasses.asdict(tile), 'width': w, 'height': h, 'pass': p, 'block': block}


@pytest.fixture(scope='module')
def addon():
    if not shutil.which('node'):
        pytest.skip('node is not installed')
    if not os.path.exists(build.NAPI_LIB):
        pytest.skip('mcr.node is not built')
    return build.NAPI_LIB


@pytest.mark.parametrize('kind', ['plain', 'textured'])
def test_protocol_through_the_node_addon(addon, tmp_path, kind):
    if kind == 'plain':
        sc, cam, o = scenes.compiled('plain_cpu', 0, build_bvh=oracle.build_bvh, width=48, height=32)
        W, H = 48, 32
    else:
        s3, cam3 = textured_scene('screen', seed=3)
        sc, cam, o = compile_scene(s3, oracle.build_bvh), compile_camera(cam3), {'globalIllumination': True}
        W, H = 48, 32
    st = trace_settings(o)
    shading2 = dataclasses.replace(sc.shading, materials=[dataclasses.replace(m, color=(m.color[1], m.color[2], m.color[0]))
                                                          for m in sc.shading.materials])
    spec = _Spec(str(tmp_path))
    tiles = [(Tile(0, 0, W, H), 0, 1), (Tile(8, 4, 24, 20), 1, 1), (Tile(0, 0, W, H), 0, 4)]
    msgs = [{'type': 'geometry', 'version': 1, 'geometry': spec.geometry(sc.geometry)}]
    if sc.shading.textures:
        msgs.append({'type': 'textures', 'items': spec.textures(sc.shading)})
    msgs.append({'type': 'shading', 'version': 1, 'shading': spec.shading(sc.shading)})
    expect = {}
    job = 0
    for tile, p, block in tiles:
        msgs.append(_tile_msg(job, 1, 1, cam, st, tile, W, H, p, block))
        expect[job] = (sc, tile, p, block)
        job += 1
    msgs.append(_tile_msg(job, 2, 1, cam, st, Tile(0, 0, 8, 8), W, H, 0, 1))  # a version not received
    expect[job] = 'the tile refers to a scene this worker has not received'
    job += 1
    msgs.append({'type': 'shading', 'version': 2, 'shading': spec.shading(shading2)})
    for tile, p, block in tiles[:2]:
        msgs.append(_tile_msg(job, 1, 2, cam, st, tile, W, H, p, block))
        expect[job] = (CompiledScene(sc.geometry, shading2), tile, p, block)
        job += 1
    with open(os.path.join(tmp_path, 'spec.json'), 'w') as f:
        json.dump({'messages': msgs}, f)
    print(f'node driver: {len(msgs)} messages, {kind}', flush=True)
    r = subprocess.run(['node', os.path.join(HERE, 'napi_driver.js'), str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    results = json.load(open(os.path.join(tmp_path, 'results.json')))
    for j, e in expect.items():
        got = results[str(j)]
        if isinstance(e, str):
            assert got == 'error:' + e
            continue
        assert got == 'tile', got
        scene, tile, p, block = e
        data = np.fromfile(os.path.join(tmp_path, f'tile_{j}.bin'), np.float32)
        ref, _ = oracle.render_tile(scene, cam, st, tile, W, H, p, block)
        assert np.array_equal(data.view(np.uint32), ref.view(np.uint32)), (kind, j)
