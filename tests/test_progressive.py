"""tests/progressive.test.ts restated: tiles, the progressive scheduler's results and its
failure handling. Scheduling logic runs on the CPU with oracle-backed executors and image;
the GPU executor path runs under the gpu marker."""
import math

import numpy as np
import pytest

import oracle
from helpers import floor, overhead_light, top_camera
from meercadraytracer_amd import three as T
from meercadraytracer_amd.compile import compile_camera, compile_scene
from meercadraytracer_amd.progressive import ProgressiveRenderer, make_tiles
from meercadraytracer_amd.raytrace import trace_settings
from progressive_support import FakeWorker, OracleExecutor, OracleHandler, OracleImage


def make_scene():  # progressive.test.ts:38-47
    s = T.Scene()
    texture = T.DataTexture(np.array([255, 255, 255, 255, 128, 128, 128, 255], np.uint8), 2, 1)
    s.add(floor(T.MeshStandardMaterial(map=texture, roughness=0.6)), overhead_light(2), T.AmbientLight(0xffffff, 1))
    box = T.Mesh(T.BoxGeometry(0.5, 0.5, 0.5), T.MeshPhongMaterial(color=0xff8800))
    box.position = [0.0, 0.0, 0.5]
    s.add(box)
    return s


def options(**kw):  # progressive.test.ts:49-54
    o = dict(width=20, height=12, settings=trace_settings({}), outputColorSpace='srgb', samples=3,
             previewBlock=1, tileSize=8, frameInterval=0)
    o.update(kw)
    return o


def oracle_raytrace(scene, camera, w, h, samples):
    sc = compile_scene(scene, oracle.build_bvh)
    data, _ = oracle.raytrace(sc, compile_camera(camera), trace_settings({}), w, h, samples, True)
    return data


def cpu_renderer(executors):
    return ProgressiveRenderer(executors, build_bvh=oracle.build_bvh, image_factory=OracleImage,
                               local_factory=lambda ctx: OracleExecutor())


def test_make_tiles_cover_centre_first():  # progressive.test.ts:56-63
    tiles = make_tiles(20, 12, 8)
    assert sum(t.width * t.height for t in tiles) == 240
    first = tiles[0]
    assert first.x <= 10 and first.x + first.width >= 10
    # the same order as the oracle's restatement, and as 1080p at 64 px
    assert [(t.x, t.y, t.width, t.height) for t in make_tiles(1920, 1080, 64)] == oracle.make_tiles(1920, 1080, 64)
    assert [(t.x, t.y, t.width, t.height) for t in make_tiles(37, 23, 5)] == oracle.make_tiles(37, 23, 5)


def test_progressive_equals_raytrace_in_page():  # progressive.test.ts:66-76
    scene, camera = make_scene(), top_camera()
    frames = []
    r = cpu_renderer([OracleExecutor()])
    done = r.render(scene, camera, options(), lambda f: frames.append(f))
    assert done is True
    assert frames[-1]['done'] and frames[-1]['samples'] == 3
    assert np.array_equal(frames[-1]['image'], oracle_raytrace(scene, camera, 20, 12, 3))


def test_progressive_through_workers_sends_scene_once():  # progressive.test.ts:78-108
    from meercadraytracer_amd.executors import WorkerExecutor
    scene, camera = make_scene(), top_camera()
    workers = [FakeWorker(OracleHandler()), FakeWorker(OracleHandler())]
    r = cpu_renderer([WorkerExecutor(w) for w in workers])
    image = {}
    assert r.render(scene, camera, options(), lambda f: image.update(img=f['image'])) is True
    assert np.array_equal(image['img'], oracle_raytrace(scene, camera, 20, 12, 3))
    count = lambda kind: sum(1 for w in workers for m in w.received if m['type'] == kind)
    assert count('geometry') == 2 and count('textures') == 2 and count('shading') == 2
    camera.position = [0.3, 0.0, 10.0]
    camera.updateMatrixWorld()
    r.render(scene, camera, options(), lambda f: None)
    assert count('geometry') == 2 and count('shading') == 2       # a camera move sends tiles only
    scene.children[0].material.color = T.Color(0x00ff00)
    r.render(scene, camera, options(), lambda f: None)
    assert count('geometry') == 2 and count('textures') == 2 and count('shading') == 4  # a colour change resends shading
    scene.children[3].position = [0.2, 0.0, 0.5]
    r.render(scene, camera, options(), lambda f: None)
    assert count('geometry') == 4                                  # moving a mesh resends geometry
    r.dispose()
    assert all(w.terminated for w in workers)


def test_block_preview_first():  # progressive.test.ts:112-118
    frames = []
    r = cpu_renderer([OracleExecutor()])
    r.render(make_scene(), top_camera(), options(previewBlock=4), lambda f: frames.append(f['samples']))
    assert frames[0] == 0 and frames[-1] == 3


def test_superseded_and_stopped_render_resolve_false():  # progressive.test.ts:120-129
    r = cpu_renderer([OracleExecutor()])
    calls = []

    def stop_on_first(f):
        calls.append(f)
        if len(calls) == 1:
            r.stop()
    assert r.render(make_scene(), top_camera(), options(), stop_on_first) is False
    seen = []

    def supersede(f):
        seen.append(f)
        if len(seen) == 1:
            assert r.render(make_scene(), top_camera(), options(samples=1), lambda g: None) is True
    assert r.render(make_scene(), top_camera(), options(), supersede) is False


def test_worker_failure_is_tolerated():  # progressive.test.ts:131-141
    from meercadraytracer_amd.executors import WorkerExecutor
    scene, camera = make_scene(), top_camera()
    good = FakeWorker(OracleHandler())
    bad = FakeWorker(OracleHandler(), fail=True)
    r = cpu_renderer([WorkerExecutor(bad), WorkerExecutor(good)])
    image = {}
    assert r.render(scene, camera, options(), lambda f: image.update(img=f['image'])) is True
    assert np.array_equal(image['img'], oracle_raytrace(scene, camera, 20, 12, 3))


def test_all_workers_failing_fall_back_in_process():  # progressive.test.ts:143-160
    from meercadraytracer_amd.executors import WorkerExecutor
    scene, camera = make_scene(), top_camera()
    r = cpu_renderer([WorkerExecutor(FakeWorker(OracleHandler(), fail=True))])
    image = {}
    with pytest.warns(UserWarning, match='tracing in-process'):
        assert r.render(scene, camera, options(), lambda f: image.update(img=f['image'])) is True
    assert np.array_equal(image['img'], oracle_raytrace(scene, camera, 20, 12, 3))


def test_executor_failure_mid_render_requeues_tiles():
    scene, camera = make_scene(), top_camera()
    flaky = OracleExecutor(concurrency=4, fail_after=5)
    r = cpu_renderer([flaky, OracleExecutor(concurrency=2)])
    image = {}
    assert r.render(scene, camera, options(), lambda f: image.update(img=f['image'])) is True
    assert np.array_equal(image['img'], oracle_raytrace(scene, camera, 20, 12, 3))


@pytest.mark.gpu
@pytest.mark.parametrize('concurrency', [1, 7, 256])
def test_gpu_executor_progressive_equals_raytrace(concurrency):
    from meercadraytracer_amd.executors import GpuExecutor
    from meercadraytracer_amd.raytrace import raytrace
    scene, camera = make_scene(), top_camera()
    frames = []
    r = ProgressiveRenderer([GpuExecutor(concurrency=concurrency)])
    assert r.render(scene, camera, options(previewBlock=4), lambda f: frames.append(f)) is True
    assert frames[0]['samples'] == 0
    want = raytrace(scene, camera, 20, 12, {'samples': 3})['data']
    assert np.array_equal(frames[-1]['image'], want)
    assert np.array_equal(want, oracle_raytrace(scene, camera, 20, 12, 3))
    r.dispose()


@pytest.mark.gpu
def test_gpu_workers_and_protocol_errors():
    from meercadraytracer_amd.executors import WorkerExecutor
    from meercadraytracer_amd.protocol import NO_SCENE, create_tile_handler
    scene, camera = make_scene(), top_camera()
    workers = [FakeWorker(create_tile_handler()), FakeWorker(create_tile_handler())]
    r = ProgressiveRenderer([WorkerExecutor(w, concurrency=3) for w in workers])
    image = {}
    assert r.render(scene, camera, options(), lambda f: image.update(img=f['image'])) is True
    assert np.array_equal(image['img'], oracle_raytrace(scene, camera, 20, 12, 3))
    h = create_tile_handler()
    msg = {'type': 'tile', 'job': 9, 'geometryVersion': 1, 'shadingVersion': 1, 'camera': compile_camera(camera),
           'settings': trace_settings({}), 'tile': make_tiles(20, 12, 8)[0], 'width': 20, 'height': 12, 'pass': 0,
           'block': 1}
    assert h(msg) == {'type': 'error', 'job': 9, 'message': NO_SCENE}
    # protocol.ts:84: a shading whose texture pixels never arrived fails its tiles, in those words
    from meercadraytracer_amd.progressive import SceneSync
    sync = SceneSync(oracle.build_bvh)
    sync.update(scene, camera)
    h(dict(type='geometry', version=1, geometry=sync.geometry))
    h(dict(type='shading', version=1, shading=sync.shading))
    tid = sync.shading.textures[0].id
    assert h(msg) == {'type': 'error', 'job': 9, 'message': f'texture {tid} was not sent'}
    h(dict(type='textures', items=[{'id': tid, 'data': sync._textures[0].data}]))
    assert h(msg)['type'] == 'tile'
    h.release()
    r.dispose()
