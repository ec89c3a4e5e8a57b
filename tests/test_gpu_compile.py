"""compileGeometry on the device (mcr_compile_geometry): every array of the compiled geometry —
triangles, normals, uvs, tangents, colours, materials, flags, sides, the BVH — and the epsilon
byte-identical to the host mirror of compile.ts (compile.py, whose BVH comes from the same
bit-exact builder), on the config scenes and on meshes with groups, a drawRange, a mirroring
transform, no index, no normals and vertex colours."""
import numpy as np
import pytest

from meercadraytracer_amd import abi, native, scenes, three
from meercadraytracer_amd.compile import (add_mesh, compile_geometry, compile_scene, compile_scene_device,
                                          compile_shading, mesh_descriptor)

pytestmark = pytest.mark.gpu

FIELDS = ('triangles', 'normals', 'uvs', 'tangents', 'colors', 'materials', 'flags', 'sides')


def _same(a, b, label):
    for f in FIELDS:
        x, y = np.asarray(getattr(a, f)), np.asarray(getattr(b, f))
        assert x.dtype == y.dtype and x.shape == y.shape, (label, f, x.shape, y.shape)
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), (label, f)
    assert a.triangleCount == b.triangleCount
    assert a.bvh.nodeCount == b.bvh.nodeCount
    assert np.array_equal(np.asarray(a.bvh.nodes), np.asarray(b.bvh.nodes)), label
    assert np.array_equal(np.asarray(a.bvh.bounds).view(np.uint32), np.asarray(b.bvh.bounds).view(np.uint32)), label
    assert a.epsilon == b.epsilon, (label, a.epsilon, b.epsilon)


@pytest.mark.parametrize('kind,seed,kw', [('plain_cpu', 0, {}), ('plain_gpu', 2, dict(quads=30, spheres=12, subdivisions=2)),
                                          ('plain_gpu', 0, {})])
def test_config_scenes(kind, seed, kw):
    ctx = native.default_context()
    if kind == 'plain_cpu':
        scene, camera, _ = scenes.plain_cpu_scene(seed)
    else:
        scene, camera, _ = scenes.plain_gpu_scene(seed, **kw)
    print(f'{kind} {seed} {kw}', flush=True)
    host = compile_scene(scene, ctx.build_bvh)
    dev = compile_scene_device(scene, ctx)
    _same(dev.geometry, host.geometry, kind)


def _mat(cls=three.MeshStandardMaterial, **kw):
    return cls(**kw)


def _part_mesh(geo, mat, world_mesh, index=None):
    """A single-material mesh over a subset of geo's index (or vertices), with world_mesh's transform."""
    g = three.BufferGeometry(geo.position, geo.normal, geo.uv, index, geo.color)
    m = three.Mesh(g, mat)
    m.matrixWorld = world_mesh.matrixWorld
    m.castShadow, m.receiveShadow = world_mesh.castShadow, world_mesh.receiveShadow
    return m


def test_groups_draw_range_mirror_colors():
    ctx = native.default_context()
    rng = np.random.Generator(np.random.PCG64(8))
    plane = three.PlaneGeometry(4, 3, 9, 7)
    plane.setAttribute('color', rng.random((len(plane.position), 4)), 4)
    a, b, c = _mat(color=0x3366aa), _mat(three.MeshPhongMaterial, flatShading=True), _mat(vertexColors=True,
                                                                                           transparent=True)
    m1 = three.Mesh(plane, a)
    m1.position, m1.scale = [0.5, -1, 2], [-1.0, 1.5, 1.0]   # mirroring: winding swapped
    m1.rotation = [0.3, 0.2, -0.1]
    m1.castShadow = True
    # a non-indexed mesh without normals (face normals), with colours of size 3
    box = three.BoxGeometry(1, 2, 3)
    pos = box.position[box.index]
    soup = three.BufferGeometry(pos, None, box.uv[box.index], None, rng.random((len(pos), 3)))
    m2 = three.Mesh(soup, c)
    m2.position = [3, 0, 0]
    m2.receiveShadow = True
    root = three.Scene()
    root.add(m1, m2)
    root.updateMatrixWorld()
    index = {id(a): 0, id(b): 1, id(c): 2}
    # Device: m1 with two groups under a drawRange, then m2 whole.
    d1 = mesh_descriptor(m1, 0, a)
    n_idx = len(plane.index)
    draw = (6, n_idx - 20)
    d1['draw'] = draw
    flags_a = d1['groups'][0][3]
    flags_b = abi.TRI_CAST_SHADOW | abi.TRI_OPAQUE
    d1['groups'] = [(0, 90, 0, flags_a, a.side, False, False), (90, None, 1, flags_b, b.side, True, False)]
    d2 = mesh_descriptor(m2, 2, c)
    dev = native.compile_meshes(ctx, [d1, d2])
    # Host mirror: the same triangles as single-material meshes over index subsets.
    end = min(n_idx, draw[0] + draw[1])
    parts = [add_mesh(_part_mesh(plane, a, m1, plane.index[max(draw[0], 0):min(end, 90)]), 0, a),
             add_mesh(_part_mesh(plane, b, m1, plane.index[max(draw[0], 90):end]), 1, b),
             add_mesh(m2, 2, c)]
    host = compile_geometry(parts, ctx.build_bvh)
    from meercadraytracer_amd.types import CompiledBvh, CompiledGeometry
    got = CompiledGeometry(triangleCount=dev['count'], triangles=dev['triangles'], normals=dev['normals'],
                           uvs=dev['uvs'], tangents=dev['tangents'], colors=dev['colors'],
                           materials=dev['materials'], flags=dev['flags'], sides=dev['sides'],
                           bvh=CompiledBvh(dev['bounds'], dev['nodes'], dev['node_count']), epsilon=dev['epsilon'])
    assert host.colors.size == host.triangleCount * 12  # colours present, white where a mesh has none
    _same(got, host, 'groups')


def test_empty_and_errors():
    ctx = native.default_context()
    out = native.compile_meshes(ctx, [])
    assert out['count'] == 0 and out['node_count'] == 0 and out['epsilon'] == 1e-5
    pos = np.zeros((3, 3), np.float32)
    bad = {'world': np.eye(4).reshape(-1), 'positions': pos, 'index': np.array([0, 1, 7], np.uint32),
           'groups': [(0, None, 0, 0, 0, False, False)]}
    with pytest.raises(native.NativeError) as e:
        native.compile_meshes(ctx, [bad])
    assert 'out of range' in str(e.value)


@pytest.mark.parametrize('kind,seed,kw', [('plain_cpu', 0, {}), ('plain_gpu', 2, dict(quads=30, spheres=12, subdivisions=2)),
                                          ('textured_gpu', 1, dict(quads=40, spheres=10, subdivisions=2))])
def test_resident_compile_renders_identically(kind, seed, kw):
    """mcr_scene_compile_geometry: the compiled arrays laid out as the scene's device records
    without leaving the device. Every float of rendered tiles equals the render of a scene given
    the downloaded arrays (mcr_compile_geometry + mcr_scene_set_geometry) and the oracle's; the
    wide BVH builds from the resident layout and renders the same as from the host layout."""
    import oracle
    from meercadraytracer_amd.compile import compile_camera, compile_scene_resident
    from meercadraytracer_amd.raytrace import trace_settings
    from meercadraytracer_amd.types import Tile
    ctx = native.default_context()
    maker = {'plain_cpu': scenes.plain_cpu_scene, 'plain_gpu': scenes.plain_gpu_scene,
             'textured_gpu': scenes.textured_gpu_scene}[kind]
    scene, camera, o = maker(seed, 64, 48, **kw) if kind != 'plain_cpu' else maker(seed, 64, 48)
    cam = compile_camera(camera)
    st = trace_settings(dict(o, maxBounces=4))
    host = compile_scene_device(scene, ctx)
    a = native.Scene(ctx)
    a.set_geometry(host.geometry, 1)
    a.set_shading(host.shading, 1)
    b = native.Scene(ctx)
    shading, info = compile_scene_resident(scene, b, 1)
    assert info['triangle_count'] == host.geometry.triangleCount and not info['host_layout']
    assert info['node_count'] == host.geometry.bvh.nodeCount and info['epsilon'] == host.geometry.epsilon
    try:
        for tile, p in [(Tile(0, 0, 64, 48), 0), (Tile(8, 4, 40, 30), 3)]:
            print(f'{kind} resident tile {tile} pass {p}', flush=True)
            ga = a.render_tile(cam, st, tile, 64, 48, p, 1)
            gb = b.render_tile(cam, st, tile, 64, 48, p, 1)
            ref, _ = oracle.render_tile(host, cam, st, tile, 64, 48, p, 1)
            assert np.array_equal(ga.view(np.uint32), gb.view(np.uint32)), kind
            assert np.array_equal(gb.view(np.uint32), ref.view(np.uint32)), kind
        ia, ib = a.build_wide(), b.build_wide()
        assert ia['node_count'] == ib['node_count'] and ia['triangle_count'] == ib['triangle_count']
        ctx.set_intersection(abi.MCR_ISECT_WIDE_FP32)
        try:
            wa = a.render_tile(cam, st, Tile(0, 0, 64, 48), 64, 48, 1, 1)
            wb = b.render_tile(cam, st, Tile(0, 0, 64, 48), 64, 48, 1, 1)
        finally:
            ctx.set_intersection(abi.MCR_ISECT_FP64)
        assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32))
    finally:
        a.release()
        b.release()


def test_resident_compile_empty_and_errors():
    ctx = native.default_context()
    s = native.Scene(ctx)
    try:
        info = s.compile_meshes([], 3)
        assert info['triangle_count'] == 0 and info['epsilon'] == 1e-5
        pos = np.zeros((3, 3), np.float32)
        bad = {'world': np.eye(4).reshape(-1), 'positions': pos, 'index': np.array([0, 1, 7], np.uint32),
               'groups': [(0, None, 0, 0, 0, False, False)]}
        with pytest.raises(native.NativeError) as e:
            s.compile_meshes([bad], 4)
        assert 'out of range' in str(e.value)
    finally:
        s.release()


def test_resident_compile_coincident_pile():
    """500 coincident triangles over a plane: buildBvh halves runs whose centroids coincide
    (bvh.ts:158-162), so the resident layout packs every leaf; the scene renders exactly as the
    oracle renders the downloaded arrays."""
    import oracle
    from meercadraytracer_amd.compile import compile_camera, compile_scene_resident
    from meercadraytracer_amd.raytrace import trace_settings
    from meercadraytracer_amd.types import Tile
    ctx = native.default_context()
    tri = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]] * 500, np.float32)
    root = three.Scene()
    pile = three.Mesh(three.BufferGeometry(tri), three.MeshStandardMaterial(color=0x88aacc))
    plane = three.Mesh(three.PlaneGeometry(6, 6), three.MeshLambertMaterial(color=0xcc8844))
    plane.position = [0.0, 0.0, -0.5]
    sun = three.DirectionalLight(0xffffff, 2)
    sun.position = [1.0, 2.0, 5.0]
    root.add(pile, plane, sun)
    root.updateMatrixWorld()
    camera = three.PerspectiveCamera(50, 1.0, 0.1, 100)
    camera.position = [0.0, 0.0, 6.0]
    camera.lookAt(0, 0, 0)
    camera.updateMatrixWorld()
    cam = compile_camera(camera)
    st = trace_settings({})
    host = compile_scene_device(root, ctx)
    b = native.Scene(ctx)
    try:
        shading, info = compile_scene_resident(root, b, 1)
        assert not info['host_layout'] and info['triangle_count'] == 502
        g = b.render_tile(cam, st, Tile(0, 0, 32, 32), 32, 32, 0, 1)
        ref, _ = oracle.render_tile(host, cam, st, Tile(0, 0, 32, 32), 32, 32, 0, 1)
        assert np.array_equal(g.view(np.uint32), ref.view(np.uint32))
    finally:
        b.release()
