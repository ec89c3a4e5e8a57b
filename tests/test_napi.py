"""The Node-API addon (meercadraytracer_amd/napi/mcr_napi.c) without a GPU: it builds against
Node's headers, loads into Node, and its host-only exports answer as the C-ABI does: the version
string, and addMesh's triangle counts (compile.ts:235-251) for meshes given as the reference's
plain objects, compared with mcr_compile_count called through ctypes."""
import ctypes as C
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from meercadraytracer_amd import abi, build, native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope='module')
def addon():
    if not shutil.which('node'):
        pytest.skip('node is not installed')
    if not os.path.exists(native.LIB_PATH):
        pytest.skip('libmcr.so is not built')
    path = build.build_napi()
    if path is None:
        pytest.skip("Node's headers are not installed")
    return path


def _node(script):
    r = subprocess.run(['node', '-e', script], capture_output=True, text=True, timeout=60,
                       cwd=os.path.join(ROOT, 'meercadraytracer_amd', 'napi'))
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_addon_loads_and_reports_version(addon):
    out = _node("console.log(require('./mcr.node').version())")
    assert out.strip() == native.lib().mcr_version().decode()


def _random_meshes(rng, n):
    meshes = []
    for _ in range(n):
        nv = int(rng.integers(3, 40))
        indexed = bool(rng.integers(0, 2))
        ni = int(rng.integers(1, 30)) * 3
        groups = []
        for _ in range(int(rng.integers(1, 4))):
            groups.append({'start': int(rng.integers(0, 20)), 'count': None if rng.integers(0, 3) == 0 else int(rng.integers(0, 40)),
                           'material': int(rng.integers(0, 3)), 'flags': 7, 'side': 0,
                           'vertexColors': False})
        color_size = int(rng.choice([0, 3, 4]))
        for g in groups:
            g['vertexColors'] = bool(color_size and rng.integers(0, 2))
        meshes.append({'vertexCount': nv, 'indexCount': ni if indexed else 0, 'colorSize': color_size,
                       'drawStart': int(rng.integers(0, 10)),
                       'drawCount': None if rng.integers(0, 2) == 0 else int(rng.integers(0, 60)), 'groups': groups})
    return meshes


def _ctypes_count(meshes):
    arr = (abi.Mesh * len(meshes))()
    keep = []
    for m, d in zip(meshes, arr):
        pos = np.zeros(m['vertexCount'] * 3, np.float32)
        idx = np.zeros(m['indexCount'], np.uint32)
        col = np.zeros(m['vertexCount'] * m['colorSize'], np.float32)
        gs = (abi.MeshGroup * len(m['groups']))()
        for g, gd in zip(m['groups'], gs):
            gd.start, gd.count = g['start'], 0xffffffff if g['count'] is None else g['count']
            gd.material, gd.flags, gd.side, gd.vertex_colors = g['material'], g['flags'], g['side'], int(g['vertexColors'])
        keep += [pos, idx, col, gs]
        d.positions = pos.ctypes.data
        d.index = idx.ctypes.data if m['indexCount'] else None
        d.colors = col.ctypes.data if m['colorSize'] else None
        d.groups = C.cast(gs, C.c_void_p)
        d.vertex_count, d.index_count, d.color_size, d.group_count = m['vertexCount'], m['indexCount'], m['colorSize'], len(gs)
        d.draw_start = m['drawStart']
        d.draw_count = 0xffffffff if m['drawCount'] is None else m['drawCount']
    n, colors = C.c_uint32(), C.c_int()
    native.check(native.lib().mcr_compile_count(C.cast(arr, C.c_void_p), len(meshes), C.byref(n), C.byref(colors)))
    return n.value, bool(colors.value)


def test_compile_count_matches_the_c_abi(addon):
    rng = np.random.Generator(np.random.PCG64(21))
    cases = [_random_meshes(rng, int(rng.integers(1, 5))) for _ in range(40)]
    script = '''
const m = require('./mcr.node');
const cases = %s;
const out = cases.map(meshes => m.compileCount(meshes.map(x => ({
  positions: new Float32Array(x.vertexCount * 3),
  index: x.indexCount ? new Uint32Array(x.indexCount) : null,
  colors: x.colorSize ? new Float32Array(x.vertexCount * x.colorSize) : null,
  colorSize: x.colorSize, drawStart: x.drawStart, drawCount: x.drawCount === null ? Infinity : x.drawCount,
  groups: x.groups.map(g => Object.assign({}, g, { count: g.count === null ? Infinity : g.count })),
}))));
console.log(JSON.stringify(out));
''' % json.dumps(cases)
    got = json.loads(_node(script))
    for meshes, g in zip(cases, got):
        n, colors = _ctypes_count(meshes)
        assert g == {'triangleCount': n, 'anyColors': colors}


def test_gpu_exports_report_errors_as_exceptions(addon):
    """Without a device (or with bad arguments) the exports throw JS errors carrying the library's
    message rather than crashing the process."""
    out = _node('''
const m = require('./mcr.node');
try { m.buildBvh(new Float32Array(9), 2); console.log('no throw'); } catch (e) { console.log('threw:' + e.message); }
''')
    assert out.startswith('threw:') and 'fewer than count * 9' in out
