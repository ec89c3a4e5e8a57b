"""The C-ABI boundary without a GPU: libmcr.so loads and exports every entry point include/mcr.h
declares, and the ctypes mirror (meercadraytracer_amd/abi.py) lays out every struct exactly as a
C compiler lays out the header's (sizes and field offsets, checked with gcc)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from meercadraytracer_amd import abi, native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'mcr.h')


def _declared():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'^\s*(?:const\s+)?(?:int|void|char|uint64_t)\s*\*?\s*(mcr_\w+)\s*\(', text,
                                 flags=re.M)))


def test_header_declares_entry_points():
    names = _declared()
    for n in ('mcr_create', 'mcr_build_bvh', 'mcr_render_tile', 'mcr_render_tiles', 'mcr_render_batch',
              'mcr_encode_tile', 'mcr_intersect_closest', 'mcr_intersect_any_count'):
        assert n in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(native.LIB_PATH):
        pytest.skip('libmcr.so is not built')
    lib = C.CDLL(native.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, f'declared but not exported: {missing}'
    lib.mcr_version.restype = C.c_char_p
    assert b'gfx950' in lib.mcr_version()


STRUCTS = {
    'mcr_geometry': abi.Geometry, 'mcr_material': abi.Material, 'mcr_light': abi.Light,
    'mcr_shading': abi.Shading, 'mcr_camera': abi.Camera, 'mcr_settings': abi.Settings,
    'mcr_tile': abi.Tile, 'mcr_tile_job': abi.TileJob, 'mcr_view': abi.View, 'mcr_counters': abi.Counters,
    'mcr_timing': abi.Timing, 'mcr_server_config': abi.ServerConfig, 'mcr_request': abi.Request,
    'mcr_request_status': abi.RequestStatus, 'mcr_server_stats': abi.ServerStats,
    'mcr_wide_info': abi.WideInfo, 'mcr_mesh_group': abi.MeshGroup, 'mcr_mesh': abi.Mesh,
    'mcr_geometry_out': abi.GeometryOut, 'mcr_texture': abi.Texture, 'mcr_environment': abi.Environment,
    'mcr_compile_info': abi.CompileInfo,
}


def test_ctypes_layout_matches_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', 'int main(void) {']
    expect = []
    for cname, py in STRUCTS.items():
        lines.append(f'  printf("%zu\\n", sizeof({cname}));')
        expect.append(C.sizeof(py))
        for field, _ in py._fields_:
            cf = 'pass' if field == 'pass_' else field
            lines.append(f'  printf("%zu\\n", offsetof({cname}, {cf}));')
            expect.append(getattr(py, field).offset)
    lines += ['  return 0;', '}']
    src = tmp_path / 'layout.c'
    src.write_text('\n'.join(lines))
    exe = tmp_path / 'layout'
    subprocess.run(['gcc', '-std=c11', '-Wall', str(src), '-o', str(exe)], check=True, capture_output=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == expect


def test_noalloc_entry_refuses_null_scene_without_a_gpu():
    """mcr_intersect_closest_noalloc checks its arguments before any HIP call."""
    if not os.path.exists(native.LIB_PATH):
        pytest.skip('libmcr.so is not built')
    L = native.lib()
    rc = L.mcr_intersect_closest_noalloc(None, None, 0, 0.0, 1.0, 1, None, None, None, None, None, None)
    assert rc == abi.MCR_E_INVALID
    assert b'null' in L.mcr_last_error()
