"""mcr_compile_count (host only, no GPU): the triangle count of addMesh's loops
(compile.ts:220-251) over drawRange and groups, checked against the loop itself restated in
Python, and the argument checks."""
import ctypes as C

import numpy as np
import pytest

from meercadraytracer_amd import abi, native


def _loop_count(total, draw, groups):
    """compile.ts:220-225, 250: for (k = from; k + 2 < to; k += 3) over each group."""
    start = max(0, draw[0])
    end = min(total, start + (np.inf if draw[1] is None else draw[1]))
    n = 0
    for g_start, g_count in groups:
        frm = max(start, g_start)
        to = min(end, g_start + (np.inf if g_count is None else g_count))
        k = frm
        while k + 2 < to:
            n += 1
            k += 3
    return n


def _count(meshes):
    arr = (abi.Mesh * len(meshes))()
    keep = []
    for i, (vertices, index_count, draw, groups, colors) in enumerate(meshes):
        pos = np.zeros((vertices, 3), np.float32)
        keep.append(pos)
        arr[i].positions, arr[i].vertex_count = pos.ctypes.data, vertices
        if index_count is not None:
            idx = np.zeros(index_count, np.uint32)
            keep.append(idx)
            arr[i].index, arr[i].index_count = idx.ctypes.data, index_count
        if colors:
            col = np.ones((vertices, 4), np.float32)
            keep.append(col)
            arr[i].colors, arr[i].color_size = col.ctypes.data, 4
        arr[i].draw_start = draw[0]
        arr[i].draw_count = abi.INF_COUNT if draw[1] is None else draw[1]
        g = (abi.MeshGroup * max(1, len(groups)))()
        for k, (s, c) in enumerate(groups):
            g[k] = abi.MeshGroup(s, abi.INF_COUNT if c is None else c, 0, 0, 0, 0, int(colors))
        keep.append(g)
        arr[i].groups, arr[i].group_count = C.cast(g, C.c_void_p), len(groups)
    n, any_colors = C.c_uint32(), C.c_int()
    native.check(native.lib().mcr_compile_count(C.cast(arr, C.c_void_p), len(meshes), C.byref(n), C.byref(any_colors)))
    return n.value, any_colors.value


def test_counts_follow_addmesh_loops():
    rng = np.random.Generator(np.random.PCG64(3))
    for trial in range(300):
        vertices = int(rng.integers(0, 40))
        index_count = None if rng.random() < 0.4 else int(rng.integers(0, 60))
        total = vertices if index_count is None else index_count
        draw = (int(rng.integers(0, 10)), None if rng.random() < 0.5 else int(rng.integers(0, 50)))
        groups = [(int(rng.integers(0, 40)), None if rng.random() < 0.3 else int(rng.integers(0, 40)))
                  for _ in range(int(rng.integers(0, 4)))]
        got, _ = _count([(vertices, index_count, draw, groups, False)])
        assert got == _loop_count(total, draw, groups), (trial, vertices, index_count, draw, groups)


def test_any_colors_and_empty():
    assert _count([]) == (0, 0)
    assert _count([(9, None, (0, None), [(0, None)], False)]) == (3, 0)
    assert _count([(9, None, (0, None), [(0, None)], False), (6, None, (0, None), [(0, None)], True)]) == (5, 1)
    # a group with colours that draws nothing does not turn colours on
    assert _count([(9, None, (9, None), [(0, None)], True)]) == (0, 0)


def test_bad_color_size():
    arr = (abi.Mesh * 1)()
    pos = np.zeros((3, 3), np.float32)
    col = np.zeros((3, 2), np.float32)
    arr[0].positions, arr[0].vertex_count = pos.ctypes.data, 3
    arr[0].colors, arr[0].color_size = col.ctypes.data, 2
    n = C.c_uint32()
    with pytest.raises(native.NativeError):
        native.check(native.lib().mcr_compile_count(C.cast(arr, C.c_void_p), 1, C.byref(n), None))
