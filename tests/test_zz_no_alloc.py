"""mcr_intersect_closest_noalloc (the entry point that allocates nothing) against
mcr_intersect_closest on the same rays in device memory: every output identical, on two calls in
a row (the second smaller, so it runs on the scratch the first left behind), the first on the
default stream and the second on a stream of the caller's; the caller waits for its stream
through the HIP runtime, as a caller of this entry point does."""
import ctypes as C

import numpy as np
import pytest

from scenes_common import BVH_CASES, geometry_from, random_rays
from meercadraytracer_amd import abi, native

pytestmark = pytest.mark.gpu


def _outputs(ctx, n):
    return {'tri': native.DeviceBuffer(ctx, n * 4), 't': native.DeviceBuffer(ctx, n * 8),
            'u': native.DeviceBuffer(ctx, n * 8), 'v': native.DeviceBuffer(ctx, n * 8),
            'front': native.DeviceBuffer(ctx, n)}


def _download(bufs, n):
    return [bufs['tri'].download(np.int32, n), bufs['t'].download(np.float64, n),
            bufs['u'].download(np.float64, n), bufs['v'].download(np.float64, n),
            bufs['front'].download(np.uint8, n)]


def _hip():
    hip = C.CDLL('libamdhip64.so')
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    return hip


def _order(b):
    return [b[k].ptr for k in ('tri', 't', 'u', 'v', 'front')]


def test_noalloc_matches_intersect_closest():
    ctx = native.default_context()
    L = native.lib()
    hip = _hip()
    own = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(own)) == 0
    for case, nrays, seed, cull, stream in (('soup20k', 40000, 3, 1, None), ('lcg3000', 5000, 4, 0, own)):
        tris, count = BVH_CASES[case]()
        g, _ = geometry_from(tris, count, ctx.build_bvh, sides=0 if cull else 2)
        scene = native.Scene(ctx)
        scene.set_geometry(g, 1)
        rays = random_rays(nrays, seed)
        print(f'noalloc case {case} rays={nrays} cull={cull}', flush=True)
        dr = native.DeviceBuffer(ctx, rays.nbytes)
        dr.upload(np.ascontiguousarray(rays))
        ref, got = _outputs(ctx, nrays), _outputs(ctx, nrays)
        for b in (ref, got):
            for x in b.values():
                x.zero()
        native.check(L.mcr_intersect_closest(ctx.handle, scene.handle, dr.ptr, nrays, 0.0, np.inf, cull,
                                             *_order(ref), 1, None))
        native.check(L.mcr_intersect_closest_noalloc(scene.handle, dr.ptr, nrays, 0.0, np.inf, cull,
                                                     *_order(got), stream))
        assert hip.hipStreamSynchronize(stream) == 0
        a, b = _download(ref, nrays), _download(got, nrays)
        assert (a[0] >= 0).sum() > nrays // 20
        for x, y in zip(a, b):
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), case
        scene.release()
    assert hip.hipStreamDestroy(own) == 0


def test_noalloc_refuses_oversized_batch():
    ctx = native.default_context()
    tris, count = BVH_CASES['five']()
    g, _ = geometry_from(tris, count, ctx.build_bvh)
    scene = native.Scene(ctx)
    scene.set_geometry(g, 1)
    buf = native.DeviceBuffer(ctx, 64)
    rc = native.lib().mcr_intersect_closest_noalloc(scene.handle, buf.ptr, abi.MCR_NOALLOC_MAX_RAYS + 1, 0.0,
                                                    np.inf, 1, None, None, None, None, None, None)
    assert rc == abi.MCR_E_INVALID
    assert b'2^31' in native.lib().mcr_last_error()
    scene.release()
