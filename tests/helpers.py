"""Shared test helpers: the reference tests' generators and scenes (tests/helpers.ts,
tests/bvh.test.ts:4-25), restated."""
import math

import numpy as np

from meercadraytracer_amd import three as T


def random_triangles(count, seed):
    """tests/bvh.test.ts:4-19: LCG s = (s*1664525 + 1013904223) >>> 0, values s / 2^32."""
    s = seed

    def rand():
        nonlocal s
        s = (s * 1664525 + 1013904223) % 4294967296
        return s / 4294967296

    data = np.zeros(count * 9, np.float32)
    for i in range(count):
        cx, cy, cz = rand() * 20 - 10, rand() * 20 - 10, rand() * 20 - 10
        v = [[cx + rand() * 2 - 1, cy + rand() * 2 - 1, cz + rand() * 2 - 1] for _ in range(3)]
        # Float32Array.set of the doubles: each value rounds to f32, then e = v1 - v0 in FP64.
        data[i * 9:i * 9 + 3] = v[0]
        data[i * 9 + 3:i * 9 + 6] = [v[1][k] - v[0][k] for k in range(3)]
        data[i * 9 + 6:i * 9 + 9] = [v[2][k] - v[0][k] for k in range(3)]
    return data


def reorder(data, order):  # tests/bvh.test.ts:21-25
    return np.ascontiguousarray(data.reshape(-1, 9)[order].reshape(-1))


def probe_rays_400():
    """tests/bvh.test.ts:85-91: 400 analytic rays."""
    from meercadraytracer_amd.compile import js_hypot
    rays = np.zeros((400, 6))
    for i in range(400):
        o = [math.sin(i) * 30, math.cos(i * 1.3) * 30, math.sin(i * 0.7) * 30]
        target = [math.sin(i * 2.1) * 8, math.cos(i * 0.4) * 8, math.sin(i * 1.7) * 8]
        d = [target[a] - o[a] for a in range(3)]
        length = js_hypot(*d)
        rays[i, :3] = o
        rays[i, 3:] = [x / length for x in d]
    return rays


def top_camera():  # tests/helpers.ts:5-11
    camera = T.OrthographicCamera(-1, 1, 1, -1, 0.1, 100)
    camera.position = [0.0, 0.0, 10.0]
    camera.lookAt(0, 0, 0)
    camera.updateMatrixWorld()
    return camera


def floor(material, size=4):  # tests/helpers.ts:14-16
    return T.Mesh(T.PlaneGeometry(size, size), material)


def overhead_light(intensity):  # tests/helpers.ts:39-43
    light = T.DirectionalLight(0xffffff, intensity)
    light.position = [0.0, 0.0, 10.0]
    return light


ONE_LEVEL = 1 / 255 + 1e-6  # tests/helpers.ts:45
