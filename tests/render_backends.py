"""render(scene, camera, options, w, h) through the oracle or through the GPU product path, as
tests/helpers.ts:26-36 renders: linear output and 1 sample unless asked."""
import math

import numpy as np
import pytest

import oracle
from meercadraytracer_amd.compile import compile_camera, compile_scene
from meercadraytracer_amd.raytrace import trace_settings

BACKENDS = [pytest.param('oracle', id='oracle'), pytest.param('gpu', id='gpu', marks=pytest.mark.gpu)]


class Rendered:
    def __init__(self, data, width, height):
        self.data, self.width, self.height = data, width, height

    def at(self, x, y):
        i = (y * self.width + x) * 4
        return [self.data[i] / 255, self.data[i + 1] / 255, self.data[i + 2] / 255, self.data[i + 3] / 255]


def raytrace(backend, scene, camera, width, height, options):
    if backend == 'oracle':
        sc = compile_scene(scene, oracle.build_bvh)
        cam = compile_camera(camera)
        o = options or {}
        samples = max(1, math.floor(o.get('samples', 16)))
        srgb = o.get('outputColorSpace', 'srgb') != 'srgb-linear'
        data, _ = oracle.raytrace(sc, cam, trace_settings(o), width, height, samples, srgb)
        return data
    from meercadraytracer_amd.raytrace import raytrace as gpu_raytrace
    return gpu_raytrace(scene, camera, width, height, options)['data']


def render(backend, scene, camera, options=None, width=16, height=16):
    o = {'outputColorSpace': 'srgb-linear', 'samples': 1}
    o.update(options or {})
    return Rendered(raytrace(backend, scene, camera, width, height, o), width, height)
