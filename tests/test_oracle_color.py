"""tests/color.test.ts restated against the oracle (colour transfer, half floats, tone mapping,
DFG table, GGX normalisation, BRDF helpers), plus pins of the portable math library against
glibc: the oracle's portable build and the device share it, so these pin both."""
import math

import numpy as np
import pytest

import oracle
from meercadraytracer_amd import abi


def mapped(value, mapping, exposure=1.0, libm=False):
    c = np.array(value, dtype=np.float64)
    oracle.load(libm).orc_tone_map(c.ctypes.data, mapping, exposure)
    return c.tolist()


@pytest.mark.parametrize('libm', [False, True])
def test_srgb_round_trip(libm):  # color.test.ts:22-25
    lib = oracle.load(libm)
    for v in [0, 0.001, 0.02, 0.2, 0.5, 0.8, 1]:
        assert abs(lib.orc_srgb_to_linear(lib.orc_linear_to_srgb(v)) - v) < 5e-5
    assert abs(lib.orc_linear_to_srgb(0.5) - 0.7354) < 5e-4


def test_half_floats():  # color.test.ts:27-32
    lib = oracle.load()
    assert lib.orc_half_to_float(0x3c00) == 1
    assert lib.orc_half_to_float(0xc000) == -2
    assert lib.orc_half_to_float(0x3800) == 0.5
    assert lib.orc_half_to_float(0x0000) == 0
    # every finite half equals numpy's decode
    h = np.arange(0, 0x10000, dtype=np.uint16)
    ref = h.view(np.float16).astype(np.float64)
    got = np.array([lib.orc_half_to_float(int(x)) for x in h[::37]])
    fin = np.isfinite(ref[::37])
    assert np.array_equal(got[fin], ref[::37][fin])


@pytest.mark.parametrize('libm', [False, True])
def test_tone_mapping_operators(libm):  # color.test.ts:34-57
    assert mapped([1, 0.5, 4], abi.NoToneMapping, libm=libm) == [1, 0.5, 4]
    assert mapped([0.5, 2, 0], abi.LinearToneMapping, 1, libm) == [0.5, 1, 0]
    assert abs(mapped([1, 3, 0], abi.ReinhardToneMapping, libm=libm)[0] - 0.5) < 5e-3
    assert abs(mapped([1, 3, 0], abi.ReinhardToneMapping, libm=libm)[1] - 0.75) < 5e-3
    assert abs(mapped([1, 0, 0], abi.ReinhardToneMapping, 3, libm)[0] - 0.75) < 5e-3
    x = 0.18 - 0.004
    want = ((x * (6.2 * x + 0.5)) / (x * (6.2 * x + 1.7) + 0.06)) ** 2.2
    assert abs(mapped([0.18] * 3, abi.CineonToneMapping, libm=libm)[0] - want) < 5e-7
    assert abs(mapped([0.5] * 3, abi.NeutralToneMapping, libm=libm)[0] - 0.46) < 5e-7
    for m in (abi.ACESFilmicToneMapping, abi.AgXToneMapping, abi.NeutralToneMapping):
        bright = mapped([20] * 3, m, libm=libm)
        dark = mapped([0.05] * 3, m, libm=libm)
        assert bright[0] <= 1.000001 and bright[0] > dark[0] and dark[0] >= 0
    grey = mapped([0.3] * 3, abi.ACESFilmicToneMapping, libm=libm)
    assert abs(grey[0] - grey[1]) < 5e-3 and abs(grey[1] - grey[2]) < 5e-3


def _dfg_table():
    """The 512 half floats of csrc/dfg_lut.h (the table the device and the oracle both compile)."""
    import os
    import re
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'meercadraytracer_amd', 'csrc',
                        'dfg_lut.h')
    body = open(path).read()
    body = body[body.index('{'):body.rindex('}')]
    vals = [int(h, 16) for h in re.findall(r'0x([0-9a-fA-F]{4})', body)]
    assert len(vals) == 512
    return vals


def test_dfg_texel_centres():  # color.test.ts:61-70
    lib = oracle.load()
    out = np.zeros(2)
    lib.orc_dfg(3.5 / 16, 5.5 / 16, out.ctypes.data)
    DFG_LUT_HALF = _dfg_table()
    assert abs(out[0] - lib.orc_half_to_float(DFG_LUT_HALF[(5 * 16 + 3) * 2])) < 5e-7
    assert abs(out[1] - lib.orc_half_to_float(DFG_LUT_HALF[(5 * 16 + 3) * 2 + 1])) < 5e-7
    lib.orc_dfg(0, 1, out.ctypes.data)
    assert out[0] + out[1] > 0.9


def test_ggx_normalised():  # color.test.ts:72-83
    lib = oracle.load()
    for alpha in (0.1, 0.5, 1):
        steps = 4000
        s = 0.0
        for i in range(steps):
            th = (i + 0.5) / steps * (math.pi / 2)
            s += lib.orc_distribution_ggx(alpha, math.cos(th)) * math.cos(th) * math.sin(th) * 2 * math.pi * (math.pi / 2 / steps)
        assert abs(s - 1) < 5e-3


def test_brdf_helpers():  # color.test.ts:85-93
    lib = oracle.load()
    assert lib.orc_fresnel_weight(0) == 1
    assert lib.orc_fresnel_weight(1) < 0.001
    assert lib.orc_smoothstep(0, 1, 0.5) == 0.5
    assert lib.orc_smoothstep(0.5, 0.5, 0.6) == 1
    assert abs(lib.orc_distance_attenuation(2, 0, 2) - 0.25) < 5e-3
    assert lib.orc_distance_attenuation(0.01, 0, 2) == 100
    assert lib.orc_distance_attenuation(10, 10, 2) == 0


def _ulps(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    ia = a.view(np.int64)
    ib = b.view(np.int64)
    ia = np.where(ia < 0, np.int64(-0x8000000000000000) - ia, ia)
    ib = np.where(ib < 0, np.int64(-0x8000000000000000) - ib, ib)
    return np.abs(ia - ib)


@pytest.mark.parametrize('fn,name,gen,max_ulp', [
    (0, 'exp', lambda r: r.uniform(-700, 700, 4000), 2),
    (0, 'exp_small', lambda r: r.uniform(-5, 5, 4000), 2),
    (1, 'log', lambda r: np.exp(r.uniform(-700, 700, 4000)), 2),
    (1, 'log_near1', lambda r: r.uniform(0.5, 2, 4000), 2),
    (3, 'sin', lambda r: r.uniform(0, 2 * math.pi, 4000), 2),
    (4, 'cos', lambda r: r.uniform(0, 2 * math.pi, 4000), 2),
    (5, 'log2', lambda r: r.uniform(1e-10, 10, 4000), 3),
])
def test_portable_math_against_glibc(fn, name, gen, max_ulp):
    rng = np.random.Generator(np.random.PCG64(hash(name) & 0xffff))
    xs = gen(rng)
    lib = oracle.load()
    got = np.array([lib.orc_math(fn, float(x), 0.0) for x in xs])
    ref = {0: np.exp, 1: np.log, 3: np.sin, 4: np.cos, 5: np.log2}[fn](xs)
    u = _ulps(got, ref)
    # sin/cos near their zeros: absolute error is what matters
    if fn in (3, 4):
        ok = (u <= max_ulp) | (np.abs(got - ref) < 1e-16)
    else:
        ok = u <= max_ulp
    assert ok.all(), f'{name}: worst {u.max()} ulp at {xs[np.argmax(u)]}'


def test_portable_pow_against_glibc():
    rng = np.random.Generator(np.random.PCG64(5))
    lib = oracle.load()
    cases = [(rng.uniform(0, 1, 3000), rng.uniform(1, 200, 3000)),       # blinnPhong dotNH^shininess
             (rng.uniform(0.003, 1.5, 3000), np.full(3000, 0.41666)),     # sRGB encode
             (np.full(3000, 2.0), rng.uniform(-13, 0, 3000)),             # fresnel 2^x
             (rng.uniform(0.01, 50, 3000), np.full(3000, 2.2))]
    for x, y in cases:
        got = np.array([lib.orc_math(2, float(a), float(b)) for a, b in zip(x, y)])
        ref = np.power(x, y)
        rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
        assert np.all((_ulps(got, ref) <= 2) | (ref < 1e-290)), rel.max()
    # ECMAScript special cases (Number::exponentiate)
    pw = lambda a, b: lib.orc_math(2, a, b)
    assert math.isnan(pw(1.0, math.nan)) and pw(math.nan, 0.0) == 1.0
    assert math.isnan(pw(-1.0, math.inf)) and pw(-8.0, 2.0) == 64.0 and math.isnan(pw(-8.0, 0.5))
    assert pw(0.0, 3.0) == 0.0 and pw(0.0, -1.0) == math.inf and pw(3.0, 2.0) == 9.0


def test_js_hypot_matches_v8_algorithm():
    """Math.hypot in V8: normalise by the largest magnitude, Kahan-sum the squares. Known
    answers: 3-4-5, zero, and agreement with math.hypot to 2 ulp on random vectors."""
    lib = oracle.load()
    assert lib.orc_math(6, 3.0, 4.0) == 5.0
    assert lib.orc_math(6, 0.0, 0.0) == 0.0
    rng = np.random.Generator(np.random.PCG64(1))
    for a, b in rng.normal(size=(2000, 2)) * 10 ** rng.uniform(-5, 5, (2000, 1)):
        assert _ulps(lib.orc_math(6, a, b), math.hypot(a, b)) <= 2
