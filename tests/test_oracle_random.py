"""Known answers for the oracle's restatement of random.ts (hash, Random.seed, Random.next), which
the renders consume draw by draw; each check is fixed by a published result or by arithmetic,
not by the oracle itself."""
import numpy as np

import oracle


def _stream(state, n):
    out = np.zeros(n, np.float64)
    oracle.load().orc_random_stream(state, n, out.ctypes.data)
    return out


def _inverse_lowbias32(x):
    """C. Wellons' published inverse of lowbias32 (hash-prospector, 2018): multipliers 0x43021123
    and 0x1d69e2a5 undo 0x846ca68b and 0x7feb352d modulo 2^32; x ^= x >> 15 ^ x >> 30 undoes
    x ^= x >> 15."""
    m = 0xffffffff
    x ^= x >> 16
    x = (x * 0x43021123) & m
    x ^= (x >> 15) ^ (x >> 30)
    x = (x * 0x1d69e2a5) & m
    x ^= x >> 16
    return x


def test_next_is_marsaglia_xorshift32():
    """random.ts:13-20 is Marsaglia's xorshift32 with shifts (13, 17, 5) (J. Stat. Softw. 8(14),
    2003): from y = 2463534242 the first output is 723471715, and next() returns state / 2^32."""
    v = _stream(2463534242, 1)
    assert v[0] * 4294967296 == 723471715
    assert v[0] == 723471715 / 4294967296


def test_next_period_structure():
    """xorshift32 is a bijection of the nonzero states: a stream never reaches 0 and its outputs
    lie in [0, 1) on multiples of 2^-32."""
    v = _stream(1, 100000)
    assert np.all(v > 0) and np.all(v < 1)
    assert np.all(v * 4294967296 == np.floor(v * 4294967296))
    assert len(np.unique(v)) == len(v)


def test_hash_round_trips_through_published_inverse():
    """random.ts:24-31 is lowbias32 (multipliers 0x7feb352d, 0x846ca68b, shifts 16, 15, 16): the
    published inverse maps every hash back to its input. A wrong multiplier or shift fails this
    on almost every input."""
    lib = oracle.load()
    rng = np.random.Generator(np.random.PCG64(5))
    xs = [0, 1, 2, 0x7fffffff, 0x80000000, 0xffffffff] + rng.integers(0, 2 ** 32, 5000).tolist()
    for x in xs:
        assert _inverse_lowbias32(lib.orc_hash(x)) == x


def test_seed_closed_forms():
    """random.ts:6-10: h = hash(a ^ hash(b ^ hash(c ^ hash(d)))), and 0 becomes 0x9e3779b9. Since
    hash(0) = 0 (every step maps 0 to 0), zero arguments drop out: seed(0,0,0,0) is 0x9e3779b9,
    seed(a,0,0,0) = hash(a), seed(0,b,0,0) = hash(hash(b)), seed(0,0,c,0) = hash^3(c) and
    seed(0,0,0,d) = hash^4(d), which fixes the nesting order of the four arguments."""
    lib = oracle.load()
    h = lib.orc_hash
    assert h(0) == 0
    assert lib.orc_seed(0, 0, 0, 0) == 0x9e3779b9
    for x in (1, 7, 12345, 0x7fffffff):
        assert lib.orc_seed(x, 0, 0, 0) == h(x)
        assert lib.orc_seed(0, x, 0, 0) == h(h(x))
        assert lib.orc_seed(0, 0, x, 0) == h(h(h(x)))
        assert lib.orc_seed(0, 0, 0, x) == h(h(h(h(x))))
    # int32 arguments mix as JS's ^ does (ToInt32): -1 is 0xffffffff.
    assert lib.orc_seed(-1, 0, 0, 0) == h(0xffffffff)


def test_seed_never_zero():
    """The zero state is replaced (xorshift32 would stay at 0 forever). hash is a bijection with
    hash^-1(0) = 0, so the outer hash is 0 only when its argument is; no seed may come out 0."""
    lib = oracle.load()
    assert _inverse_lowbias32(0) == 0
    rng = np.random.Generator(np.random.PCG64(9))
    for a, b, c, d in rng.integers(-2 ** 31, 2 ** 31, (2000, 4)).tolist():
        assert lib.orc_seed(a, b, c, d) != 0
