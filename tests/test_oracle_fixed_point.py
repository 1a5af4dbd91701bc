"""Pins of the oracle's fixed-point conversion (IAC/Settings.h:168-175).

FloatToInt(float) = (long)floorf((f + 0.0f) / 0.001f + 0.5f), narrowed to
unsigned int by ACMesh.cpp:390-391.  The checks compute the same quantity
with numpy's IEEE float32 (independent of the C build: catches x87 excess
precision, FMA contraction or a double-precision slip) and pin the fact the
GPU kernel's fast path relies on: val(p) == 1000*p exactly for integer
p < 21,054 and not at p = 21,054.
"""
import numpy as np

import oracle as O

f32 = np.float32


def np_val(p):
    p = np.asarray(p, dtype=np.float32)
    t = (p / f32(0.001)) + f32(0.5)
    assert t.dtype == np.float32
    return (np.floor(t).astype(np.int64) & 0xFFFFFFFF).astype(np.uint64)


def test_fast_path_exhaustive():
    p = np.arange(0, 21054, dtype=np.float32)
    assert np.array_equal(np_val(p), (np.arange(21054, dtype=np.uint64) * 1000))
    # the first integer where the float formula departs from 1000*p
    assert int(np_val(f32(21054))) != 21054 * 1000


def test_oracle_matches_numpy_float32():
    L = O.lib()
    rng = np.random.default_rng(0)
    ints = np.concatenate([np.arange(0, 22000), rng.integers(22000, 2 ** 24, 3000)]).astype(np.float32)
    floats = rng.gamma(1.0, 30.0, 3000).astype(np.float32)
    for p in np.concatenate([ints, floats, [f32(4294967.0), f32(65535.0 * 65535.0)]]):
        assert L.ora_pair_val(float(p)) == int(np_val(p)), p


def test_unsigned_wraparound():
    # 65535^2 -> p/0.001 ~ 4.29e12: the long is truncated modulo 2^32
    p = f32(65535.0) * f32(65535.0)
    full = int(np.floor(f32(p / f32(0.001)) + f32(0.5)))
    assert O.pair_val(p) == full % (1 << 32)
    assert full > (1 << 32)


def test_float_to_int_double_overload():
    L = O.lib()
    step = float(f32(0.001))  # (double)0.001f
    for v in [0.0, 0.25, 0.4999999, 1.0, 10526.9, 10527.0, 21054.0, 123456.789]:
        exp = int(np.floor(v / step + 0.5))
        assert L.ora_float_to_int_d(v) == exp
    # the double path differs from 1000*t once t >= 10,527 (the OpenCL host re-quantisation)
    assert L.ora_float_to_int_d(10527.0) != 10527000
    assert L.ora_float_to_int_d(10526.0) == 10526000


def test_int_to_float():
    L = O.lib()
    for i in [0, 1, 999, 1000, 123456789, 2 ** 40]:
        assert L.ora_int_to_float_u64(i) == float(f32(f32(i) * f32(0.001)))


def test_round_half_away_by_biased_truncation():
    """The pair engine's fast path rounds q * vox with trunc(x + copysign(0.49999997f, x))
    (float32 addition, round to nearest).  It equals roundf for every float with
    |x| < 2^23 (checked exhaustively once, 2.5e9 values); here: every half-integer and
    its two float neighbours up to 2^21, plus 4M random floats over the whole range."""
    import numpy as np

    h = np.float32(0.49999997)
    assert h == np.nextafter(np.float32(0.5), np.float32(0))

    def ref(x):
        x64 = x.astype(np.float64)
        return np.trunc(x64 + np.copysign(0.5, x64))

    def fast(x):
        return np.trunc((x + np.copysign(h, x)).astype(np.float32)).astype(np.float64)

    half = (np.arange(0, 1 << 21, dtype=np.float64) + 0.5).astype(np.float32)
    cand = np.concatenate([half, np.nextafter(half, np.float32(0)), np.nextafter(half, np.float32(1 << 23))])
    cand = np.concatenate([cand, -cand])
    assert np.array_equal(fast(cand), ref(cand))
    rng = np.random.default_rng(0)
    bits = rng.integers(0, 1 << 32, 4 << 20, dtype=np.uint64).astype(np.uint32).view(np.float32)
    x = bits[np.isfinite(bits) & (np.abs(bits) < np.float32(1 << 23))]
    assert np.array_equal(fast(x), ref(x))
