"""Pin of the AC1D kernels' division by the bin width (IAC/AC1D.cpp:212, 272:
q = q / dq_per_Step in float).

The kernels compute (float)((double)q * RN(1/(double)dq)) instead of the float
division.  The claim: the exact quotient of two floats is never within 2^-49
(relative) of a float rounding midpoint, while the double product errs by at
most 2^-52, so both round to the same float whenever the quotient is 0 or a
normal float.  numpy's float32 division is IEEE (correctly rounded) and is the
reference; a plausible slip (a float reciprocal, a float product, a truncating
conversion) fails these checks.
"""
import numpy as np

import incohautocorrelate_amd as iac
from incohautocorrelate_amd import workloads as W

f32, f64 = np.float32, np.float64


def fast(q, dq):
    return (q.astype(f64) * (f64(1.0) / f64(dq))).astype(f32)


def exact(q, dq):
    out = q / f32(dq)
    assert out.dtype == np.float32
    return out


def test_random_quotients():
    rng = np.random.default_rng(0)
    for _ in range(20):
        dq = f32(2.0 ** rng.uniform(-40, 40))
        q = (rng.uniform(0.5, 1.0, 1_000_000) * 2.0 ** rng.integers(-60, 60, 1_000_000)).astype(f32)
        assert np.array_equal(fast(q, dq).view(np.uint32), exact(q, dq).view(np.uint32))


def test_near_midpoints():
    """q = RN(m dq) and its float neighbours, m a float rounding midpoint: the
    quotients that land closest to where the float rounding flips."""
    rng = np.random.default_rng(1)
    for _ in range(40):
        dq = f32(rng.uniform(1e-4, 1e-2))
        f = rng.uniform(1.0, 4096.0, 200_000).astype(f32)
        m = f.astype(f64) + np.spacing(f).astype(f64) / 2  # midpoints above f
        q0 = (m * f64(dq)).astype(f32)
        for k in range(-2, 3):
            q = (q0.view(np.int32) + k).view(f32)
            assert np.array_equal(fast(q, dq).view(np.uint32), exact(q, dq).view(np.uint32))


def test_bin_widths_of_real_detectors():
    """Every pair distance of a 96^2 detector against the AC1D bin widths the
    kernels see (4096, 1000 and 600 bins)."""
    _, km, mq = W.detector(96)
    k = km.reshape(-1, 3)[::7]
    d = k[:, None, :] - k[None, :, :]
    q = np.sqrt(((d[..., 0] * d[..., 0]) + (d[..., 1] * d[..., 1])) + (d[..., 2] * d[..., 2])).ravel()
    assert q.dtype == np.float32
    for size in (4096, 1000, 600):
        _, dq = iac.ac1d_shape(mq, size)
        assert np.array_equal(fast(q, dq).view(np.uint32), exact(q, dq).view(np.uint32))


def test_fixed_point_step():
    """The AC_uw fixed point FloatToInt(float) (IAC/Settings.h:168-175) divides by the
    step 0.001f; the kernels take the same short division.  Random magnitudes, values
    around every half-integer quotient (where floor(t + 0.5f) flips), and the full
    formula against the oracle's C restatement."""
    import oracle as O

    step = f32(0.001)
    rng = np.random.default_rng(2)
    q = (rng.uniform(0.5, 1.0, 2_000_000) * 2.0 ** rng.integers(-40, 40, 2_000_000)).astype(f32)
    q = np.concatenate([q, -q])
    assert np.array_equal(fast(q, step).view(np.uint32), exact(q, step).view(np.uint32))
    # p near (k + 1/2) * step and its neighbours: the quotient lands next to a rounding flip
    k = rng.integers(0, 4_000_000, 300_000).astype(np.float64)
    p0 = ((k + 0.5) * f64(step)).astype(f32)
    for d in range(-3, 4):
        p = (p0.view(np.int32) + d).view(f32)
        assert np.array_equal(fast(p, step).view(np.uint32), exact(p, step).view(np.uint32))
    L = O.lib()
    sample = np.concatenate([p0[:2000], q[:2000], rng.gamma(1.0, 40.0, 2000).astype(f32)])
    for p in sample:
        t = fast(np.array([p], f32), step)[0] + f32(0.5)
        short = int(np.floor(t).astype(np.int64)) & 0xFFFFFFFF
        assert L.ora_pair_val(float(p)) == short, p
