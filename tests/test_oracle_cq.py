"""Pins of the oracle's C(q) small, merge and normalisation
(IAC/kernel.cl:27-344, IAC/Detector.cpp:1184-1534, IAC/RunIAC.cpp:120-158).

- A flat field (pixel mask as intensity, Multiplicator 1) makes every small
  C(q) voxel an exact pair count (invariant printed by IAC/UnitTest.cpp:215-230);
  the total is the number of ordered pixel pairs inside Max_Q, counted
  independently with numpy.
- The axis permutation (k_A, k_B, k_C) only relabels mesh axes.
- Merging with the identity and weight 1 embeds the small mesh in the big
  mesh with a constant ratio (the disabled self-check of Detector.cpp:1538-1568).
- Merge_ACandCQ's closed form, zeros where C(q) <= 0.
"""
import numpy as np
import pytest

import oracle as O

f32 = np.float32


def inside_count(km, max_q):
    k = km.astype(np.float32)
    q = (k[:, None, :] - k[None, :, :]).astype(np.float32)
    s2 = (q[..., 0] * q[..., 0] + q[..., 1] * q[..., 1]) + q[..., 2] * q[..., 2]
    return int((~(np.sqrt(s2) > f32(max_q))).sum())


@pytest.mark.parametrize("zoom", [1.0, 2.0])
def test_flat_field_counts_pairs(zoom):
    _, km, mq = O.detector(24)
    s = O.mesh_shape_small_cofq(mq, 15, zoom)
    flat = np.ones(km.shape[0], np.float32)
    mesh, dropped = O.cq_small(flat, km, s, 0, 1.0)
    assert dropped == 0
    assert int(mesh.sum()) == inside_count(km, s.max_q)
    # point symmetric; every pixel pairs with itself at q = 0
    assert np.array_equal(mesh, mesh[::-1])
    c = s.center[0] + s.center[1] * s.size_ab + s.center[2] * s.size_ab ** 2
    assert mesh[c] >= km.shape[0]


def test_multiplicator_rule():
    # smallest 10^k >= 1e-10 with <I>^2 10^k >= 1, times 1e5
    assert O.cq_small_multiplicator(np.full(100, 0.5, np.float32)) == pytest.approx(1e6)
    assert O.cq_small_multiplicator(np.full(100, 3.0, np.float32)) == pytest.approx(1e5)
    assert O.cq_small_multiplicator(np.full(100, 1e-3, np.float32)) == pytest.approx(1e11)


def test_values_are_truncated_products():
    _, km, mq = O.detector(12)
    s = O.mesh_shape_small_cofq(mq, 11)
    I = np.zeros(144, np.float32)
    I[[5, 77]] = [0.3, 0.7]
    m = 1e6
    mesh, _ = O.cq_small(I, km, s, 0, m)
    vals = sorted(int(v) for v in mesh[mesh > 0])
    a, b = f32(0.3), f32(0.7)
    expect = sorted([int(float(a * a) * m) + int(float(b * b) * m)] * 1 + [int(float(a * b) * m)] * 2)
    # self pairs share the centre voxel; (a,b) and (b,a) sit at +-q
    assert vals == sorted(expect)


def test_axis_permutation_only_relabels():
    """A mesh built with axes (k_B, k_A, k_C) is the transpose of the (k_A, k_B, k_C) one."""
    _, km, mq = O.detector(20)
    s = O.mesh_shape_small_cofq(mq, 13)
    rng = np.random.default_rng(0)
    I = rng.gamma(1.0, 0.5, km.shape[0]).astype(np.float32)
    m1, _ = O.cq_small(I, km, s, 0, 1e5)
    t = O.MeshShape.from_any(s)
    t.k_a, t.k_b = s.k_b, s.k_a
    m2, _ = O.cq_small(I, km, t, 0, 1e5)
    S, SC = s.size_ab, s.size_c
    assert np.array_equal(m2.reshape(SC, S, S), m1.reshape(SC, S, S).transpose(0, 2, 1))


def test_linear_conserves_mass_within_truncation():
    _, km, mq = O.detector(16)
    s = O.mesh_shape_small_cofq(mq, 11)
    I = np.ones(256, np.float32)
    m0, _ = O.cq_small(I, km, s, 0, 1e6)
    m1, d1 = O.cq_small(I, km, s, 1, 1e6)
    assert d1 == 0
    npairs = int(m0.sum()) // 10 ** 6
    assert 0 <= int(m0.sum()) - int(m1.sum()) <= 8 * npairs  # truncation of 8 corners per pair


def test_merge_identity_embeds_small_mesh():
    _, km, mq = O.detector(24)
    s = O.mesh_shape_small_cofq(mq, 15)
    rng = np.random.default_rng(1)
    I = rng.gamma(1.0, 0.5, km.shape[0]).astype(np.float32)
    mult = O.cq_small_multiplicator(I)
    small_u, _ = O.cq_small(I, km, s, 0, mult)
    small = small_u.astype(np.float64) / mult
    rot = np.eye(3, dtype=np.float32).reshape(1, 9)
    rw, m = O.merge_prepare(rot, np.array([1.0], np.float32), small)
    assert rw[0, 9] == 1.0
    big, dropped = O.merge_cq(small, s, rw, 0, m)
    assert dropped == 0
    S, SC = s.size_ab, s.size_c
    g = big.reshape(S, S, S)
    emb = g[(S - 1) // 2 - (SC - 1) // 2:(S - 1) // 2 + (SC - 1) // 2 + 1]
    assert emb.sum() == g.sum()  # nothing outside the embedded slab
    nz = small.reshape(SC, S, S) > 0
    ratio = emb[nz] / small.reshape(SC, S, S)[nz]
    assert np.allclose(ratio, m, rtol=1e-9)


def test_merge_prepare_multiplicator_and_weights():
    small = np.full(1000, 2.0)
    rot = np.tile(np.eye(3, dtype=np.float32).reshape(9), (3, 1))
    mi = np.array([0.5, 0.25, 0.123456], np.float32)
    rw, m = O.merge_prepare(rot, mi, small)
    # 1e-12 * 10^k until maxW^2 * <C> * M >= 1 (0.25 * 2 -> 1e1... ), / OOM(3 events -> 10), * 1e10
    mm = 1e-12
    while 1 > 0.5 * 0.5 * 2.0 * mm:
        mm *= 10
    mm = mm / 10 * 1e10
    mm = round(mm) if mm > 0.6 else mm
    assert m == mm
    red = f32(round(f32(m))) if m > 0.6 else f32(m)
    for i, w in enumerate(mi):
        wq = f32(np.floor(f32(f32(f32(w) * f32(w)) * red) + 0.5))
        assert rw[i, 9] == f32(wq / red)


def test_merge_ac_cq_closed_form():
    ac = np.array([0, 1000, 2500, 7, 123456789], np.uint64)
    cq = np.array([1.0, 0.0, 2.0, -1.0, 3.0])
    out = O.merge_ac_cq(ac, cq)
    expect = [0.0, 0.0, float(f32(2500) * f32(0.001)) / 2.0, 0.0, float(f32(123456789) * f32(0.001)) / 3.0]
    assert np.array_equal(out, np.array(expect))


def test_fp64_oracle_agrees_with_fixed_point():
    """The FP64 dense oracle and the uint64 fixed point (huge M) agree: same
    binning, values differ only by the float product and the truncation."""
    _, km, mq = O.detector(32)
    s = O.mesh_shape_small_cofq(mq, 15)
    rng = np.random.default_rng(3)
    I = rng.gamma(1.0, 0.5, km.shape[0]).astype(np.float32)
    f64 = O.cq_small_f64(I, km, s)
    u, _ = O.cq_small(I, km, s, 0, 1e12)
    fx = u.astype(np.float64) / 1e12
    nz = f64 > 0
    assert np.array_equal(nz, fx > 0)
    assert np.abs(fx[nz] - f64[nz]).max() / f64[nz].min() < 1e-6
    # merge of the FP64 mesh with the identity embeds it unchanged
    rw = np.zeros(10, np.float32)
    rw[[0, 4, 8, 9]] = 1
    big = O.merge_f64(f64, s, rw)
    assert big.sum() == pytest.approx(f64.sum(), rel=1e-12)
