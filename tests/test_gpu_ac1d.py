"""GPU parity of AC1D (angular-averaged C(q), AC_uw, AC) and of the device
sparsifier, against the oracle."""
import numpy as np
import pytest

import oracle as O
from gpu_helpers import mesh_np, to_dev
from incohautocorrelate_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("interp,n,bright", [(0, 96, False), (1, 96, False), (1, 77, False), (0, 77, True),
                                             (1, 64, True)])
def test_ac1d_cq_vs_oracle(ctx, dev, interp, n, bright):
    """n = 77: 5929 pixels, a ragged last tile row and column; the per-block tile
    ranges cross triangle rows.  bright: a few pixels 3e4 times the mean, so that
    weighted values pass 2^32 and the full 64-bit truncation is taken."""
    import torch

    import incohautocorrelate_amd as iac

    _, km, mq = W.detector(n)
    size = 600
    max_q, dq = iac.ac1d_shape(mq, size)
    assert (max_q, dq) == O.ac1d_shape(mq, size)
    rng = np.random.default_rng(interp)
    I = rng.gamma(1.0, 0.5, n * n).astype(np.float32)
    if bright:
        I[rng.integers(0, n * n, 5)] = 1.5e4
    mult = iac.ac1d_cq_multiplicator(I)
    assert mult == O.ac1d_cq_multiplicator(I)
    ref, dref = O.ac1d_cq(I, km, size, max_q, dq, interp, mult)
    h = torch.zeros(size, dtype=torch.int64, device=dev)
    ctx.ac1d_cq(to_dev(km, dev), to_dev(I, dev), size, max_q, dq, interp, mult, h)
    assert dref == 0
    if bright:  # the bright self pairs weigh more than 2^32
        assert float(I.max()) ** 2 * mult >= 2 ** 32
    assert np.array_equal(mesh_np(h), ref)


def test_ac1d_acuw_nearest_bit_exact(ctx, dev):
    """Integer photon counts: the reference's double sums are exact integers."""
    import torch

    import incohautocorrelate_amd as iac

    n = 256
    _, km, mq = W.detector(n)
    size = 1000
    max_q, dq = iac.ac1d_shape(mq, size)
    d = W.lognormal_photon_frames(n * n, 30, 300, 1.0, 1, 3000, seed=4)
    d["cnt"] = np.random.default_rng(1).integers(1, 6, d["cnt"].size).astype(np.uint16)
    for spf in (2, 1):
        ref = np.zeros(size)
        off = d["frame_off"]
        for f in range(off.size - 1):
            hl = O.hit_list_from_photons(km, d["pix"][off[f]:off[f + 1]], d["cnt"][off[f]:off[f + 1]])
            ref, _ = O.ac1d_acuw(hl, size, max_q, dq, 0, spf, ac=ref)
        h = torch.zeros(size, dtype=torch.float64, device=dev)
        ctx.ac1d_acuw(to_dev(km, dev), to_dev(d["pix"], dev), to_dev(d["cnt"], dev), to_dev(off, dev),
                      int(off[-1]), size, max_q, dq, 0, h, self_pair_factor=spf)
        assert np.array_equal(h.cpu().numpy(), ref)


def test_ac1d_acuw_linear_within_double_rounding(ctx, dev):
    import torch

    import incohautocorrelate_amd as iac

    n = 256
    _, km, mq = W.detector(n)
    size = 1000
    max_q, dq = iac.ac1d_shape(mq, size)
    d = W.poisson_photon_frames(n * n, 40, 0.02, seed=5)
    ref = np.zeros(size)
    off = d["frame_off"]
    for f in range(off.size - 1):
        hl = O.hit_list_from_photons(km, d["pix"][off[f]:off[f + 1]], d["cnt"][off[f]:off[f + 1]])
        ref, _ = O.ac1d_acuw(hl, size, max_q, dq, 1, 2, ac=ref)
    h = torch.zeros(size, dtype=torch.float64, device=dev)
    ctx.ac1d_acuw(to_dev(km, dev), to_dev(d["pix"], dev), to_dev(d["cnt"], dev), to_dev(off, dev),
                  int(off[-1]), size, max_q, dq, 1, h)
    got = h.cpu().numpy()
    # the same float terms summed in another order: relative error ~ n * 2^-53
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-9)
    assert abs(got.sum() - ref.sum()) <= 1e-9 * ref.sum()


def test_ac1d_calc_ac(ctx, dev):
    import torch

    a = np.array([1.0, 0.0, 4.0, 2.0, 0.0])
    c = np.array([2.0, 0.0, 0.0, -1.0, 1.0])
    ref = O.ac1d_calc_ac(a, c)
    out = torch.zeros(5, dtype=torch.float64, device=dev)
    ctx.ac1d_calc_ac(torch.from_numpy(a).to(dev), torch.from_numpy(c).to(dev), out)
    assert np.array_equal(out.cpu().numpy(), ref)  # 0/0 -> 0, x/0 -> inf as in the reference


def test_sparse_hit_list_vs_oracle(ctx, dev):
    import torch

    n = 128
    _, km, _ = W.detector(n)
    rng = np.random.default_rng(6)
    frames = (rng.gamma(0.3, 4.0, (5, n * n)) - 0.5).astype(np.float32)
    mask = (rng.random(n * n) > 0.05).astype(np.int32)
    T, S = 0.8, 1.7
    rec = torch.zeros((5 * n * n, 4), dtype=torch.float32, device=dev)
    off = torch.zeros(6, dtype=torch.int64, device=dev)
    nt = ctx.sparse_hit_list(to_dev(km, dev), to_dev(frames, dev), to_dev(mask, dev), T, S, rec, off)
    got_off = off.cpu().numpy()
    got = rec[:nt].cpu().numpy()
    total = 0
    for f in range(5):
        ref = O.sparse_hit_list(frames[f], km, T, S, mask)
        assert got_off[f + 1] - got_off[f] == ref.shape[0]
        assert np.array_equal(got[got_off[f]:got_off[f + 1]], ref)
        total += ref.shape[0]
    assert nt == total


def test_sparsified_frames_through_ac3d(ctx, dev):
    """Dense Poisson frames -> device sparsify -> records batch == oracle mesh."""
    import torch

    import incohautocorrelate_amd as iac

    n = 128
    _, km, mq = W.detector(n)
    shape = iac.mesh_shape_big(mq, 63)
    d = W.dense_poisson_frames(n * n, 20, 0.02, seed=7)
    frames = d["frames"].astype(np.float32)
    rec = torch.zeros((20 * n * n, 4), dtype=torch.float32, device=dev)
    off = torch.zeros(21, dtype=torch.int64, device=dev)
    nt = ctx.sparse_hit_list(to_dev(km, dev), to_dev(frames, dev), None, 0.0, 1.0, rec, off)
    mesh = torch.zeros(shape.size_ab ** 3, dtype=torch.int64, device=dev)
    ctx.ac3d_records_batch(to_dev(km, dev), rec, off, nt, shape, mesh)
    ref, _ = O.ac3d_frames(km, d["pix"], d["cnt"], d["frame_off"], shape)
    assert np.array_equal(mesh_np(mesh), ref)


def test_short_sqrt_exhaustive(ctx):
    """The AC1D kernels' |q| square root against the library sqrtf on every one of
    the 2^32 float bit patterns (zero, subnormals, the small-input branch, inf, NaN)."""
    assert ctx.selfcheck_sqrt(0, 1 << 32) == 0


@pytest.mark.parametrize("interp", [0, 1])
def test_ac1d_cq_and_acuw_with_drops(ctx, dev, interp):
    """A bin width half of AC1D::Initialize's: pairs past the histogram are dropped and
    counted (the range-checked kernels), both for C(q) and for AC_uw, against the oracle."""
    import torch

    import incohautocorrelate_amd as iac

    n = 80
    _, km, mq = W.detector(n)
    size = 500
    max_q, dq0 = iac.ac1d_shape(mq, size)
    dq = float(np.float32(dq0 * 0.5))
    rng = np.random.default_rng(40 + interp)
    I = rng.gamma(1.0, 0.5, n * n).astype(np.float32)
    mult = iac.ac1d_cq_multiplicator(I)
    ref, dref = O.ac1d_cq(I, km, size, max_q, dq, interp, mult)
    assert dref > 0
    h = torch.zeros(size, dtype=torch.int64, device=dev)
    before = ctx.dropped()
    ctx.ac1d_cq(to_dev(km, dev), to_dev(I, dev), size, max_q, dq, interp, mult, h)
    assert np.array_equal(mesh_np(h), ref)
    assert ctx.dropped() - before == dref

    d = W.lognormal_photon_frames(n * n, 8, 150, 1.0, 2, 600, seed=9)
    d["cnt"] = rng.integers(1, 4, d["cnt"].size).astype(np.uint16)
    off = d["frame_off"]
    ref2 = np.zeros(size)
    dref2 = 0
    for f in range(off.size - 1):
        hl = O.hit_list_from_photons(km, d["pix"][off[f]:off[f + 1]], d["cnt"][off[f]:off[f + 1]])
        ref2, dd = O.ac1d_acuw(hl, size, max_q, dq, interp, 2, ac=ref2)
        dref2 += dd
    assert dref2 > 0
    h2 = torch.zeros(size, dtype=torch.float64, device=dev)
    before = ctx.dropped()
    ctx.ac1d_acuw(to_dev(km, dev), to_dev(d["pix"], dev), to_dev(d["cnt"], dev), to_dev(off, dev), int(off[-1]),
                  size, max_q, dq, interp, h2)
    got2 = h2.cpu().numpy()
    if interp == 0:
        assert np.array_equal(got2, ref2)
    else:
        assert np.allclose(got2, ref2, rtol=1e-12, atol=0)
    assert ctx.dropped() - before == dref2


@pytest.mark.parametrize("interp", [0, 1])
def test_ac1d_acuw_sparse_1024_4096_bins(ctx, dev, interp):
    """AC1D AC_uw at the sparse benchmark's shape: 1024^2 detector, Poisson(0.002) photon
    frames (~2,100 photons each), 4096 bins; the first 16 frames against the oracle's CPU
    branch (AC1D.cpp:256-293): bit for bit with nearest bins (integer weights, exact in
    double), to the double rounding of the add order with linear bins."""
    import torch

    import incohautocorrelate_amd as iac

    n, size, nf = 1024, 4096, 16
    _, km, mq = W.detector(n)
    max_q, dq = iac.ac1d_shape(mq, size)
    d = W.poisson_photon_frames(n * n, nf, 0.002, seed=2)
    off = d["frame_off"]
    ref = np.zeros(size)
    for f in range(nf):
        hl = O.hit_list_from_photons(km, d["pix"][off[f]:off[f + 1]], d["cnt"][off[f]:off[f + 1]])
        ref, _ = O.ac1d_acuw(hl, size, max_q, dq, interp, 2, ac=ref)
    h = torch.zeros(size, dtype=torch.float64, device=dev)
    ctx.ac1d_acuw(to_dev(km, dev), to_dev(d["pix"], dev), to_dev(d["cnt"], dev), to_dev(off, dev),
                  int(off[-1]), size, max_q, dq, interp, h)
    got = h.cpu().numpy()
    assert ref.sum() > 0 and np.count_nonzero(ref) > 1000
    if interp == 0:
        assert np.array_equal(got, ref)
    else:
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-9)
