"""GPU parity of C(q) small, dense frame groups, the rotation merge and the
final AC/C(q) division against the oracle (bit-exact uint64 meshes)."""
import numpy as np
import pytest

import oracle as O
from gpu_helpers import mesh_np, to_dev
from incohautocorrelate_amd import workloads as W

pytestmark = pytest.mark.gpu


def _zeros(n, dev, dtype=None):
    import torch

    return torch.zeros(n, dtype=dtype or torch.int64, device=dev)


@pytest.mark.parametrize("interp", [0, 1])
@pytest.mark.parametrize("n,perp", [(24, 15), (64, 31), (128, 63)])
def test_cq_small_vs_oracle(ctx, dev, interp, n, perp):
    import incohautocorrelate_amd as iac

    _, km, mq = W.detector(n)
    s = iac.mesh_shape_small_cofq(mq, perp)
    rng = np.random.default_rng(n + interp)
    I = rng.gamma(1.0, 0.5, n * n).astype(np.float32)
    I[rng.random(n * n) < 0.1] = 0.0  # masked pixels
    mult = iac.cq_small_multiplicator(I)
    assert mult == O.cq_small_multiplicator(I)
    ref, dref = O.cq_small(I, km, s, interp, mult)
    m = _zeros(ref.size, dev)
    ctx.cq_small(to_dev(km, dev), to_dev(I, dev), s, m, interp, mult)
    assert dref == 0
    assert np.array_equal(mesh_np(m), ref)


def test_cq_small_flat_field_zoomed(ctx, dev):
    import incohautocorrelate_amd as iac

    _, km, mq = W.detector(96)
    s = iac.mesh_shape_small_cofq(mq, 41, 2.0)
    flat = np.ones(96 * 96, np.float32)
    ref, _ = O.cq_small(flat, km, s, 0, 1.0)
    m = _zeros(ref.size, dev)
    ctx.cq_small(to_dev(km, dev), to_dev(flat, dev), s, m, 0, 1.0)
    assert np.array_equal(mesh_np(m), ref)


@pytest.mark.parametrize("interp", [0, 1])
def test_cq_dense_frame_groups(ctx, dev, interp):
    """Frames of one group sum into that group's mesh (orientation groups)."""
    import incohautocorrelate_amd as iac

    n = 48
    _, km, mq = W.detector(n)
    s = iac.mesh_shape_small_cofq(mq, 21)
    frames = W.gamma_frames(n * n, 7, 1.0, 0.5, seed=3)
    goff = np.array([0, 3, 3, 7], np.uint32)  # an empty group in the middle
    mult = 1e9
    meshes = _zeros(3 * s.size_ab ** 2 * s.size_c, dev)
    ctx.cq_dense_frames(to_dev(km, dev), to_dev(frames, dev), goff, s, meshes, interp, mult)
    got = mesh_np(meshes).reshape(3, -1)
    for g in range(3):
        ref = np.zeros(got.shape[1], np.uint64)
        for f in range(goff[g], goff[g + 1]):
            ref, _ = O.cq_small(frames[f], km, s, interp, mult, mesh=ref)
        assert np.array_equal(got[g], ref), g


@pytest.mark.parametrize("interp", [0, 1])
def test_merge_vs_oracle(ctx, dev, interp):
    import torch

    import incohautocorrelate_amd as iac

    n = 64
    _, km, mq = W.detector(n)
    s = iac.mesh_shape_small_cofq(mq, 31)
    rng = np.random.default_rng(5)
    I = rng.gamma(1.0, 0.5, n * n).astype(np.float32)
    mult = O.cq_small_multiplicator(I)
    small_u, _ = O.cq_small(I, km, s, 0, mult)
    small = small_u.astype(np.float64) / mult
    nev = 40
    rots = W.random_rotations(nev, 9)
    mi = rng.gamma(2.0, 0.3, nev).astype(np.float32)
    rw, m2 = iac.merge_prepare(rots, mi, small)
    rw_o, m2_o = O.merge_prepare(rots, mi, small)
    assert m2 == m2_o and np.array_equal(rw, rw_o)
    ref, dref = O.merge_cq(small, s, rw, interp, m2)
    big = _zeros(s.size_ab ** 3, dev)
    ctx.merge_cq(torch.from_numpy(small).to(dev), s, to_dev(rw, dev), interp, m2, big)
    assert np.array_equal(mesh_np(big), ref)
    # the reference writes out-of-cube corners out of bounds; both paths drop them
    assert dref >= 0


def test_u64_to_f64_and_merge_ac_cq(ctx, dev):
    import torch

    rng = np.random.default_rng(2)
    ac = rng.integers(0, 2 ** 40, 5000, dtype=np.uint64)
    cq = rng.normal(1.0, 1.0, 5000)
    cq[::17] = 0.0
    ref = O.merge_ac_cq(ac, cq)
    out = torch.zeros(5000, dtype=torch.float64, device=dev)
    ctx.merge_ac_cq(to_dev(ac, dev), torch.from_numpy(cq).to(dev), out)
    assert np.array_equal(out.cpu().numpy(), ref)
    d = torch.zeros(5000, dtype=torch.float64, device=dev)
    ctx.u64_to_f64(to_dev(ac, dev), 1e6, d)
    assert np.array_equal(d.cpu().numpy(), ac.astype(np.float64) / 1e6)


def test_dense_float_rotation_merge_vs_fp64(ctx, dev):
    """The dense float path at reduced size: Gamma(1, 0.5) frames grouped by
    orientation -> uint64 small C(q) meshes (M = 1e10) -> double -> merge of
    mesh o with orientation o (weight 1) into the lab mesh -> double; checked
    against the FP64 oracle at 1e-6 relative error per populated voxel."""
    import torch

    import incohautocorrelate_amd as iac

    n, nf, norient = 48, 12, 4
    _, km, mq = W.detector(n)
    s = iac.mesh_shape_small_cofq(mq, 21)
    frames = W.gamma_frames(n * n, nf, 1.0, 0.5, seed=12)
    rots = W.random_rotations(norient, 13)
    order = np.argsort(np.arange(nf) % norient, kind="stable")  # frame f -> orientation f mod norient
    frames_sorted = frames[order]
    goff = np.searchsorted(np.sort(np.arange(nf) % norient), np.arange(norient + 1)).astype(np.uint32)
    m1 = 1e10
    nsmall = s.size_ab ** 2 * s.size_c
    small_u = torch.zeros(norient * nsmall, dtype=torch.int64, device=dev)
    ctx.cq_dense_frames(to_dev(km, dev), to_dev(frames_sorted, dev), goff, s, small_u, 0, m1)
    small = torch.zeros(norient * nsmall, dtype=torch.float64, device=dev)
    ctx.u64_to_f64(small_u, m1, small)
    rw = np.zeros((norient, 10), np.float32)
    rw[:, :9] = rots
    rw[:, 9] = 1.0
    m2 = 1e6
    big_u = torch.zeros(s.size_ab ** 3, dtype=torch.int64, device=dev)
    ctx.merge_cq_multi(small, s, to_dev(rw, dev), 0, m2, big_u)
    big = torch.zeros(s.size_ab ** 3, dtype=torch.float64, device=dev)
    ctx.u64_to_f64(big_u, m2, big)
    got = big.cpu().numpy()
    ref = np.zeros(s.size_ab ** 3)
    for o in range(norient):
        sm = np.zeros(nsmall)
        for f in range(nf):
            if f % norient == o:
                sm = O.cq_small_f64(frames[f], km, s, mesh=sm)
        ref = O.merge_f64(sm, s, rw[o], big=ref)
    pop = ref > 1e-3 * ref.max()
    assert pop.sum() > 100
    rel = np.abs(got[pop] - ref[pop]) / ref[pop]
    assert rel.max() <= 1e-6, rel.max()
    assert abs(got.sum() - ref.sum()) <= 1e-6 * ref.sum()


def _cfg3_setup(dev, nframes, norient, seed):
    """The dense-float configuration at its shape: 256^2 detector, Gamma(1, 0.5) frames,
    frame f in orientation f mod norient, small C(q) 129x129x3, lab mesh 129^3."""
    import incohautocorrelate_amd as iac

    n = 256
    _, km, mq = W.detector(n)
    s = iac.mesh_shape_small_cofq(mq, 127)
    frames = W.gamma_frames(n * n, nframes, 1.0, 0.5, seed=seed)
    orient = np.arange(nframes) % norient
    order = np.argsort(orient, kind="stable")
    goff = np.searchsorted(orient[order], np.arange(norient + 1)).astype(np.uint32)
    rw = np.zeros((norient, 10), np.float32)
    rw[:, :9] = W.random_rotations(norient, seed + 1)
    rw[:, 9] = 1.0
    return km, s, frames, orient, order, goff, rw


def _cfg3_gpu(ctx, dev, km, s, frames, order, goff, rw, m1, m2):
    import torch

    norient = len(goff) - 1
    nsmall = s.size_ab ** 2 * s.size_c
    small_u = torch.zeros(norient * nsmall, dtype=torch.int64, device=dev)
    ctx.cq_dense_frames(to_dev(km, dev), to_dev(frames[order], dev), goff, s, small_u, 0, m1)
    small = torch.zeros(norient * nsmall, dtype=torch.float64, device=dev)
    ctx.u64_to_f64(small_u, m1, small)
    big_u = torch.zeros(s.size_ab ** 3, dtype=torch.int64, device=dev)
    ctx.merge_cq_multi(small, s, to_dev(rw, dev), 0, m2, big_u)
    return mesh_np(small_u), small.cpu().numpy(), mesh_np(big_u)


@pytest.mark.timeout(400)
def test_cfg3_shape_reference_multiplicator_bit_exact(ctx, dev):
    """256^2 Gamma(1, 0.5) frames, 2 orientations x 2 frames, with the reference's
    fixed point: the C(q) small Multiplicator of Detector.cpp:1215-1227 (the smallest
    10^k >= 1e-10 with <I>^2 10^k >= 1, times 1e5: 1e6 for these frames) and the merge
    Multiplicator of Detector.cpp:1395-1413 from the merged weights.  The uint64 small
    meshes and the uint64 lab mesh equal the oracle bit for bit."""
    km, s, frames, orient, order, goff, rw = _cfg3_setup(dev, 4, 2, seed=31)
    m1 = O.cq_small_multiplicator(frames.mean(axis=0))
    assert m1 == 1e6
    nsmall = s.size_ab ** 2 * s.size_c
    ref_small = np.zeros((2, nsmall), np.uint64)
    for f in range(4):
        O.cq_small(frames[f], km, s, 0, m1, mesh=ref_small[orient[f]])
    small_d = ref_small.astype(np.float64) / m1
    # the merge Multiplicator from the first orientation's mesh and unit weights
    _, m2 = O.merge_prepare(rw[:, :9], np.ones(2, np.float32), small_d[0])
    got_small, got_small_d, got_big = _cfg3_gpu(ctx, dev, km, s, frames, order, goff, rw, m1, m2)
    assert np.array_equal(got_small.reshape(2, nsmall), ref_small)
    assert np.array_equal(got_small_d.reshape(2, nsmall), small_d)
    ref_big = np.zeros(s.size_ab ** 3, np.uint64)
    for o in range(2):
        ref_big += O.merge_cq(small_d[o], s, rw[o:o + 1], 0, m2)[0]
    assert np.array_equal(got_big, ref_big)
    assert ref_big.sum() > 0


@pytest.mark.timeout(400)
def test_cfg3_shape_fp64_1e6(ctx, dev):
    """The same shape with the finer fixed point (small mesh M = 1e10, merge M = 1e6)
    against the FP64 oracle: every populated lab voxel within 1e-6 relative.  With the
    reference's 1e6 the truncation of each pair value alone moves a voxel by about
    2e-6 (a bias of half a unit in 1e6 per pair on values near 0.25), above that bar."""
    km, s, frames, orient, order, goff, rw = _cfg3_setup(dev, 4, 2, seed=37)
    _, _, got_big = _cfg3_gpu(ctx, dev, km, s, frames, order, goff, rw, 1e10, 1e6)
    got = got_big.astype(np.float64) / 1e6
    nsmall = s.size_ab ** 2 * s.size_c
    ref = np.zeros(s.size_ab ** 3)
    for o in range(2):
        sm = np.zeros(nsmall)
        for f in range(4):
            if orient[f] == o:
                sm = O.cq_small_f64(frames[f], km, s, mesh=sm)
        ref = O.merge_f64(sm, s, rw[o], big=ref)
    pop = ref > 1e-3 * ref.max()
    assert pop.sum() > 1000
    rel = np.abs(got[pop] - ref[pop]) / ref[pop]
    print(f"max relative error over {pop.sum()} populated voxels: {rel.max():.3e}")
    assert rel.max() <= 1e-6, rel.max()


@pytest.mark.parametrize("mode,cell,box,region", [(1, 4, 4096, 0), (1, 8, 4096, 0), (1, 6, 1920, 0), (2, 4, 4096, 0),
                                                (2, 6, 1920, 0), (1, 3, 1920, 2), (2, 3, 1920, 2), (2, 3, 1024, 0),
                                                (2, 5, 4096, 3)])
def test_cq_dense_pair_modes(ctx, dev, mode, cell, box, region):
    """Nearest-neighbour C(q) of dense frames gives the oracle's bits with lanes on
    consecutive pairs and with a pair range per thread, at several cell sizes, box
    classes and task regions: in-box tasks bin inline (mode 2 adding runs of one
    voxel in a register), the others through a checked box whose misses go to
    global memory."""
    import incohautocorrelate_amd as iac

    n = 96
    _, km, mq = W.detector(n)
    s = iac.mesh_shape_small_cofq(mq, 47)
    frames = W.gamma_frames(n * n, 3, 1.0, 0.5, seed=21)
    frames[1, ::7] = 0.0  # zero pixels are dropped by the cell sort
    goff = np.array([0, 3], np.uint32)
    mult = iac.cq_small_multiplicator(frames.mean(axis=0))
    ref = np.zeros(s.size_ab ** 2 * s.size_c, np.uint64)
    for f in range(3):
        ref, _ = O.cq_small(frames[f], km, s, 0, mult, mesh=ref)
    ctx.set_option(iac.OPT_PAIR_MODE, mode)
    ctx.set_option(iac.OPT_CELL_VOX, cell)
    ctx.set_option(iac.OPT_BOX_CAP, box)
    ctx.set_option(iac.OPT_TASK_REGION, region)
    try:
        mesh = _zeros(s.size_ab ** 2 * s.size_c, dev)
        ctx.cq_dense_frames(to_dev(km, dev), to_dev(frames, dev), goff, s, mesh, 0, mult)
        got = mesh_np(mesh)
    finally:
        ctx.set_option(iac.OPT_PAIR_MODE, 0)
        ctx.set_option(iac.OPT_CELL_VOX, 0)
        ctx.set_option(iac.OPT_BOX_CAP, 0)
        ctx.set_option(iac.OPT_TASK_REGION, 0)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("interp", [0, 1])
def test_cq_heavy_cell(ctx, dev, interp):
    """Cells of 64 voxels make the whole 80^2 detector one cell: one item of 20M
    distinct pixel pairs, past 2^32 / group size, so the frame bins item by item.
    Bit-exact against the oracle."""
    import incohautocorrelate_amd as iac

    n = 80
    _, km, mq = W.detector(n)
    s = iac.mesh_shape_small_cofq(mq, 21)
    rng = np.random.default_rng(5)
    I = rng.gamma(1.0, 0.5, n * n).astype(np.float32)
    mult = iac.cq_small_multiplicator(I)
    ref, _ = O.cq_small(I, km, s, interp, mult)
    print(f"C(q) heavy cell: interp {interp}", flush=True)
    ctx.set_option(iac.OPT_CELL_VOX, 64)
    try:
        m = _zeros(ref.size, dev)
        ctx.cq_small(to_dev(km, dev), to_dev(I, dev), s, m, interp, mult)
    finally:
        ctx.set_option(iac.OPT_CELL_VOX, 0)
    assert np.array_equal(mesh_np(m), ref)


@pytest.mark.timeout(400)
def test_cfg3_full_size_halves_and_symmetry(ctx, dev):
    """The dense-float configuration at its full size, where the oracle is too slow:
    2,000 frames of 256^2 over 360 orientations (launches sliced into whole chunks).
    The 360 uint64 C(q) small meshes equal the sum of two passes over the frame halves
    (frames are independent, uint64 sums exact), each is point symmetric, and so is the
    merged lab mesh (the merge rotates integral voxel offsets, which negation commutes
    with)."""
    import torch

    km, s, frames, orient, order, goff, rw = _cfg3_setup(dev, 2000, 360, seed=41)
    nsmall = s.size_ab ** 2 * s.size_c
    print("cfg3 full size: 2,000 frames, 360 orientations", flush=True)
    got_small, _, got_big = _cfg3_gpu(ctx, dev, km, s, frames, order, goff, rw, 1e10, 1e6)
    # halves: frames f < 1000 and f >= 1000, each grouped by orientation
    d_km = to_dev(km, dev)
    acc = torch.zeros(360 * nsmall, dtype=torch.int64, device=dev)
    for half in (np.arange(1000), np.arange(1000, 2000)):
        o = orient[half]
        hord = half[np.argsort(o, kind="stable")]
        hgoff = np.searchsorted(np.sort(o, kind="stable"), np.arange(361)).astype(np.uint32)
        ctx.cq_dense_frames(d_km, to_dev(frames[hord], dev), hgoff, s, acc, 0, 1e10)
    assert np.array_equal(mesh_np(acc), got_small)
    sm = got_small.reshape(360, nsmall)
    assert np.array_equal(sm, sm[:, ::-1])
    assert np.array_equal(got_big, got_big[::-1])
    assert got_big.sum() > 0
