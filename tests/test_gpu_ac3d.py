"""GPU parity of the 3D AC_uw pair accumulation against the CPU oracle.

Integer meshes must be bit-identical (uint64 accumulation is order free, so
this holds iff every pair's voxel and fixed-point value match the reference
CPU path of Detector::AutoCorrelateSparseList / ACMesh::Atomic_Add_q_Entry).
"""
import numpy as np
import pytest

import oracle as O
from gpu_helpers import gpu_batch, gpu_single, mesh_np
from incohautocorrelate_amd import workloads as W

pytestmark = pytest.mark.gpu

MODES = [(i1, i2, dm) for i1 in (0, 1) for i2 in (0, 1) for dm in (True, False)]


def _rand_hitlist(rng, km, n, cmax=5, floats=False):
    pix = np.sort(rng.choice(km.shape[0], n, replace=False))
    if floats:
        vals = rng.gamma(1.0, 0.5, n).astype(np.float32)
    else:
        vals = rng.integers(1, cmax + 1, n).astype(np.float32)
    hl = np.empty((n, 4), np.float32)
    hl[:, :3] = km[pix]
    hl[:, 3] = vals
    return hl


@pytest.mark.parametrize("i1,i2,dm", MODES)
def test_single_frame_modes(ctx, dev, i1, i2, dm):
    import incohautocorrelate_amd as iac

    _, km, mq = O.detector(24)
    shape = iac.mesh_shape_big(mq, 17)
    rng = np.random.default_rng(100 + 4 * i1 + 2 * i2 + int(dm))
    rots = [None] + list(W.random_rotations(3, 7))
    for trial, R in enumerate(rots):
        hl = _rand_hitlist(rng, km, int(rng.integers(1, 150)), floats=(trial == 2))
        ref, dref = O.ac3d_sparse(hl, shape, R, i1, i2, dm)
        before = ctx.dropped()
        got = mesh_np(gpu_single(ctx, dev, hl, shape, R, i1, i2, dm))
        assert np.array_equal(got, ref), f"mode {(i1, i2, dm)} trial {trial}"
        assert ctx.dropped() - before == dref


def test_cfg1_plain_small_full(ctx, dev):
    """128^2 detector, 100 frames Poisson(0.01), 65^3 mesh: every frame, bit exact."""
    import incohautocorrelate_amd as iac
    from incohautocorrelate_amd import workloads as W

    _, km, mq = W.detector(128)
    shape = iac.mesh_shape_big(mq, 63)
    d = W.dense_poisson_frames(128 * 128, 100, 0.01, seed=1)
    ref, dref = O.ac3d_frames(km, d["pix"], d["cnt"], d["frame_off"], shape)
    got = mesh_np(gpu_batch(ctx, dev, km, d["pix"], d["cnt"], d["frame_off"], shape))
    assert dref == 0
    assert np.array_equal(got, ref)
    assert got.sum() > 0


def test_cfg2_sparse_subset(ctx, dev):
    """1024^2 detector, Poisson(0.002), 257^3 mesh: the first 24 frames, bit exact."""
    import incohautocorrelate_amd as iac
    from incohautocorrelate_amd import workloads as W

    _, km, mq = W.detector(1024)
    shape = iac.mesh_shape_big(mq, 255)
    d = W.poisson_photon_frames(1024 * 1024, 24, 0.002, seed=2)
    ref, _ = O.ac3d_frames(km, d["pix"], d["cnt"], d["frame_off"], shape)
    got = mesh_np(gpu_batch(ctx, dev, km, d["pix"], d["cnt"], d["frame_off"], shape))
    assert np.array_equal(got, ref)


def test_cfg2_full_size_engine_tiles_halves(ctx, dev):
    """The sparse configuration at its full size (10,000 frames of the 1024^2 detector,
    Poisson(0.002), 257^3 mesh), where the oracle is too slow: the LDS engine's mesh equals
    the global-atomic tile kernel's (an independent path, pair by pair), equals the sum of
    two half-batch passes (frames are independent), is point symmetric about the centre,
    and its centre holds at least 1000 sum c^2 (every photon's self pair).  Nothing is
    dropped."""
    import torch

    import incohautocorrelate_amd as iac
    from incohautocorrelate_amd import workloads as W

    _, km, mq = W.detector(1024)
    shape = iac.mesh_shape_big(mq, 255)
    d = W.poisson_photon_frames(1024 * 1024, 10000, 0.002, seed=2)
    pix, cnt, off = d["pix"], d["cnt"], d["frame_off"]
    print(f"cfg2 full size: {int(off[-1])} photons in 10,000 frames", flush=True)
    dropped0 = ctx.dropped()
    full = mesh_np(gpu_batch(ctx, dev, km, pix, cnt, off, shape))
    h = int(off[5000])
    m = gpu_batch(ctx, dev, km, pix[:h], cnt[:h], off[:5001], shape)
    m = gpu_batch(ctx, dev, km, pix[h:], cnt[h:], off[5000:] - np.uint64(h), shape, mesh=m)
    assert np.array_equal(mesh_np(m), full)
    ctx.set_option(iac.OPT_AC3D_KERNEL, 1)
    try:
        tiles = mesh_np(gpu_batch(ctx, dev, km, pix, cnt, off, shape))
    finally:
        ctx.set_option(iac.OPT_AC3D_KERNEL, 0)
    assert np.array_equal(tiles, full)
    assert ctx.dropped() == dropped0
    assert np.array_equal(full, full[::-1])  # point symmetry: voxel v and nvox - 1 - v
    c = cnt.astype(np.uint64)
    assert full[full.size // 2] >= np.uint64(1000) * int((c * c).sum())


def test_batch_equals_per_frame_uneven(ctx, dev):
    """Uneven frames (lognormal sizes, empty and one-photon frames) through the
    persistent batch equal per-frame calls and the oracle."""
    import torch

    import incohautocorrelate_amd as iac
    from incohautocorrelate_amd import workloads as W

    _, km, mq = W.detector(256)
    shape = iac.mesh_shape_big(mq, 127)
    d = W.lognormal_photon_frames(256 * 256, 40, 300, 1.0, 10, 3000, seed=5)
    pix, cnt, off = d["pix"], d["cnt"], d["frame_off"]
    # an empty frame (repeated boundary) and a one-photon frame (split boundary)
    off = np.sort(np.concatenate([off, [off[1], off[2] + 1]])).astype(np.uint64)
    rots = W.random_rotations(off.size - 1, 3)
    batch = mesh_np(gpu_batch(ctx, dev, km, pix, cnt, off, shape, rots=rots))
    m = torch.zeros(shape.size_ab ** 3, dtype=torch.int64, device=dev)
    for f in range(off.size - 1):
        hl = O.hit_list_from_photons(km, pix[off[f]:off[f + 1]], cnt[off[f]:off[f + 1]])
        gpu_single(ctx, dev, hl, shape, rots[f], mesh=m)
    per = mesh_np(m)
    assert np.array_equal(batch, per)
    ref, _ = O.ac3d_frames(km, pix, cnt, off, shape, rots=rots)
    assert np.array_equal(batch, ref)


def test_large_counts_fixed_point(ctx, dev):
    """Products beyond 21,054 leave the 1000*p fast path; products beyond ~4.29e6
    wrap in the reference's unsigned int (IAC/ACMesh.cpp:390-391)."""
    import incohautocorrelate_amd as iac

    _, km, mq = O.detector(32)
    shape = iac.mesh_shape_big(mq, 21)
    rng = np.random.default_rng(9)
    hl = _rand_hitlist(rng, km, 60)
    hl[:, 3] = rng.choice([1, 7, 145, 146, 3000, 20000, 65535], size=60).astype(np.float32)
    for i2 in (0, 1):
        ref, _ = O.ac3d_sparse(hl, shape, None, 0, i2, True)
        got = mesh_np(gpu_single(ctx, dev, hl, shape, None, 0, i2, True))
        assert np.array_equal(got, ref)


def test_maxq_cut_boundary(ctx, dev):
    """Pairs whose |q| straddles Max_Q (diagonal corners of a full detector are
    cut: the cut is Euclidean while Max_Q is the largest axis extent)."""
    import incohautocorrelate_amd as iac

    _, km, mq = O.detector(64)
    shape = iac.mesh_shape_big(mq, 31)
    idx = [0, 63, 64 * 63, 64 * 64 - 1, 32 * 64 + 32, 5, 64 * 5, 64 * 58 + 63, 64 * 63 + 58]
    hl = np.empty((len(idx), 4), np.float32)
    hl[:, :3] = km[idx]
    hl[:, 3] = 1
    ref, _ = O.ac3d_sparse(hl, shape)
    got = mesh_np(gpu_single(ctx, dev, hl, shape))
    assert np.array_equal(got, ref)
    # the far corners are cut: fewer than all ordered pairs land in the mesh
    assert ref.sum() // 1000 < len(idx) ** 2


def test_empty_and_single(ctx, dev):
    import incohautocorrelate_amd as iac

    _, km, mq = O.detector(16)
    shape = iac.mesh_shape_big(mq, 15)
    assert mesh_np(gpu_single(ctx, dev, np.zeros((0, 4), np.float32), shape)).sum() == 0
    hl = np.array([[*km[17], 3.0]], np.float32)
    got = mesh_np(gpu_single(ctx, dev, hl, shape))
    c = shape.center[0]
    assert got[c + c * shape.size_ab + c * shape.size_ab ** 2] == 9000
    assert got.sum() == 9000


@pytest.mark.parametrize("i1", [0, 1])
def test_batch_rotations_on_engine(ctx, dev, i1):
    """Per-frame rotations through the LDS engine (double map, nearest second level,
    nearest or linear first level): frames bin under the identity and each box voxel
    is rotated when it flushes.  A run of one repeated rotation (chunks over several
    frames), identity frames, a distinct rotation per frame, an empty and a one-photon
    frame: the oracle's bits, the tile kernel's bits, nothing dropped."""
    import torch

    import incohautocorrelate_amd as iac
    from incohautocorrelate_amd import workloads as W

    _, km, mq = W.detector(256)
    shape = iac.mesh_shape_big(mq, 127)
    d = W.lognormal_photon_frames(256 * 256, 48, 300, 1.0, 10, 3000, seed=8)
    pix, cnt, off = d["pix"], d["cnt"], d["frame_off"]
    off = np.sort(np.concatenate([off, [off[20], off[21] + 1]])).astype(np.uint64)
    nf = off.size - 1
    rots = W.random_rotations(nf, 17).astype(np.float32)
    rots[:10] = rots[0]
    rots[10:15] = np.eye(3, dtype=np.float32).reshape(9)
    ref, dref = O.ac3d_frames(km, pix, cnt, off, shape, rots=rots, interp1=i1)
    assert dref == 0
    got = {}
    try:
        for kernel in (2, 1):
            ctx.set_option(iac.OPT_AC3D_KERNEL, kernel)
            before = ctx.dropped()
            got[kernel] = mesh_np(gpu_batch(ctx, dev, km, pix, cnt, off, shape, rots=rots, interp1=i1))
            assert ctx.dropped() == before
    finally:
        ctx.set_option(iac.OPT_AC3D_KERNEL, 0)
    assert np.array_equal(got[2], ref)
    assert np.array_equal(got[1], ref)


@pytest.mark.parametrize("i2,dm", [(1, True), (0, False)])
def test_batch_rotations_engine_refuses(ctx, dev, i2, dm):
    """Per-frame rotations whose voxel is not a function of the first-level voxel (a
    linear second level, or no double map) stay on the tile kernel: forcing the LDS
    engine is a precondition error, and the automatic choice matches the oracle."""
    import incohautocorrelate_amd as iac
    from incohautocorrelate_amd import workloads as W

    _, km, mq = W.detector(64)
    shape = iac.mesh_shape_big(mq, 31)
    d = W.lognormal_photon_frames(64 * 64, 6, 40, 1.0, 2, 200, seed=3)
    rots = W.random_rotations(6, 5)
    ctx.set_option(iac.OPT_AC3D_KERNEL, 2)
    try:
        with pytest.raises(iac.IACError):
            gpu_batch(ctx, dev, km, d["pix"], d["cnt"], d["frame_off"], shape, rots=rots, interp2=i2,
                      double_map=dm)
    finally:
        ctx.set_option(iac.OPT_AC3D_KERNEL, 0)
    got = mesh_np(gpu_batch(ctx, dev, km, d["pix"], d["cnt"], d["frame_off"], shape, rots=rots, interp2=i2,
                            double_map=dm))
    ref, _ = O.ac3d_frames(km, d["pix"], d["cnt"], d["frame_off"], shape, rots=rots, interp2=i2, double_map=dm)
    assert np.array_equal(got, ref)
