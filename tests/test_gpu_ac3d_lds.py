"""GPU parity of the LDS cell-box pipeline (nearest-neighbour binning, identity
rotation) against the oracle and against the global-atomic tile kernel."""
import numpy as np
import pytest

import oracle as O
from gpu_helpers import gpu_batch, mesh_np
from incohautocorrelate_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture
def lds(ctx):
    import incohautocorrelate_amd as iac

    ctx.set_option(iac.OPT_AC3D_KERNEL, 2)
    yield ctx
    ctx.set_option(iac.OPT_AC3D_KERNEL, 0)
    ctx.set_option(iac.OPT_CELL_VOX, 0)
    ctx.set_option(iac.OPT_FRAME_CHUNK, 0)
    ctx.set_option(iac.OPT_PAIR_PREFETCH, 0)
    ctx.set_option(iac.OPT_BOX_CAP, 0)
    ctx.set_option(iac.OPT_CELL_VOX_V, 0)
    ctx.set_option(iac.OPT_TASK_REGION, 0)
    ctx.set_option(iac.OPT_PAIR_MODE, 0)


def test_lds_cfg1_full(lds, dev):
    import incohautocorrelate_amd as iac

    _, km, mq = W.detector(128)
    shape = iac.mesh_shape_big(mq, 63)
    d = W.dense_poisson_frames(128 * 128, 100, 0.01, seed=1)
    ref, _ = O.ac3d_frames(km, d["pix"], d["cnt"], d["frame_off"], shape)
    got = mesh_np(gpu_batch(lds, dev, km, d["pix"], d["cnt"], d["frame_off"], shape))
    assert np.array_equal(got, ref)


def test_lds_cfg2_subset(lds, dev):
    import incohautocorrelate_amd as iac

    _, km, mq = W.detector(1024)
    shape = iac.mesh_shape_big(mq, 255)
    d = W.poisson_photon_frames(1024 * 1024, 24, 0.002, seed=2)
    ref, _ = O.ac3d_frames(km, d["pix"], d["cnt"], d["frame_off"], shape)
    got = mesh_np(gpu_batch(lds, dev, km, d["pix"], d["cnt"], d["frame_off"], shape))
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("cell_vox,chunk,prefetch,box", [(2, 1, 8, 4096), (3, 7, 4, 4096), (4, 128, 1, 4096),
                                                         (8, 64, 8, 4096), (16, 300, 4, 4096), (6, 64, 1, 4096),
                                                         (5, 32, 1, 1920), (6, 64, 1, 1920), (9, 16, 1, 1920),
                                                         (3, 32, 1, 1024), (4, 8, 1, 1024), (4, 32, 1, 2048),
                                                         (3008, 32, 1, 1920), (12004, 16, 1, 4096)])
@pytest.mark.parametrize("region,mode", [(0, 1), (0, 2), (5, 1)])
def test_lds_tuning_invariance(lds, dev, cell_vox, chunk, prefetch, box, region, mode):
    """Every cell size, frame chunk, pair batch depth, task region and pair walk
    (lanes on consecutive pairs, a range per thread) gives the same bits (the
    tile kernel's)."""
    import incohautocorrelate_amd as iac

    _, km, mq = W.detector(1024)
    shape = iac.mesh_shape_big(mq, 255)
    d = W.poisson_photon_frames(1024 * 1024, 200, 0.002, seed=3)
    lds.set_option(iac.OPT_AC3D_KERNEL, 1)
    tile = mesh_np(gpu_batch(lds, dev, km, d["pix"], d["cnt"], d["frame_off"], shape))
    lds.set_option(iac.OPT_AC3D_KERNEL, 2)
    lds.set_option(iac.OPT_CELL_VOX, cell_vox % 1000)  # u side; v side in the thousands (0 = same)
    lds.set_option(iac.OPT_CELL_VOX_V, cell_vox // 1000)
    lds.set_option(iac.OPT_FRAME_CHUNK, chunk)
    lds.set_option(iac.OPT_PAIR_PREFETCH, prefetch)
    lds.set_option(iac.OPT_BOX_CAP, box)
    lds.set_option(iac.OPT_TASK_REGION, region)
    lds.set_option(iac.OPT_PAIR_MODE, mode)
    got = mesh_np(gpu_batch(lds, dev, km, d["pix"], d["cnt"], d["frame_off"], shape))
    assert np.array_equal(got, tile)


def test_lds_uneven_batch_and_edges(lds, dev):
    """Lognormal frame sizes with empty and one-photon frames, large counts."""
    import incohautocorrelate_amd as iac

    _, km, mq = W.detector(256)
    shape = iac.mesh_shape_big(mq, 127)
    d = W.lognormal_photon_frames(256 * 256, 60, 400, 1.0, 10, 5000, seed=8)
    off = np.sort(np.concatenate([d["frame_off"], [0, d["frame_off"][5], d["frame_off"][9] + 1]]))
    off = off.astype(np.uint64)
    cnt = d["cnt"].copy()
    rng = np.random.default_rng(1)
    cnt[rng.choice(cnt.size, 50, replace=False)] = rng.choice([3, 146, 20000, 65535], 50)
    ref, _ = O.ac3d_frames(km, d["pix"], cnt, off, shape)
    got = mesh_np(gpu_batch(lds, dev, km, d["pix"], cnt, off, shape))
    assert np.array_equal(got, ref)


def test_lds_single_cell_hot_spot(lds, dev):
    """All photons of every frame inside one small patch: one task carries
    all the work and the centre voxel is hammered."""
    import incohautocorrelate_amd as iac

    n = 512
    _, km, mq = W.detector(n)
    shape = iac.mesh_shape_big(mq, 101)
    rng = np.random.default_rng(4)
    pix, off = [], [0]
    for _ in range(30):
        rows = rng.integers(250, 262, 300)
        cols = rng.integers(250, 262, 300)
        p = np.unique(rows * n + cols)
        pix.append(p)
        off.append(off[-1] + p.size)
    pix = np.concatenate(pix).astype(np.uint32)
    cnt = rng.integers(1, 4, pix.size).astype(np.uint16)
    off = np.array(off, np.uint64)
    ref, _ = O.ac3d_frames(km, pix, cnt, off, shape)
    got = mesh_np(gpu_batch(lds, dev, km, pix, cnt, off, shape))
    assert np.array_equal(got, ref)


def test_lds_accumulates_into_existing_mesh(lds, dev):
    import incohautocorrelate_amd as iac

    _, km, mq = W.detector(128)
    shape = iac.mesh_shape_big(mq, 63)
    d = W.dense_poisson_frames(128 * 128, 40, 0.01, seed=6)
    h = d["frame_off"][20]
    first = (d["pix"][:h], d["cnt"][:h], d["frame_off"][:21])
    second = (d["pix"][h:], d["cnt"][h:], d["frame_off"][20:] - h)
    m = gpu_batch(lds, dev, km, *first, shape)
    m = gpu_batch(lds, dev, km, *second, shape, mesh=m)
    ref, _ = O.ac3d_frames(km, d["pix"], d["cnt"], d["frame_off"], shape)
    assert np.array_equal(mesh_np(m), ref)


@pytest.mark.parametrize("box,mode", [(1920, 1), (4096, 1), (1920, 2)])
def test_lds_dense_and_sparse_frames(lds, dev, box, mode):
    """Dense frames (groups of more than 2^24 pairs, items of ~15k pairs) between
    sparse and empty ones, in one batch."""
    import incohautocorrelate_amd as iac

    _, km, mq = W.detector(256)
    shape = iac.mesh_shape_big(mq, 127)
    rng = np.random.default_rng(11)
    sizes = [150, 8000, 40, 0, 6000, 1, 300]
    pix, off = [], [0]
    for n in sizes:
        p = np.sort(rng.choice(256 * 256, n, replace=False)) if n else np.zeros(0, np.int64)
        pix.append(p)
        off.append(off[-1] + n)
    pix = np.concatenate(pix).astype(np.uint32)
    cnt = rng.integers(1, 3, pix.size).astype(np.uint16)
    off = np.array(off, np.uint64)
    ref, _ = O.ac3d_frames(km, pix, cnt, off, shape)
    lds.set_option(iac.OPT_CELL_VOX, 8)
    lds.set_option(iac.OPT_BOX_CAP, box)
    lds.set_option(iac.OPT_PAIR_MODE, mode)
    got = mesh_np(gpu_batch(lds, dev, km, pix, cnt, off, shape))
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("zoom", [0.75, 1.5, 2.0])
def test_lds_zoomed_mesh_cut(lds, dev, zoom):
    """A mesh zoomed in (q_zoom 1.5, 2) cuts through many tasks: tasks beyond the cut
    sphere are dropped, tasks across it test every pair, tasks inside it skip the test;
    zoomed out (0.75) every task lies inside.  All equal the oracle bit for bit, on the
    sparse 1024^2 detector and on dense frames (items of many pairs)."""
    import incohautocorrelate_amd as iac

    _, km, mq = W.detector(1024)
    shape = iac.mesh_shape_big(mq, 127, zoom)
    print(f"zoomed mesh q_zoom={zoom}: sparse 1024^2, 12 frames", flush=True)
    d = W.poisson_photon_frames(1024 * 1024, 12, 0.002, seed=7)
    ref, _ = O.ac3d_frames(km, d["pix"], d["cnt"], d["frame_off"], shape)
    got = mesh_np(gpu_batch(lds, dev, km, d["pix"], d["cnt"], d["frame_off"], shape))
    assert np.array_equal(got, ref)
    _, km2, mq2 = W.detector(128)
    shape2 = iac.mesh_shape_big(mq2, 63, zoom)
    print(f"zoomed mesh q_zoom={zoom}: dense 128^2, 20 frames", flush=True)
    d2 = W.dense_poisson_frames(128 * 128, 20, 0.05, seed=8)
    ref2, _ = O.ac3d_frames(km2, d2["pix"], d2["cnt"], d2["frame_off"], shape2)
    got2 = mesh_np(gpu_batch(lds, dev, km2, d2["pix"], d2["cnt"], d2["frame_off"], shape2))
    assert np.array_equal(got2, ref2)


@pytest.mark.parametrize("box,mode,interp1,chunk", [(1920, 1, 0, 1), (4096, 1, 0, 1), (1920, 2, 0, 0),
                                                    (1920, 1, 1, 1)])
def test_lds_heavy_cell_vs_oracle(lds, dev, box, mode, interp1, chunk):
    """Cells of 64 voxels on a 31-voxel mesh hold the whole 96^2 detector: a frame of
    5,000 photons is one item of 12.5M distinct (25M ordered) pairs, at or past
    2^32 / group size, so its chunk bins item by item; with one-frame chunks the
    other frames still walk.  Bit-exact against the oracle, nearest and linear."""
    import incohautocorrelate_amd as iac

    _, km, mq = W.detector(96)
    shape = iac.mesh_shape_big(mq, 31)
    rng = np.random.default_rng(21)
    sizes = [300, 5000, 40, 0, 700]
    pix, off = [], [0]
    for n in sizes:
        p = np.sort(rng.choice(96 * 96, n, replace=False)) if n else np.zeros(0, np.int64)
        pix.append(p)
        off.append(off[-1] + n)
    pix = np.concatenate(pix).astype(np.uint32)
    cnt = rng.integers(1, 3, pix.size).astype(np.uint16)
    off = np.array(off, np.uint64)
    print(f"heavy cell: box {box} mode {mode} interp1 {interp1} chunk {chunk}", flush=True)
    ref, _ = O.ac3d_frames(km, pix, cnt, off, shape, interp1=interp1)
    lds.set_option(iac.OPT_CELL_VOX, 64)
    lds.set_option(iac.OPT_BOX_CAP, box)
    lds.set_option(iac.OPT_PAIR_MODE, mode)
    lds.set_option(iac.OPT_FRAME_CHUNK, chunk)
    got = mesh_np(gpu_batch(lds, dev, km, pix, cnt, off, shape, interp1=interp1))
    assert np.array_equal(got, ref)


def test_lds_heavy_group_past_2_32(lds, dev):
    """One frame of 95,000 unit photons on three k points, all in one cell: the cell's
    4.5e9 distinct pairs pass 2^32 inside one group.  The mesh is a quadratic form in
    the photon counts c_i per point, sum c_i E_i + sum_{i<j} c_i c_j P_ij +
    sum C(c_i, 2) Q_i, whose terms the oracle gives from frames of one or two photons."""
    import incohautocorrelate_amd as iac

    _, km, mq = W.detector(96)
    shape = iac.mesh_shape_big(mq, 31)
    pts = km.reshape(-1, 3)[[48 * 96 + 48, 48 * 96 + 51, 52 * 96 + 47]]
    counts = [93500, 1000, 500]
    kmap = np.repeat(pts, counts, axis=0).astype(np.float32)
    n = kmap.shape[0]

    def oracle(classes):
        m, dropped = O.ac3d_frames(pts, np.asarray(classes, np.uint32), np.ones(len(classes), np.uint16),
                                   np.array([0, len(classes)], np.uint64), shape)
        assert dropped == 0
        return m

    E = [oracle([i]) for i in range(3)]
    want = np.zeros_like(E[0])
    for i in range(3):
        c = np.uint64(counts[i])
        want += c * E[i] + (c * (c - np.uint64(1)) // np.uint64(2)) * (oracle([i, i]) - np.uint64(2) * E[i])
        for j in range(i + 1, 3):
            want += c * np.uint64(counts[j]) * (oracle([i, j]) - E[i] - E[j])
    lds.set_option(iac.OPT_CELL_VOX, 64)
    print(f"heavy group: one frame of {n} photons on 3 points", flush=True)
    got = mesh_np(gpu_batch(lds, dev, kmap, np.arange(n, dtype=np.uint32), np.ones(n, np.uint16),
                            np.array([0, n], np.uint64), shape))
    assert want[want.size // 2] > 0
    assert np.array_equal(got, want)


@pytest.mark.parametrize("interp1", [0, 1])
def test_lds_heavy_cell_rotated(lds, dev, interp1):
    """Per-frame rotations on chunks that bin item by item: a 5,000-photon frame in one
    64-voxel cell between smaller frames, each frame its own rotation (rotated at the box
    flush and on the way out of the box).  Bit-exact against the oracle, nothing dropped."""
    import incohautocorrelate_amd as iac

    _, km, mq = W.detector(96)
    shape = iac.mesh_shape_big(mq, 31)
    rng = np.random.default_rng(23)
    sizes = [200, 5000, 1, 600]
    pix, off = [], [0]
    for n in sizes:
        pix.append(np.sort(rng.choice(96 * 96, n, replace=False)))
        off.append(off[-1] + n)
    pix = np.concatenate(pix).astype(np.uint32)
    cnt = rng.integers(1, 3, pix.size).astype(np.uint16)
    off = np.array(off, np.uint64)
    rots = W.random_rotations(len(sizes), 29).astype(np.float32)
    print(f"heavy cell rotated: interp1 {interp1}", flush=True)
    ref, dref = O.ac3d_frames(km, pix, cnt, off, shape, rots=rots, interp1=interp1)
    assert dref == 0
    lds.set_option(iac.OPT_CELL_VOX, 64)
    before = lds.dropped()
    got = mesh_np(gpu_batch(lds, dev, km, pix, cnt, off, shape, rots=rots, interp1=interp1))
    assert lds.dropped() == before
    assert np.array_equal(got, ref)
