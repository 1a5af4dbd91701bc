"""The frame queue (iac_ac3d_queue_*): Detector::AutoCorrelateSparseList once per event
(RunIAC.cpp:524-550) binned in batches on the LDS engine.  Every case equals the CPU
oracle bit for bit: host and device pushes, flushes forced by a small queue, empty frames,
per-frame rotations (on the engine, and on the tile kernel where a rotation could leave
the engine's guarantee), every interpolation mode, and the sparse 1024^2 detector."""
import numpy as np
import pytest

import oracle as O
from gpu_helpers import mesh_np, to_dev
from incohautocorrelate_amd import workloads as W

pytestmark = pytest.mark.gpu


def _frames(n, nframes, seed, empty=(3,)):
    _, km, mq = W.detector(n)
    data = W.lognormal_photon_frames(n * n, nframes, 50, 1.0, 2, 400, seed=seed)
    off = data["frame_off"]
    hls = []
    for f in range(nframes):
        a, b = int(off[f]), int(off[f + 1])
        hl = O.hit_list_from_photons(km, data["pix"][a:b], data["cnt"][a:b])
        hls.append(hl[:0] if f in empty else hl)
    return km, mq, hls


def _oracle(hls, shape, rots, i1, i2, dm):
    mesh = np.zeros(shape.size_ab ** 2 * shape.size_c, np.uint64)
    for f, hl in enumerate(hls):
        if len(hl):
            O.ac3d_sparse(hl, shape, None if rots is None else rots[f], i1, i2, dm, mesh=mesh)
    return mesh


@pytest.mark.parametrize("i1,i2,dm", [(0, 0, True), (1, 1, True), (0, 1, True), (1, 0, False)])
def test_queue_host_pushes_equal_oracle(ctx, dev, i1, i2, dm):
    import torch

    import incohautocorrelate_amd as iac

    km, mq, hls = _frames(96, 30, seed=3 + i1 + 2 * i2)
    shape = iac.mesh_shape_big(mq, 47)
    m = torch.zeros(shape.size_ab ** 2 * shape.size_c, dtype=torch.int64, device=dev)
    # a small queue: flushes on the frame limit (7) and on the record limit (2,500)
    q = ctx.ac3d_queue(to_dev(km, dev), shape, m, i1, i2, dm, max_records=2500, max_frames=7)
    for hl in hls:
        q.push(hl)
    q.destroy()
    assert np.array_equal(mesh_np(m), _oracle(hls, shape, None, i1, i2, dm))


def test_queue_rotations_engine_and_tiles(ctx, dev):
    """Per-frame rotations: nearest-neighbour second level on the engine (rotated at the
    box flush); a linear second level takes the tile kernel for that flush."""
    import torch

    import incohautocorrelate_amd as iac

    km, mq, hls = _frames(96, 24, seed=11, empty=())
    rots = W.random_rotations(len(hls), seed=4).reshape(-1, 9).astype(np.float32)
    for i2 in (0, 1):
        shape = iac.mesh_shape_big(mq, 47)
        m = torch.zeros(shape.size_ab ** 2 * shape.size_c, dtype=torch.int64, device=dev)
        q = ctx.ac3d_queue(to_dev(km, dev), shape, m, 0, i2, True, max_frames=10)
        for hl, r in zip(hls, rots):
            q.push(hl, r)
        q.flush()
        assert q.pending() == (0, 0)
        q.destroy()
        assert np.array_equal(mesh_np(m), _oracle(hls, shape, rots, 0, i2, True))


def test_queue_device_pushes_and_source_switch(ctx, dev):
    import torch

    import incohautocorrelate_amd as iac

    km, mq, hls = _frames(80, 16, seed=21)
    shape = iac.mesh_shape_big(mq, 41)
    m = torch.zeros(shape.size_ab ** 2 * shape.size_c, dtype=torch.int64, device=dev)
    q = ctx.ac3d_queue(to_dev(km, dev), shape, m)
    for f, hl in enumerate(hls):
        if f % 5 < 3:  # device records, host records: each switch flushes the pending frames
            q.push_device(to_dev(hl, dev) if len(hl) else None)
        else:
            q.push(hl)
    # the last switch (host frames 13, 14 -> device frame 15) flushed all but frame 15
    assert q.pending() == (1, len(hls[15]))
    q.destroy()
    assert np.array_equal(mesh_np(m), _oracle(hls, shape, None, 0, 0, True))


def test_queue_limits(ctx, dev):
    import torch

    import incohautocorrelate_amd as iac

    km, mq, hls = _frames(64, 4, seed=1, empty=())
    shape = iac.mesh_shape_big(mq, 31)
    m = torch.zeros(shape.size_ab ** 2 * shape.size_c, dtype=torch.int64, device=dev)
    big = max(hls, key=len)
    q = ctx.ac3d_queue(to_dev(km, dev), shape, m, max_records=len(big) - 1, max_frames=2)
    with pytest.raises(iac.IACError):
        q.push(big)
    q.push(big[: len(big) - 1])
    assert q.pending() == (1, len(big) - 1)
    q.destroy()
    ref = _oracle([big[: len(big) - 1]], shape, None, 0, 0, True)
    assert np.array_equal(mesh_np(m), ref)


def test_queue_sparse_1024(ctx, dev):
    """The sparse benchmark detector (1024^2, Poisson 0.002, 257^3 mesh): 24 frames pushed
    one by one from the host equal the oracle."""
    import torch

    import incohautocorrelate_amd as iac

    _, km, mq = W.detector(1024)
    shape = iac.mesh_shape_big(mq, 255)
    data = W.poisson_photon_frames(1024 * 1024, 24, 0.002, seed=2)
    off = data["frame_off"]
    m = torch.zeros(shape.size_ab ** 2 * shape.size_c, dtype=torch.int64, device=dev)
    q = ctx.ac3d_queue(to_dev(km, dev), shape, m, max_frames=10)
    for f in range(24):
        a, b = int(off[f]), int(off[f + 1])
        q.push(O.hit_list_from_photons(km, data["pix"][a:b], data["cnt"][a:b]))
    q.destroy()
    ref, _ = O.ac3d_frames(km, data["pix"], data["cnt"], off, shape)
    assert np.array_equal(mesh_np(m), ref)
