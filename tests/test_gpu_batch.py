"""Prepared frame batches (iac_ac3d_batch_*): one launch or a replayed graph equals
the unbatched per-frame GPU kernel, the one-call batch and the CPU oracle,
bit for bit -- including the uneven lognormal workload at full size."""
import numpy as np
import pytest

import oracle as O
from gpu_helpers import mesh_np, to_dev
from incohautocorrelate_amd import workloads as W

pytestmark = pytest.mark.gpu


def _dev_frames(data, dev):
    return (to_dev(data["pix"], dev), to_dev(data["cnt"], dev), to_dev(data["frame_off"], dev))


def _per_frame(ctx, dev, km, data, shape, frames):
    """The unbatched GPU path: iac_ac3d_sparse once per frame (tile kernel)."""
    import torch

    m = torch.zeros(shape.size_ab ** 2 * shape.size_c, dtype=torch.int64, device=dev)
    off = data["frame_off"]
    eye = np.eye(3, dtype=np.float32).reshape(9)
    for f in frames:
        a, b = int(off[f]), int(off[f + 1])
        if a == b:
            continue
        hl = O.hit_list_from_photons(km, data["pix"][a:b], data["cnt"][a:b])
        ctx.ac3d_sparse(to_dev(hl, dev), eye, shape, m)
    return mesh_np(m)


@pytest.mark.parametrize("interp1,interp2,dm", [(0, 0, True), (1, 1, True), (0, 1, False)])
def test_batch_equals_per_frame_and_oracle(ctx, dev, interp1, interp2, dm):
    import torch

    import incohautocorrelate_amd as iac

    n = 128
    _, km, mq = W.detector(n)
    shape = iac.mesh_shape_big(mq, 63)
    data = W.lognormal_photon_frames(n * n, 40, 60, 1.0, 2, 600, seed=5 + interp1)
    n_total = int(data["frame_off"][-1])
    d_km = to_dev(km, dev)
    pix, cnt, off = _dev_frames(data, dev)
    b = ctx.ac3d_batch(d_km, pix, cnt, off, n_total, shape, interp1, interp2, dm)
    m = torch.zeros(shape.size_ab ** 2 * shape.size_c, dtype=torch.int64, device=dev)
    b.launch(m, zero_mesh=True)
    got = mesh_np(m)
    ref, _ = O.ac3d_frames(km, data["pix"], data["cnt"], data["frame_off"], shape, None, interp1, interp2, dm)
    assert np.array_equal(got, ref)
    if (interp1, interp2, dm) == (0, 0, True):
        assert np.array_equal(got, _per_frame(ctx, dev, km, data, shape, range(40)))
    b.destroy()


def test_graph_replays_match_eager(ctx, dev):
    """Replays of the captured graph, with eager work and syncs in between, each
    equal the eager launch (the mesh is cleared inside the graph by a kernel)."""
    import torch

    import incohautocorrelate_amd as iac

    n = 192
    _, km, mq = W.detector(n)
    shape = iac.mesh_shape_big(mq, 95)
    data = W.lognormal_photon_frames(n * n, 64, 150, 1.0, 5, 1500, seed=8)
    n_total = int(data["frame_off"][-1])
    d_km = to_dev(km, dev)
    pix, cnt, off = _dev_frames(data, dev)
    b = ctx.ac3d_batch(d_km, pix, cnt, off, n_total, shape)
    eager = torch.zeros(shape.size_ab ** 3, dtype=torch.int64, device=dev)
    b.launch(eager, zero_mesh=True)
    want = mesh_np(eager)
    m = torch.zeros_like(eager)
    b.capture(m, zero_mesh=True)
    for r in range(4):
        m.fill_(12345)  # stale content the graph must clear
        b.launch(m, zero_mesh=True)
        torch.cuda.synchronize()
        assert np.array_equal(mesh_np(m), want), f"replay {r}"
        # eager work between replays
        other = torch.zeros_like(eager)
        ctx.ac3d_sparse_batch(d_km, pix, cnt, off, n_total, shape, other)
        torch.cuda.synchronize()
        assert np.array_equal(mesh_np(other), want)
    # new photon counts in the same buffers: the graph reads the current contents
    cnt2 = np.minimum(data["cnt"].astype(np.int64) * 3, 65535).astype(np.uint16)
    cnt.copy_(to_dev(cnt2, dev))
    b.launch(m, zero_mesh=True)
    ref, _ = O.ac3d_frames(km, data["pix"], cnt2, data["frame_off"], shape)
    assert np.array_equal(mesh_np(m), ref)
    b.destroy()


def test_accumulating_launches(ctx, dev):
    import torch

    import incohautocorrelate_amd as iac

    n = 96
    _, km, mq = W.detector(n)
    shape = iac.mesh_shape_big(mq, 47)
    data = W.poisson_photon_frames(n * n, 30, 0.01, 3)
    n_total = int(data["frame_off"][-1])
    pix, cnt, off = _dev_frames(data, dev)
    b = ctx.ac3d_batch(to_dev(km, dev), pix, cnt, off, n_total, shape)
    m = torch.zeros(shape.size_ab ** 3, dtype=torch.int64, device=dev)
    for _ in range(3):
        b.launch(m, zero_mesh=False)
    ref, _ = O.ac3d_frames(km, data["pix"], data["cnt"], data["frame_off"], shape)
    assert np.array_equal(mesh_np(m), ref * np.uint64(3))
    b.destroy()


def test_uneven_lognormal_full_size(ctx, dev):
    """4,096 frames on a 1024^2 detector, photons/frame lognormal (median 1,500,
    sigma 1, clipped to 10..20,000): the graph-replayed batch equals the
    global-atomic tile kernel over all frames, the per-frame kernel on a frame
    subset and the CPU oracle on a smaller subset."""
    import torch

    import incohautocorrelate_amd as iac

    n = 1024
    _, km, mq = W.detector(n)
    shape = iac.mesh_shape_big(mq, 255)
    data = W.lognormal_photon_frames(n * n, 4096, 1500, 1.0, 10, 20000, seed=0)
    n_total = int(data["frame_off"][-1])
    print(f"uneven workload: {n_total} photons, max frame {np.diff(data['frame_off'].astype(np.int64)).max()}")
    d_km = to_dev(km, dev)
    pix, cnt, off = _dev_frames(data, dev)
    nvox = shape.size_ab ** 2 * shape.size_c
    dropped0 = ctx.dropped()  # the context's counter is cumulative over the session
    b = ctx.ac3d_batch(d_km, pix, cnt, off, n_total, shape)
    m = torch.zeros(nvox, dtype=torch.int64, device=dev)
    b.capture(m, zero_mesh=True)
    b.launch(m, zero_mesh=True)
    got = mesh_np(m)
    tiles = torch.zeros(nvox, dtype=torch.int64, device=dev)
    ctx.set_option(iac.OPT_AC3D_KERNEL, 1)
    try:
        ctx.ac3d_sparse_batch(d_km, pix, cnt, off, n_total, shape, tiles)
    finally:
        ctx.set_option(iac.OPT_AC3D_KERNEL, 0)
    assert np.array_equal(got, mesh_np(tiles))
    assert got.sum() > 0 and ctx.dropped() == dropped0
    b.destroy()
    # subsets: per-frame GPU kernel (24 frames incl. the largest) and the CPU oracle (6 frames)
    sizes = np.diff(data["frame_off"].astype(np.int64))
    sub = sorted(set(range(16)) | set(np.argsort(sizes)[-8:].tolist()))
    sub_off = np.zeros(len(sub) + 1, np.uint64)
    sub_pix, sub_cnt = [], []
    for i, f in enumerate(sub):
        a, c = int(data["frame_off"][f]), int(data["frame_off"][f + 1])
        sub_pix.append(data["pix"][a:c])
        sub_cnt.append(data["cnt"][a:c])
        sub_off[i + 1] = sub_off[i] + (c - a)
    sd = {"pix": np.concatenate(sub_pix), "cnt": np.concatenate(sub_cnt), "frame_off": sub_off}
    spix, scnt, soff = _dev_frames(sd, dev)
    bs = ctx.ac3d_batch(d_km, spix, scnt, soff, int(sub_off[-1]), shape)
    ms = torch.zeros(nvox, dtype=torch.int64, device=dev)
    bs.launch(ms, zero_mesh=True)
    assert np.array_equal(mesh_np(ms), _per_frame(ctx, dev, km, data, shape, sub))
    bs.destroy()
    six = {"pix": sd["pix"][: int(sub_off[6])], "cnt": sd["cnt"][: int(sub_off[6])], "frame_off": sub_off[:7]}
    ref, _ = O.ac3d_frames(km, six["pix"], six["cnt"], six["frame_off"], shape)
    p6, c6, o6 = _dev_frames(six, dev)
    b6 = ctx.ac3d_batch(d_km, p6, c6, o6, int(sub_off[6]), shape)
    m6 = torch.zeros(nvox, dtype=torch.int64, device=dev)
    b6.launch(m6, zero_mesh=True)
    assert np.array_equal(mesh_np(m6), ref)
    b6.destroy()


@pytest.mark.parametrize("mode,cell,box", [(1, 5, 4096), (1, 4, 1920), (2, 4, 4096), (1, 3, 1024)])
def test_linear_modes_pair_walks(ctx, dev, mode, cell, box):
    """Linear AC_uw (ordered pairs, 8 or 64 corner adds each; a linear first level with a
    nearest second level bins in-box tasks inline) through the engine with lanes on
    consecutive pairs and with a pair range per thread: the oracle's bits."""
    import torch

    import incohautocorrelate_amd as iac

    n = 96
    _, km, mq = W.detector(n)
    shape = iac.mesh_shape_big(mq, 47)
    data = W.lognormal_photon_frames(n * n, 12, 50, 1.0, 1, 400, seed=31)
    n_total = int(data["frame_off"][-1])
    ref = {}
    for i1, i2 in ((1, 1), (0, 1), (1, 0)):
        ref[i1, i2], _ = O.ac3d_frames(km, data["pix"], data["cnt"], data["frame_off"], shape, None, i1, i2, True)
    ctx.set_option(iac.OPT_AC3D_KERNEL, 2)
    ctx.set_option(iac.OPT_PAIR_MODE, mode)
    ctx.set_option(iac.OPT_CELL_VOX, cell)
    ctx.set_option(iac.OPT_BOX_CAP, box)
    try:
        pix, cnt, off = _dev_frames(data, dev)
        for (i1, i2), r in ref.items():
            m = torch.zeros(shape.size_ab ** 2 * shape.size_c, dtype=torch.int64, device=dev)
            ctx.ac3d_sparse_batch(to_dev(km, dev), pix, cnt, off, n_total, shape, m, interp1=i1, interp2=i2)
            assert np.array_equal(mesh_np(m), r), (i1, i2)
    finally:
        for o in (iac.OPT_AC3D_KERNEL, iac.OPT_PAIR_MODE, iac.OPT_CELL_VOX, iac.OPT_BOX_CAP):
            ctx.set_option(o, 0)


@pytest.mark.parametrize("mode", [1, 2])
def test_nn_linear_double_map_symmetric(ctx, dev, mode):
    """Double map, nearest first level, linear second level, identity rotation: the
    engine bins it point-symmetrically on the voxel itself, valued FloatToInt(double).
    Float weights and integer counts whose products pass 10,527 (where the double and
    float fixed points part) pin that value against the oracle's ordered 8-corner loop."""
    import torch

    import incohautocorrelate_amd as iac

    n = 96
    _, km, mq = W.detector(n)
    shape = iac.mesh_shape_big(mq, 47)
    rng = np.random.default_rng(77)
    hls = []
    for f in range(6):
        npho = int(rng.integers(20, 120))
        pix = rng.choice(n * n, npho, replace=False)
        w = (rng.gamma(1.0, 3.0, npho) if f % 2 == 0 else rng.integers(1, 300, npho)).astype(np.float32)
        hl = np.empty((npho, 4), np.float32)
        hl[:, :3] = km[pix]
        hl[:, 3] = w
        hls.append(hl)
    ref = np.zeros(shape.size_ab ** 3, np.uint64)
    ref_nn = np.zeros(shape.size_ab ** 3, np.uint64)
    for hl in hls:
        ref, _ = O.ac3d_sparse(hl, shape, None, 0, 1, True, mesh=ref)
        ref_nn, _ = O.ac3d_sparse(hl, shape, None, 0, 0, True, mesh=ref_nn)
    assert not np.array_equal(ref, ref_nn)  # the value function matters for these weights
    rec = np.concatenate(hls)
    off = np.concatenate([[0], np.cumsum([len(h) for h in hls])]).astype(np.uint64)
    ctx.set_option(iac.OPT_AC3D_KERNEL, 2)
    ctx.set_option(iac.OPT_PAIR_MODE, mode)
    try:
        m = torch.zeros(shape.size_ab ** 3, dtype=torch.int64, device=dev)
        ctx.ac3d_records_batch(to_dev(km, dev), to_dev(rec, dev), to_dev(off, dev), len(rec), shape, m,
                               interp1=0, interp2=1)
        assert np.array_equal(mesh_np(m), ref)
    finally:
        ctx.set_option(iac.OPT_AC3D_KERNEL, 0)
        ctx.set_option(iac.OPT_PAIR_MODE, 0)


@pytest.mark.parametrize("interp1,interp2", [(0, 0), (1, 1)])
def test_library_tuner(ctx, dev, interp1, interp2):
    """iac_ac3d_tune times every candidate on a frame prefix, applies the fastest, caches it
    per (k-map, mesh, modes); every setting it tries gives the oracle's mesh (the tuner is
    also how the benchmark picks its settings)."""
    import torch

    import incohautocorrelate_amd as iac

    n = 128
    _, km, mq = W.detector(n)
    shape = iac.mesh_shape_big(mq, 63)
    data = W.lognormal_photon_frames(n * n, 60, 80, 1.0, 2, 800, seed=23)
    d_km = to_dev(km, dev)
    pix, cnt, off = _dev_frames(data, dev)
    c = iac.Context(0)
    try:
        best, cands = c.ac3d_tune(d_km, pix, cnt, off, shape, interp1, interp2, True, prefix_frames=40)
        assert len(cands) >= 5 and all(x["ms"] > 0 for x in cands)
        assert best["ms"] == min(x["ms"] for x in cands)
        if interp1 == 1:  # linear binning: the 4096-voxel boxes lead the candidates
            assert cands[0]["box_cap"] == 4096
        again, none = c.ac3d_tune(d_km, pix, cnt, off, shape, interp1, interp2, True, prefix_frames=40)
        assert none == [] and again["ms"] == -1.0
        assert {k: v for k, v in again.items() if k != "ms"} == {k: v for k, v in best.items() if k != "ms"}
        ref, _ = O.ac3d_frames(km, data["pix"], data["cnt"], data["frame_off"], shape, None, interp1, interp2, True)
        for x in cands:  # each candidate setting, applied by hand, is bit-exact
            c.set_option(iac.OPT_PAIR_MODE, x["pair_mode"])
            c.set_option(iac.OPT_CELL_VOX, x["cell_vox"])
            c.set_option(iac.OPT_BOX_CAP, x["box_cap"])
            c.set_option(iac.OPT_FRAME_CHUNK, x["frame_chunk"])
            m = torch.zeros(shape.size_ab ** 2 * shape.size_c, dtype=torch.int64, device=dev)
            c.ac3d_sparse_batch(d_km, pix, cnt, off, int(data["frame_off"][-1]), shape, m,
                                interp1=interp1, interp2=interp2, double_map=True)
            assert np.array_equal(mesh_np(m), ref), x
    finally:
        c.close()
