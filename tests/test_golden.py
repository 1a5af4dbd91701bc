"""Golden fixtures (tests/golden/golden.npz, written by make_golden.py after its
brute-force and closed-form cross-checks): the oracle must still produce them
(CPU), and the GPU path must reproduce them without consulting the oracle."""
from pathlib import Path

import numpy as np
import pytest

G = np.load(Path(__file__).resolve().parent / "golden" / "golden.npz", allow_pickle=False)

AC3D_MODES = {"nn": (0, 0, 1, False), "lin": (1, 1, 1, False), "nn_lin_single": (0, 1, 0, False),
              "lin_nn": (1, 0, 1, False), "nn_rot": (0, 0, 1, True)}


def _dense(prefix, nvox):
    m = np.zeros(nvox, np.uint64)
    m[G[prefix + "_idx"]] = G[prefix + "_val"]
    return m


@pytest.mark.parametrize("mode", list(AC3D_MODES))
def test_oracle_reproduces_ac3d(mode):
    import oracle as O
    from incohautocorrelate_amd import workloads as W

    i1, i2, dm, rot = AC3D_MODES[mode]
    _, km, mq = W.detector(16)
    assert np.array_equal(km, G["ac3d_kmap"])
    shape = O.mesh_shape_big(mq, 15)
    R = np.tile(G["ac3d_rot"], (3, 1)) if rot else None
    mesh, _ = O.ac3d_frames(km, G["ac3d_pix"], G["ac3d_cnt"], G["ac3d_off"], shape, R, i1, i2, bool(dm))
    assert np.array_equal(mesh, _dense(f"ac3d_{mode}", mesh.size))


def test_oracle_reproduces_cq_merge_ac1d_sim_post():
    import oracle as O
    from incohautocorrelate_amd import workloads as W

    _, km, mq = W.detector(14)
    ss = O.mesh_shape_small_cofq(mq, 21)
    nvs = ss.size_ab ** 2 * ss.size_c
    cqf, _ = O.cq_small(np.ones(196, np.float32), km, ss, 0, 1.0)
    assert np.array_equal(cqf, _dense("cq_flat", nvs))
    mult = float(G["cq_mult"][0])
    assert mult == O.cq_small_multiplicator(G["cq_I"])
    cqr, _ = O.cq_small(G["cq_I"], km, ss, 1, mult)
    assert np.array_equal(cqr, _dense("cq_lin", nvs))
    small = cqr.astype(np.float64) / mult
    rw, m2 = O.merge_prepare(G["merge_rots"], G["merge_mi"], small)
    assert np.array_equal(rw, G["merge_rw"]) and m2 == float(G["merge_mult"][0])
    big, _ = O.merge_cq(small, ss, rw, 0, m2)
    assert np.array_equal(big, _dense("merge", ss.size_ab ** 3))
    max_q, dq, m1 = G["ac1d_par"]
    h, _ = O.ac1d_cq(G["ac1d_I"], G["ac1d_kmap"], 64, float(max_q), float(dq), 1, float(m1))
    assert np.array_equal(h, G["ac1d_lin"])
    f64 = O.simulate_fp64(G["sim_pm"], G["sim_em"], G["sim_ph"], 1.5, out_kind=1)
    assert np.array_equal(f64, G["sim_int"])
    assert np.array_equal(O.sim_postprocess(G["post_in"], 30.0, True, 0.0, 1.0, seed=99, frame0=4),
                          G["post_poisson"])
    assert np.array_equal(O.sim_postprocess(G["post_in"], 30.0, True, 0.5, 2.0, seed=99, frame0=4),
                          G["post_noise"])


@pytest.mark.gpu
@pytest.mark.parametrize("mode", list(AC3D_MODES))
def test_gpu_matches_golden_ac3d(ctx, dev, mode):
    import torch

    import incohautocorrelate_amd as iac
    from gpu_helpers import mesh_np, to_dev
    from incohautocorrelate_amd import workloads as W

    i1, i2, dm, rot = AC3D_MODES[mode]
    _, km, mq = W.detector(16)
    shape = iac.mesh_shape_big(mq, 15)
    nvox = shape.size_ab ** 2 * shape.size_c
    m = torch.zeros(nvox, dtype=torch.int64, device=dev)
    rots = to_dev(np.tile(G["ac3d_rot"], (3, 1)).astype(np.float32), dev) if rot else None
    ctx.ac3d_sparse_batch(to_dev(km, dev), to_dev(G["ac3d_pix"], dev), to_dev(G["ac3d_cnt"], dev),
                          to_dev(G["ac3d_off"], dev), int(G["ac3d_off"][-1]), shape, m, rots, i1, i2, bool(dm))
    assert np.array_equal(mesh_np(m), _dense(f"ac3d_{mode}", nvox))


@pytest.mark.gpu
def test_gpu_matches_golden_cq_merge_ac1d_post(ctx, dev):
    import torch

    import incohautocorrelate_amd as iac
    from gpu_helpers import mesh_np, to_dev
    from incohautocorrelate_amd import workloads as W

    _, km, mq = W.detector(14)
    ss = iac.mesh_shape_small_cofq(mq, 21)
    nvs = ss.size_ab ** 2 * ss.size_c
    m = torch.zeros(nvs, dtype=torch.int64, device=dev)
    ctx.cq_small(to_dev(km, dev), to_dev(np.ones(196, np.float32), dev), ss, m, 0, 1.0)
    assert np.array_equal(mesh_np(m), _dense("cq_flat", nvs))
    mult = float(G["cq_mult"][0])
    m.zero_()
    ctx.cq_small(to_dev(km, dev), to_dev(G["cq_I"], dev), ss, m, 1, mult)
    cqr = mesh_np(m)
    assert np.array_equal(cqr, _dense("cq_lin", nvs))
    small = torch.from_numpy(cqr.astype(np.float64) / mult).to(dev)
    big = torch.zeros(ss.size_ab ** 3, dtype=torch.int64, device=dev)
    ctx.merge_cq(small, ss, to_dev(G["merge_rw"], dev), 0, float(G["merge_mult"][0]), big)
    assert np.array_equal(mesh_np(big), _dense("merge", ss.size_ab ** 3))
    max_q, dq, m1 = G["ac1d_par"]
    h = torch.zeros(64, dtype=torch.int64, device=dev)
    ctx.ac1d_cq(to_dev(G["ac1d_kmap"], dev), to_dev(G["ac1d_I"], dev), 64, float(max_q), float(dq), 1, float(m1), h)
    assert np.array_equal(mesh_np(h), G["ac1d_lin"])
    post = torch.from_numpy(G["post_in"].copy()).to(dev)
    ctx.sim_postprocess(post, 30.0, True, 0.5, 2.0, seed=99, frame0=4)
    assert np.array_equal(post.cpu().numpy(), G["post_noise"])
