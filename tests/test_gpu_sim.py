"""GPU simulator against the FP64 oracle: the split-precision complex GEMM (bf16 /
fp16 hi + lo, FP32 accumulation) within 1e-3 normwise relative intensity error,
subsampling, modes, GEMM row chunks, the reference-arithmetic FP32 direct sum,
and bit-exact rescale / Poisson / noise post-processing."""
import numpy as np
import pytest

import oracle as O
from gpu_helpers import to_dev
from incohautocorrelate_amd import workloads as W

pytestmark = pytest.mark.gpu

WL = 1.5  # Angstrom; emitters in Angstrom, pixel map in um (only directions matter)


def _setup(n, nem, nframes, modes=1, seed=0, edge=500.0):
    pm, _, _ = W.detector(n)
    em = W.emitters_cube(nem, edge, seed)
    ph = W.emitter_phases(nframes * modes, nem, seed + 1)
    return pm, em, ph


def _rel(got, ref):
    return float(np.linalg.norm(got - ref) / np.linalg.norm(ref))


def _gemm(ctx, dev, pm, em, ph, mode, modes=1, out_kind=0, subsampling=0, us=(0, 0, 0), vs=(0, 0, 0)):
    import torch

    plan = ctx.sim_plan(to_dev(pm, dev), torch.from_numpy(np.ascontiguousarray(em)).to(dev), WL, mode,
                        subsampling, us, vs)
    nframes = ph.shape[0] // modes
    out = torch.zeros((nframes, pm.shape[0]), dtype=torch.float32, device=dev)
    ctx.sim_run(plan, torch.from_numpy(np.ascontiguousarray(ph)).to(dev), out, modes, out_kind)
    res = out.cpu().numpy()
    plan.destroy()
    return res


@pytest.mark.parametrize("mode", [1, 2])  # BF16X3, FP16X3
@pytest.mark.parametrize("out_kind", [0, 1])
def test_split_gemm_vs_fp64(ctx, dev, mode, out_kind):
    pm, em, ph = _setup(32, 600, 6, seed=mode)
    ref = O.simulate_fp64(pm, em, ph, WL, out_kind=out_kind)
    got = _gemm(ctx, dev, pm, em, ph, mode, out_kind=out_kind)
    err = _rel(got, ref)
    print(f"mode {mode} kind {out_kind}: normwise relative error {err:.3e}")
    assert err <= 1e-3
    # and per pixel, relative to the frame's mean intensity
    assert np.max(np.abs(got - ref)) <= 1e-2 * ref.mean()


def test_split_beats_single_product(ctx, dev):
    pm, em, ph = _setup(24, 400, 4, seed=7)
    ref = O.simulate_fp64(pm, em, ph, WL, out_kind=1)
    e3 = _rel(_gemm(ctx, dev, pm, em, ph, 1, out_kind=1), ref)
    e1 = _rel(_gemm(ctx, dev, pm, em, ph, 3, out_kind=1), ref)
    f1 = _rel(_gemm(ctx, dev, pm, em, ph, 4, out_kind=1), ref)
    print(f"bf16x3 {e3:.2e}  bf16 {e1:.2e}  fp16 {f1:.2e}")
    assert e1 > 5 * e3  # the lo products carry the precision
    assert f1 < e1      # fp16 keeps 3 more mantissa bits than bf16


def test_subsampling_modes_and_chunks(ctx, dev):
    import incohautocorrelate_amd as iac

    pm, em, ph = _setup(32, 300, 3, modes=2, seed=11)
    us, vs = (12.5, 0.0, 0.0), (0.0, 12.5, 0.0)  # a 75 um pixel in 3 x 3 subpixels
    ref = O.simulate_fp64(pm, em, ph, WL, subsampling=1, u_step=np.array(us), v_step=np.array(vs), modes=2)
    got = _gemm(ctx, dev, pm, em, ph, 1, modes=2, subsampling=1, us=us, vs=vs)
    assert _rel(got, ref) <= 1e-3
    # four launches of 256 whole pixels each (2304 rows): every (row, column) sums the same
    # products in the same order, so the result is the same bit for bit
    ctx.set_option(iac.OPT_SIM_CHUNK_ROWS, 100)
    try:
        got2 = _gemm(ctx, dev, pm, em, ph, 1, modes=2, subsampling=1, us=us, vs=vs)
    finally:
        ctx.set_option(iac.OPT_SIM_CHUNK_ROWS, 0)
    assert np.array_equal(got2, got)


@pytest.mark.parametrize("n,nframes", [(13, 5), (20, 130)])
def test_ragged_shapes(ctx, dev, n, nframes):
    """Pixel counts that are not a multiple of 4 (scalar stores) or of the 256-row block, frame
    counts across the 128-column block edge, an emitter count that is not a multiple of 32."""
    pm, em, ph = _setup(n, 77, nframes, seed=n)
    ref = O.simulate_fp64(pm, em, ph, WL, out_kind=1)
    got = _gemm(ctx, dev, pm, em, ph, 2, out_kind=1)
    assert _rel(got, ref) <= 1e-3


@pytest.mark.parametrize("mode", [1, 2, 4])  # BF16X3, FP16X3, FP16
def test_simulator_shape_512_10k(ctx, dev, mode):
    """The simulator at its benchmark shape: 512^2 detector (75 um at 0.5 m), 10,000 emitters in
    a 500 A cube, lambda 1.5 A; 8 frames through the GEMM; 8 central detector rows x 4 frames
    against the FP64 oracle, both output kinds, at <= 1e-3 normwise relative error."""
    import torch

    pm, em, ph = _setup(512, 10000, 8, seed=2)
    p0, npx = 252 * 512, 8 * 512
    d_pm = to_dev(pm, dev)
    d_em = torch.from_numpy(np.ascontiguousarray(em)).to(dev)
    d_ph = torch.from_numpy(np.ascontiguousarray(ph)).to(dev)
    plan = ctx.sim_plan(d_pm, d_em, WL, mode)
    out = torch.zeros((8, pm.shape[0]), dtype=torch.float32, device=dev)
    try:
        for kind in (0, 1):
            ref = O.simulate_fp64(pm, em, ph, WL, out_kind=kind, pixels=(p0, npx), frames=(0, 4))
            ctx.sim_run(plan, d_ph, out, 1, kind)
            got = out[:4, p0:p0 + npx].cpu().numpy().astype(np.float64)
            err = _rel(got, ref)
            print(f"mode {mode} kind {kind}: {err:.3e}")
            assert err <= 1e-3
            assert np.isfinite(out.cpu().numpy()).all()
    finally:
        plan.destroy()


def test_direct_fp32_vs_fp64(ctx, dev):
    import torch

    pm, em, ph = _setup(32, 500, 3, seed=5)
    em32 = em.astype(np.float32)
    ph32 = ph.astype(np.float32)
    # the reference casts positions and phases to float (Simulator.cpp:730-737)
    ref = O.simulate_fp64(pm, em32.astype(np.float64), ph32.astype(np.float64), WL, out_kind=1)
    out = torch.zeros((3, pm.shape[0]), dtype=torch.float32, device=dev)
    ctx.sim_direct_fp32(to_dev(pm, dev), to_dev(em32, dev), to_dev(ph32, dev), WL, out, out_kind=1)
    err = _rel(out.cpu().numpy(), ref)
    print(f"direct fp32 normwise relative error {err:.3e}")
    assert err <= 1e-2


def test_direct_fp32_subsampling_modes(ctx, dev):
    import torch

    pm, em, ph = _setup(12, 64, 2, modes=3, seed=9, edge=100.0)
    em32, ph32 = em.astype(np.float32), ph.astype(np.float32)
    us, vs = (12.5, 0.0, 0.0), (0.0, 12.5, 0.0)
    ref = O.simulate_fp64(pm, em32.astype(np.float64), ph32.astype(np.float64), WL, subsampling=1,
                          u_step=np.array(us), v_step=np.array(vs), modes=3)
    out = torch.zeros((2, pm.shape[0]), dtype=torch.float32, device=dev)
    ctx.sim_direct_fp32(to_dev(pm, dev), to_dev(em32, dev), to_dev(ph32, dev), WL, out, modes=3,
                        subsampling=1, u_step=us, v_step=vs)
    assert _rel(out.cpu().numpy(), ref) <= 1e-3


@pytest.mark.parametrize("npix,poisson,noise,vpp", [(3000, True, 0.0, 1.0), (1024, True, 0.3, 2.5),
                                                    (5000, False, 0.5, 1.0), (700, True, 0.0, 1.0)])
def test_postprocess_bit_exact(ctx, dev, npix, poisson, noise, vpp):
    import torch

    rng = np.random.default_rng(npix)
    I = rng.gamma(1.0, 1.0, (5, npix)).astype(np.float32)
    n_exp = 0.01 * npix if npix != 700 else 40.0 * npix  # the last case exercises Box-Muller
    ref = O.sim_postprocess(I, n_exp, poisson, noise, vpp, seed=77, frame0=3)
    d = torch.from_numpy(I.copy()).to(dev)
    ctx.sim_postprocess(d, n_exp, poisson, noise, vpp, seed=77, frame0=3)
    got = d.cpu().numpy()
    assert np.array_equal(got, ref)


def test_postprocess_counts_after_gemm(ctx, dev):
    """The simulation chain at small size: GEMM field -> rescale to 0.01 photons / pixel -> Poisson."""
    import torch

    pm, em, ph = _setup(64, 256, 16, seed=21)
    plan = ctx.sim_plan(to_dev(pm, dev), torch.from_numpy(em).to(dev), WL, 1)
    out = torch.zeros((16, pm.shape[0]), dtype=torch.float32, device=dev)
    ctx.sim_run(plan, torch.from_numpy(ph).to(dev), out, 1, 0)
    pre = out.cpu().numpy()
    ctx.sim_postprocess(out, 0.01 * pm.shape[0], True, seed=5)
    got = out.cpu().numpy()
    assert np.array_equal(got, O.sim_postprocess(pre, 0.01 * pm.shape[0], True, seed=5))
    tot = got.sum(axis=1)
    assert abs(tot.mean() - 40.96) < 5 * np.sqrt(40.96 / 16)
    plan.destroy()
