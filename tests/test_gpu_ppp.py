"""GPU parity of the LAP photon finder chain (iac_ppp_lap: pixel mask, gain correction,
largest-adjacent-pixel photon finder, per-event photon counts; PPP::ProcessData_PF_LAP,
IAC/PPP.cpp:469-495) against the oracle, bit for bit."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ADU = 37.0


def layout(name):
    """(npix, panels) -- panels tile the first sum(fs*ss) pixels, as the reference sizes them."""
    if name == "one":
        return 64 * 48, [(0, 64, 48)]
    if name == "eight":  # eight 32x12 modules in a shuffled list, 101 pixels past the panels
        pan = [(k * 384, 32, 12) for k in range(8)]
        order = [3, 0, 7, 1, 5, 2, 6, 4]
        return 8 * 384 + 101, [pan[k] for k in order]
    if name == "odd":  # one-column, narrow, strip-sized and one-row panels, 5 pixels past them
        return 2491 + 63 + 5, [(0, 1, 70), (70, 125, 3), (445, 62, 33), (2491, 63, 1)]
    if name == "even":  # even starts and widths (two columns per lane): 126, 2 and 248 wide
        return 9700, [(8950, 248, 3), (0, 126, 70), (8820, 2, 65)]
    if name == "modules":  # a 1024x1024 detector as eight 1024x128 modules
        return 1024 * 1024, [(k * 1024 * 128, 1024, 128) for k in range(8)]
    raise ValueError(name)


def frames_for(npix, nframes, seed):
    rng = np.random.default_rng(seed)
    photons = rng.poisson(0.4, (nframes, npix))
    return (ADU * (photons + rng.uniform(-0.3, 1.0, (nframes, npix)))).astype(np.float32)


def maps_for(npix, seed):
    rng = np.random.default_rng(seed + 100)
    mask = (rng.uniform(size=npix) > 0.05).astype(np.int32)
    offset = rng.uniform(0.0, 3.0, npix).astype(np.float32)
    gain = rng.uniform(0.8, 1.2, npix).astype(np.float32)
    gain[rng.integers(0, npix, 20)] = 0.0
    gain[rng.integers(0, npix, 20)] = -1.0
    return mask, offset, gain


def run_gpu(ctx, dev, I, panels, mask, offset, gain, gain_only, inplace, **kw):
    import torch

    d_in = torch.from_numpy(I).to(dev)
    d_out = d_in if inplace else torch.full_like(d_in, float("nan"))
    nf = I.shape[0]
    pc = torch.full((nf,), -7, dtype=torch.int32, device=dev)
    mi = torch.full((nf,), -7.0, dtype=torch.float32, device=dev)
    t = (lambda a: None if a is None else torch.from_numpy(a).to(dev))
    ctx.ppp_lap(d_in, d_out, panels, ADU, mask=t(mask), offset=t(offset), gain=t(gain), gain_only=gain_only,
                photon_count=pc, mean_intensity=mi, **kw)
    torch.cuda.synchronize()
    return d_out.cpu().numpy(), pc.cpu().numpy(), mi.cpu().numpy()


@pytest.mark.parametrize("name,use_mask,use_gain,gain_only,inplace,nframes", [
    ("one", False, False, False, False, 5),
    ("eight", True, False, False, False, 17),
    ("eight", True, True, False, True, 33),
    ("one", True, True, True, False, 3),
    ("one", False, True, False, True, 16),
    ("eight", False, False, False, True, 1),
    ("odd", True, True, False, False, 9),
    ("odd", False, False, False, False, 4),
    ("even", True, True, False, False, 7),
    ("even", False, False, False, True, 3),
    ("even", True, False, False, False, 1),
])
def test_lap_vs_oracle(ctx, dev, name, use_mask, use_gain, gain_only, inplace, nframes):
    npix, panels = layout(name)
    print(f"lap case {name} mask={use_mask} gain={use_gain} gain_only={gain_only} inplace={inplace} "
          f"nframes={nframes}", flush=True)
    I = frames_for(npix, nframes, seed=nframes)
    mask, offset, gain = maps_for(npix, nframes)
    mask = mask if use_mask else None
    offset, gain = (offset, gain) if use_gain else (None, None)
    ref, rpc, rmi = O.ppp_lap(I, panels, adu=ADU, mask=mask, offset=offset, gain=gain, gain_only=gain_only)
    out, pc, mi = run_gpu(ctx, dev, I, panels, mask, offset, gain, gain_only, inplace)
    assert np.array_equal(out.view(np.int32), ref.view(np.int32))
    assert np.array_equal(pc, rpc)
    assert np.array_equal(mi.view(np.int32), rmi.view(np.int32))
    if not gain_only:  # the finder did something: some pixels gained a photon from a neighbour
        assert (ref > 0).any()


def test_lap_full_detector_modules(ctx, dev):
    """1024x1024 detector as eight 1024x128 modules, 20 frames, mask and gain maps."""
    npix, panels = layout("modules")
    I = frames_for(npix, 20, seed=5)
    mask, offset, gain = maps_for(npix, 5)
    ref, rpc, rmi = O.ppp_lap(I, panels, adu=ADU, mask=mask, offset=offset, gain=gain)
    out, pc, mi = run_gpu(ctx, dev, I, panels, mask, offset, gain, False, False)
    assert np.array_equal(out.view(np.int32), ref.view(np.int32))
    assert np.array_equal(pc, rpc) and np.array_equal(mi.view(np.int32), rmi.view(np.int32))


@pytest.mark.parametrize("groups", [1, 3, 16, 1000])
def test_lap_block_order_groups(dev, groups):
    """IAC_OPT_LAP_GROUPS only reorders the blocks: 37 frames (19 two-frame groups: not a
    multiple of 3 or 16, fewer than 1000) on the odd and the even layouts equal the oracle."""
    import incohautocorrelate_amd as iac

    c = iac.Context(0)
    try:
        c.set_option(iac.OPT_LAP_GROUPS, groups)
        for name in ("odd", "even"):
            npix, panels = layout(name)
            print(f"lap groups={groups} layout={name}", flush=True)
            I = frames_for(npix, 37, seed=groups)
            mask, offset, gain = maps_for(npix, groups)
            ref, rpc, rmi = O.ppp_lap(I, panels, adu=ADU, mask=mask, offset=offset, gain=gain)
            out, pc, mi = run_gpu(c, dev, I, panels, mask, offset, gain, False, False)
            assert np.array_equal(out.view(np.int32), ref.view(np.int32))
            assert np.array_equal(pc, rpc) and np.array_equal(mi.view(np.int32), rmi.view(np.int32))
    finally:
        c.close()


def test_lap_thresholds(ctx, dev):
    """Other seed / combined thresholds than the defaults."""
    npix, panels = layout("eight")
    I = frames_for(npix, 7, seed=11)
    ref, rpc, _ = O.ppp_lap(I, panels, adu=ADU, seed=0.3, comb=0.7)
    out, pc, _ = run_gpu(ctx, dev, I, panels, None, None, None, False, False, seed_threshold=0.3,
                         combined_threshold=0.7)
    assert np.array_equal(out.view(np.int32), ref.view(np.int32)) and np.array_equal(pc, rpc)


def test_lap_rejects_bad_panels(ctx, dev):
    import torch

    import incohautocorrelate_amd as iac

    d = torch.zeros((2, 100), dtype=torch.float32, device=dev)
    o = torch.zeros_like(d)
    with pytest.raises(iac.IACError) as e:  # overlapping panels
        ctx.ppp_lap(d, o, [(0, 10, 5), (40, 10, 5)])
    assert e.value.code == -1
    with pytest.raises(iac.IACError) as e:  # a panel past the panels' pixel count
        ctx.ppp_lap(d, o, [(10, 10, 5)])
    assert e.value.code == -6
    with pytest.raises(iac.IACError) as e:  # more pixels than the frame
        ctx.ppp_lap(d, o, [(0, 10, 11)])
    assert e.value.code == -6
    with pytest.raises(iac.IACError) as e:
        ctx.ppp_lap(d, o, [(k, 1, 1) for k in range(129)])
    assert e.value.code == -8
    with pytest.raises(iac.IACError) as e:  # offset without gain
        ctx.ppp_lap(d, o, [], offset=torch.zeros(100, device=dev))
    assert e.value.code == -1
    ctx.ppp_lap(d[:0], o[:0], [(0, 10, 10)])  # no frames: nothing to do


@pytest.mark.parametrize("kind", ["whole", "huge_value", "huge_sum", "negative_whole"])
def test_photon_count_paths(ctx, dev, kind):
    """The per-event photon counts: exact integer sums where every value is whole and the
    magnitudes sum to at most 2^24, the serial float chain otherwise (values above 2^24, or
    partial sums that round), with and without the LAP finder."""
    rng = np.random.default_rng(21)
    npix, nf = 64 * 40 + 3, 6
    I = rng.poisson(2.0, (nf, npix)).astype(np.float32)
    if kind == "huge_value":
        I[1, 17] = 2.0 ** 25 + 2
    elif kind == "huge_sum":  # each value whole and below 2^24, the running sum rounds
        I[2, ::50] = 2.0 ** 20 + 1
    elif kind == "negative_whole":
        I[3, ::7] = -5.0
    for panels in ([], [(0, 64, 40)]):
        ref, rpc, rmi = O.ppp_lap(I, panels, adu=1.0)
        import torch

        d_in = torch.from_numpy(I).to(dev)
        d_out = torch.empty_like(d_in)
        pc = torch.empty(nf, dtype=torch.int32, device=dev)
        mi = torch.empty(nf, dtype=torch.float32, device=dev)
        ctx.ppp_lap(d_in, d_out, panels, 1.0, photon_count=pc, mean_intensity=mi)
        torch.cuda.synchronize()
        assert np.array_equal(d_out.cpu().numpy().view(np.int32), ref.view(np.int32))
        assert np.array_equal(pc.cpu().numpy(), rpc), (kind, panels)
        assert np.array_equal(mi.cpu().numpy().view(np.int32), rmi.view(np.int32))
    if kind == "huge_sum":  # the serial chain really differs from the exact sum here
        assert int(rpc[2]) != int(I[2].astype(np.int64).sum())


def test_photon_count_past_int_range(ctx, dev):
    """Sums past 2^31 convert as the reference's x86 build does: 0x80000000, on both the LAP
    chain (gain-only frames of large ADU values) and the averaged-intensity pass."""
    import torch

    I = np.full((3, 1000), 3.0e6, np.float32)
    I[1] = 1.0  # a frame that stays in range
    ref, rpc, rmi = O.ppp_lap(I, [], gain_only=True)
    assert rpc[0] == -2 ** 31 and rpc[1] == 1000
    d = torch.from_numpy(I).to(dev)
    pc = torch.empty(3, dtype=torch.int32, device=dev)
    mi = torch.empty(3, dtype=torch.float32, device=dev)
    ctx.ppp_lap(d, torch.empty_like(d), [], gain_only=True, photon_count=pc, mean_intensity=mi)
    torch.cuda.synchronize()
    assert np.array_equal(pc.cpu().numpy(), rpc) and np.array_equal(mi.cpu().numpy().view(np.int32), rmi.view(np.int32))
    # averaged intensity with photon discretisation (step 1): 1000 pixels of 3e6 photons each
    avg, omi, opc = O.average_intensity(I, 0.0, 1.0)
    assert opc[0] == -2 ** 31
    g_avg = torch.empty(1000, dtype=torch.float32, device=dev)
    g_pc = torch.empty(3, dtype=torch.int32, device=dev)
    g_mi = torch.empty(3, dtype=torch.float32, device=dev)
    ctx.average_intensity(d, None, 0.0, 1.0, g_avg, g_mi, g_pc)
    torch.cuda.synchronize()
    assert np.array_equal(g_pc.cpu().numpy(), opc)
