"""iac_ac1d_acuw_records_noalloc: the AC1D AC_uw pass over hit-list records with its work
memory in a static device array (no allocation).  Called through the C ABI, it equals the
allocating entry point iac_ac1d_acuw_records on two calls in a row, the second smaller
(fewer frames, fewer bins), so nothing the first call left in the shared array leaks into
the second; integer photon counts with nearest-neighbour bins make both exact (bit for
bit), linear bins agree to the double rounding of the atomic add order (1e-12).  Sizes
beyond the array are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from gpu_helpers import to_dev
from incohautocorrelate_amd import workloads as W

pytestmark = pytest.mark.gpu


def _records(km, n, nframes, seed):
    d = W.lognormal_photon_frames(n * n, nframes, 200, 1.0, 1, 1500, seed=seed)
    d["cnt"] = np.random.default_rng(seed).integers(1, 5, d["cnt"].size).astype(np.uint16)
    off = d["frame_off"]
    rec = np.concatenate([O.hit_list_from_photons(km, d["pix"][off[f]:off[f + 1]], d["cnt"][off[f]:off[f + 1]])
                          for f in range(nframes)])
    return rec, off


def _run(ctx, entry, d_rec, d_off, nframes, n_total, size, max_q, dq, interp, hist):
    from incohautocorrelate_amd import _abi

    fn = getattr(_abi.lib(), entry)
    return fn(ctx.handle, C.c_void_p(d_rec.data_ptr()), C.c_void_p(d_off.data_ptr()), nframes, n_total, size,
              max_q, dq, interp, 2, C.c_void_p(hist.data_ptr()))


@pytest.mark.parametrize("interp", [0, 1])
def test_noalloc_equals_allocating_entry(ctx, dev, interp):
    import torch

    import incohautocorrelate_amd as iac

    n = 192
    _, km, mq = W.detector(n)
    for nframes, size, seed in ((48, 1200, 3), (9, 400, 4)):
        max_q, dq = iac.ac1d_shape(mq, size)
        rec, off = _records(km, n, nframes, seed + 10 * interp)
        d_rec, d_off = to_dev(rec, dev), to_dev(off, dev)
        want = torch.zeros(size, dtype=torch.float64, device=dev)
        got = torch.zeros(size, dtype=torch.float64, device=dev)
        assert _run(ctx, "iac_ac1d_acuw_records", d_rec, d_off, nframes, len(rec), size, max_q, dq, interp,
                    want) == 0
        assert _run(ctx, "iac_ac1d_acuw_records_noalloc", d_rec, d_off, nframes, len(rec), size, max_q, dq,
                    interp, got) == 0
        torch.cuda.synchronize()
        w, g = want.cpu().numpy(), got.cpu().numpy()
        assert w.sum() > 0
        if interp == 0:
            assert np.array_equal(g, w)
        else:
            assert np.allclose(g, w, rtol=1e-12, atol=1e-9)


def test_noalloc_refuses_oversize(ctx, dev):
    import torch

    import incohautocorrelate_amd as iac

    _, km, mq = W.detector(64)
    max_q, dq = iac.ac1d_shape(mq, 100)
    off = torch.zeros(65538, dtype=torch.int64, device=dev)  # 65,537 frames, all empty
    rec = torch.zeros((1, 4), dtype=torch.float32, device=dev)
    hist = torch.zeros(100, dtype=torch.float64, device=dev)
    IAC_ERR_LIMIT = -8
    assert _run(ctx, "iac_ac1d_acuw_records_noalloc", rec, off, 65537, 1, 100, max_q, dq, 0, hist) == IAC_ERR_LIMIT
    big = torch.zeros(6145, dtype=torch.float64, device=dev)
    assert _run(ctx, "iac_ac1d_acuw_records_noalloc", rec, off[:3], 2, 1, 6145, max_q, dq, 0, big) == IAC_ERR_LIMIT
