"""GPU checks of the measurement probes the bench reports beside the engine: every
atomic add reaches the mesh, in each address pattern."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode,hot", [(0, 0), (1, 5000), (3, 0)])
def test_global_atomic_probe_counts_every_add(ctx, dev, mode, hot):
    import torch

    nvox = 65 ** 3
    m = torch.zeros(nvox, dtype=torch.int64, device=dev)
    ms = ctx.probe_atomics(m, 1 << 20, mode, hot)
    assert ms > 0
    assert int(m.sum()) == 1 << 20
    touched = int((m != 0).sum())
    if mode == 1:
        assert touched <= hot
    if mode == 3:
        # runs of 8 on whole 64-byte segments: nothing past the last whole segment
        assert int(m[nvox // 8 * 8:].abs().sum()) == 0


@pytest.mark.parametrize("mode", [2, 4])
def test_lds_atomic_probe_flushes_its_boxes(ctx, dev, mode):
    import torch

    nvox = 65 ** 3
    m = torch.zeros(nvox, dtype=torch.int64, device=dev)
    ctx.probe_atomics(m, 1 << 20, mode, 0)
    total = int(m.sum())
    assert 0 < total <= 1 << 20
