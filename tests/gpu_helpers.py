"""Small helpers shared by the GPU parity tests (upload, run, download)."""
from __future__ import annotations

import numpy as np


def to_dev(a, dev):
    import torch

    a = np.ascontiguousarray(a)
    if a.dtype == np.uint64:
        a = a.view(np.int64)
    elif a.dtype == np.uint32:
        a = a.view(np.int32)
    elif a.dtype == np.uint16:
        a = a.view(np.int16)
    return torch.from_numpy(a).to(dev)


def mesh_np(t) -> np.ndarray:
    return t.cpu().numpy().view(np.uint64)


def gpu_single(ctx, dev, hl, shape, rot=None, interp1=0, interp2=0, double_map=True, mesh=None):
    import torch

    nvox = shape.size_ab * shape.size_ab * shape.size_c
    m = torch.zeros(nvox, dtype=torch.int64, device=dev) if mesh is None else mesh
    R = np.eye(3, dtype=np.float32).reshape(9) if rot is None else np.asarray(rot, np.float32)
    h = to_dev(np.asarray(hl, np.float32).reshape(-1, 4), dev) if len(hl) else None
    ctx.ac3d_sparse(h, R, shape, m, interp1, interp2, double_map)
    return m


def gpu_batch(ctx, dev, kmap, pix, cnt, off, shape, rots=None, interp1=0, interp2=0,
              double_map=True, mesh=None):
    import torch

    nvox = shape.size_ab * shape.size_ab * shape.size_c
    m = torch.zeros(nvox, dtype=torch.int64, device=dev) if mesh is None else mesh
    n = int(off[-1])
    ctx.ac3d_sparse_batch(to_dev(kmap, dev), to_dev(np.asarray(pix, np.uint32), dev),
                          to_dev(np.asarray(cnt, np.uint16), dev),
                          to_dev(np.asarray(off, np.uint64), dev), n, shape, m,
                          rots=None if rots is None else to_dev(np.asarray(rots, np.float32), dev),
                          interp1=interp1, interp2=interp2, double_map=double_map)
    return m
