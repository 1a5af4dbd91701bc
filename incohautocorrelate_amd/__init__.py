"""MI355X-native pixel-pair autocorrelation path of IncohAutoCorrelate.

The compute lives in libiac_hip.so (hand-written gfx950 HIP kernels behind the
C ABI of include/iac_hip.h).  This module is the Python view of that ABI used
by the tests and the benchmark: device buffers are torch tensors on the GPU
(plumbing only), work is enqueued on torch's current HIP stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import IACError, MeshShape, check, fptr, lib

__all__ = [
    "IACError", "MeshShape", "lib", "generate_pixel_map", "calc_kmap", "mesh_shape_big",
    "mesh_shape_small_cofq", "Context", "NEAREST", "LINEAR",
]

NEAREST = 0
LINEAR = 1

# iac_ctx_set_option keys (include/iac_hip.h)
OPT_AC3D_KERNEL = 1
OPT_CELL_VOX = 2
OPT_FRAME_CHUNK = 3


def _f32(a, n=None) -> np.ndarray:
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if n is not None and arr.size != n:
        raise ValueError(f"expected {n} floats, got {arr.size}")
    return arr


def generate_pixel_map(size_a: int, size_b: int, pixel_size: float, center, vec_a, vec_b) -> np.ndarray:
    """Simulator::GeneratePixelMap (Simulator.cpp:495-519): [size_a*size_b, 3] float32."""
    out = np.empty((size_a * size_b, 3), np.float32)
    c, a, b = _f32(center, 3), _f32(vec_a, 3), _f32(vec_b, 3)
    check(lib().iac_generate_pixel_map(size_a, size_b, float(pixel_size), fptr(c), fptr(a), fptr(b),
                                       fptr(out)), "iac_generate_pixel_map")
    return out


def calc_kmap(pixmap: np.ndarray):
    """Detector::Calc_kMap (Detector.cpp:310-361): (kmap, max_k, min_k, max_q)."""
    pm = _f32(pixmap).reshape(-1, 3)
    kmap = np.empty_like(pm)
    mx, mn, mq = (np.empty(3, np.float32) for _ in range(3))
    check(lib().iac_calc_kmap(fptr(pm), pm.shape[0], fptr(kmap), fptr(mx), fptr(mn), fptr(mq)),
          "iac_calc_kmap")
    return kmap, mx, mn, mq


def mesh_shape_big(max_q, edge_size: int, q_zoom: float = 1.0) -> MeshShape:
    """ACMesh::CreateBigMeshForDetector (ACMesh.cpp:74-106)."""
    s = MeshShape()
    mq = _f32(max_q, 3)
    check(lib().iac_mesh_shape_big(fptr(mq), int(edge_size), float(q_zoom), C.byref(s)),
          "iac_mesh_shape_big")
    return s


def mesh_shape_small_cofq(max_q, perp_size: int, q_zoom: float = 1.0) -> MeshShape:
    """ACMesh::CreateSmallMesh_CofQ_ForDetector (ACMesh.cpp:140-206)."""
    s = MeshShape()
    mq = _f32(max_q, 3)
    check(lib().iac_mesh_shape_small_cofq(fptr(mq), int(perp_size), float(q_zoom), C.byref(s)),
          "iac_mesh_shape_small_cofq")
    return s


def _ptr(t) -> int | None:
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device (cuda) tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return t.data_ptr()


class Context:
    """One iac_ctx bound to a device; enqueues on torch's current stream."""

    def __init__(self, device: int = 0, torch_stream: bool = True):
        self._h = C.c_void_p()
        check(lib().iac_ctx_create(int(device), C.byref(self._h)), "iac_ctx_create")
        self.device = device
        if torch_stream:
            import torch

            s = torch.cuda.current_stream(device)
            check(lib().iac_ctx_set_stream(self._h, C.c_void_p(s.cuda_stream)), "iac_ctx_set_stream")

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def close(self) -> None:
        if self._h:
            lib().iac_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def sync(self) -> None:
        check(lib().iac_ctx_sync(self._h), "iac_ctx_sync")

    def set_option(self, option: int, value: int) -> None:
        check(lib().iac_ctx_set_option(self._h, int(option), int(value)), "iac_ctx_set_option")

    def enable_timing(self, on: bool = True) -> None:
        check(lib().iac_ctx_enable_timing(self._h, int(bool(on))), "iac_ctx_enable_timing")

    def kernel_time(self, reset: bool = True) -> tuple[float, int]:
        """(summed ms, launches) of the timed dominant kernels since the last reset."""
        ms, n = C.c_double(), C.c_uint64()
        check(lib().iac_ctx_timing(self._h, C.byref(ms), C.byref(n), int(bool(reset))),
              "iac_ctx_timing")
        return float(ms.value), int(n.value)

    def dropped(self) -> int:
        v = C.c_uint64()
        check(lib().iac_dropped_updates(self._h, C.byref(v)), "iac_dropped_updates")
        return int(v.value)

    # ---- 3D AC_uw ------------------------------------------------------------
    def ac3d_sparse(self, hitlist, rot, shape: MeshShape, mesh, interp1: int = NEAREST,
                    interp2: int = NEAREST, double_map: bool = True) -> None:
        """Detector::AutoCorrelateSparseList for one frame's {kx,ky,kz,value} list."""
        n = hitlist.shape[0] if hitlist is not None else 0
        r = _f32(rot, 9)
        check(lib().iac_ac3d_sparse(self._h, _ptr(hitlist) if n else None, n, fptr(r), C.byref(shape),
                                    interp1, interp2, int(bool(double_map)), _ptr(mesh)),
              "iac_ac3d_sparse")

    def ac3d_sparse_batch(self, kmap, pix, cnt, frame_off, n_total: int, shape: MeshShape, mesh,
                          rots=None, interp1: int = NEAREST, interp2: int = NEAREST,
                          double_map: bool = True) -> None:
        """All frames of (pixel, count) photon lists in one persistent launch."""
        nframes = frame_off.shape[0] - 1
        check(lib().iac_ac3d_sparse_batch(self._h, _ptr(kmap), kmap.shape[0], _ptr(pix), _ptr(cnt),
                                          _ptr(frame_off), nframes, int(n_total), _ptr(rots),
                                          C.byref(shape), interp1, interp2, int(bool(double_map)),
                                          _ptr(mesh)),
              "iac_ac3d_sparse_batch")
