"""ctypes binding of libiac_hip.so (the C ABI declared in include/iac_hip.h).

The shared library is the product; this module only declares its symbols so
Python tests and the benchmark can call it.  Loading fails loudly when the
library has not been built: there is no CPU fallback behind it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libiac_hip.so"


class MeshShape(C.Structure):
    """ACMesh::MeshShape (IncohAutoCorrelate/ACMesh.h:21-29)."""

    _fields_ = [
        ("k_a", C.c_uint32), ("k_b", C.c_uint32), ("k_c", C.c_uint32),
        ("size_ab", C.c_uint32), ("size_c", C.c_uint32),
        ("center", C.c_uint32 * 3),
        ("dq_per_voxel", C.c_float),
        ("voxel_per_dq", C.c_uint32),
        ("max_q", C.c_float),
    ]

    def as_dict(self) -> dict:
        return {
            "k_a": self.k_a, "k_b": self.k_b, "k_c": self.k_c,
            "size_ab": self.size_ab, "size_c": self.size_c,
            "center": tuple(self.center), "dq_per_voxel": self.dq_per_voxel,
            "voxel_per_dq": self.voxel_per_dq, "max_q": self.max_q,
        }


class TuneResult(C.Structure):
    """iac_tune_result: one setting of the pair engine and its pair-kernel time."""

    _fields_ = [("pair_mode", C.c_int32), ("cell_vox", C.c_int32), ("cell_vox_v", C.c_int32),
                ("box_cap", C.c_int32), ("task_region", C.c_int32), ("frame_chunk", C.c_int32),
                ("ms", C.c_float)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Panel(C.Structure):
    """iac_panel: PPP::DetectorPanel (IncohAutoCorrelate/PPP.h:14-18)."""

    _fields_ = [("first_ind", C.c_uint32), ("fs", C.c_uint32), ("ss", C.c_uint32)]


P = C.c_void_p
F3 = C.POINTER(C.c_float)
U32 = C.c_uint32
U64 = C.c_uint64
I32 = C.c_int

# name -> (restype, argtypes); mirrors include/iac_hip.h line by line
SIGNATURES: dict[str, tuple] = {
    "iac_strerror": (C.c_char_p, [I32]),
    "iac_last_error": (C.c_char_p, []),
    "iac_abi_version": (I32, []),
    "iac_device_count": (I32, [C.POINTER(I32)]),
    "iac_ctx_create": (I32, [I32, C.POINTER(P)]),
    "iac_ctx_destroy": (I32, [P]),
    "iac_ctx_set_stream": (I32, [P, P]),
    "iac_ctx_use_own_stream": (I32, [P]),
    "iac_ctx_get_stream": (P, [P]),
    "iac_ctx_sync": (I32, [P]),
    "iac_dev_alloc": (I32, [P, C.c_size_t, C.POINTER(P)]),
    "iac_dev_free": (I32, [P, P]),
    "iac_memcpy_h2d": (I32, [P, P, P, C.c_size_t]),
    "iac_memcpy_d2h": (I32, [P, P, P, C.c_size_t]),
    "iac_memset_zero": (I32, [P, P, C.c_size_t]),
    "iac_pinned_alloc": (I32, [C.c_size_t, C.POINTER(P)]),
    "iac_pinned_free": (I32, [P]),
    "iac_ctx_enable_timing": (I32, [P, I32]),
    "iac_ctx_set_option": (I32, [P, I32, C.c_int64]),
    "iac_ctx_timing": (I32, [P, C.POINTER(C.c_double), C.POINTER(U64), I32]),
    "iac_dropped_updates": (I32, [P, C.POINTER(U64)]),
    "iac_generate_pixel_map": (I32, [I32, I32, C.c_float, F3, F3, F3, F3]),
    "iac_calc_kmap": (I32, [F3, U32, F3, F3, F3, F3]),
    "iac_mesh_shape_big": (I32, [F3, I32, C.c_float, C.POINTER(MeshShape)]),
    "iac_mesh_shape_small_cofq": (I32, [F3, I32, C.c_float, C.POINTER(MeshShape)]),
    "iac_ac3d_sparse": (I32, [P, P, U32, F3, C.POINTER(MeshShape), I32, I32, I32, P]),
    "iac_ac3d_sparse_batch": (I32, [P, P, U32, P, P, P, U32, U64, P, C.POINTER(MeshShape),
                                    I32, I32, I32, P]),
    "iac_ac3d_records_batch": (I32, [P, P, U32, P, P, U32, U64, P, C.POINTER(MeshShape), I32, I32, I32, P]),
    "iac_sparse_hit_list": (I32, [P, P, U32, P, U32, P, C.c_float, C.c_float, P, P, C.POINTER(U64)]),
    "iac_ppp_lap": (I32, [P, P, P, U32, U32, P, P, P, C.POINTER(Panel), U32, C.c_float, C.c_float, C.c_float, I32,
                          P, P]),
    "iac_ac1d_shape": (I32, [F3, U32, C.c_float, F3, F3]),
    "iac_ac1d_cq_multiplicator": (I32, [F3, U32, C.POINTER(C.c_double)]),
    "iac_ac1d_cq": (I32, [P, P, U32, P, U32, C.c_float, C.c_float, I32, C.c_double, P]),
    "iac_ac1d_acuw": (I32, [P, P, U32, P, P, P, U32, U64, U32, C.c_float, C.c_float, I32, I32, P]),
    "iac_ac1d_calc_ac": (I32, [P, P, P, U32, P]),
    "iac_cq_small_multiplicator": (I32, [F3, U32, C.POINTER(C.c_double)]),
    "iac_cq_small": (I32, [P, P, U32, P, C.POINTER(MeshShape), I32, C.c_double, P]),
    "iac_cq_dense_frames": (I32, [P, P, U32, P, U32, C.POINTER(U32), U32, C.POINTER(MeshShape), I32,
                                  C.c_double, P]),
    "iac_u64_to_f64": (I32, [P, P, C.c_size_t, C.c_double, P]),
    "iac_merge_prepare": (I32, [F3, F3, U32, C.POINTER(C.c_double), C.c_size_t, F3,
                                C.POINTER(C.c_double)]),
    "iac_merge_cq": (I32, [P, P, C.POINTER(MeshShape), P, U32, I32, C.c_double, P]),
    "iac_merge_cq_multi": (I32, [P, P, C.POINTER(MeshShape), P, U32, I32, C.c_double, P]),
    "iac_merge_ac_cq": (I32, [P, P, P, C.c_size_t, P]),
    "iac_sim_plan_create": (I32, [P, P, U32, I32, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double, P, U32,
                                  I32, C.POINTER(P)]),
    "iac_sim_plan_destroy": (I32, [P]),
    "iac_ac3d_batch_create": (I32, [P, P, U32, P, P, P, U32, U64, C.POINTER(MeshShape), I32, I32, I32,
                                    C.POINTER(P)]),
    "iac_ac3d_batch_launch": (I32, [P, P, P, I32]),
    "iac_ac3d_batch_capture": (I32, [P, P, P, I32]),
    "iac_ac3d_batch_destroy": (I32, [P]),
    "iac_ac3d_tune": (I32, [P, P, U32, P, P, P, U32, C.POINTER(MeshShape), I32, I32, I32, U32, I32,
                            C.POINTER(TuneResult), C.POINTER(TuneResult), U32, C.POINTER(U32)]),
    "iac_ac3d_queue_create": (I32, [P, P, U32, C.POINTER(MeshShape), I32, I32, I32, P, U64, U32, C.POINTER(P)]),
    "iac_ac3d_queue_push": (I32, [P, P, U32, F3]),
    "iac_ac3d_queue_push_device": (I32, [P, P, U32, F3]),
    "iac_ac3d_queue_flush": (I32, [P]),
    "iac_ac3d_queue_pending": (I32, [P, C.POINTER(U32), C.POINTER(U64)]),
    "iac_ac3d_queue_destroy": (I32, [P]),
    "iac_average_intensity": (I32, [P, P, U32, U32, P, C.c_float, C.c_float, P, P, P]),
    "iac_intensity_accumulate": (I32, [P, P, U32, U32, P, C.c_float, C.c_float, P, P, P, P]),
    "iac_intensity_finish": (I32, [P, U32, U32, C.c_float, P, P, P]),
    "iac_ac1d_acuw_records": (I32, [P, P, P, U32, U64, U32, C.c_float, C.c_float, I32, I32, P]),
    "iac_ac1d_acuw_records_noalloc": (I32, [P, P, P, U32, U64, U32, C.c_float, C.c_float, I32, I32, P]),
    "iac_sim_run": (I32, [P, P, P, U32, U32, I32, P]),
    "iac_sim_direct_fp32": (I32, [P, P, U32, I32, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float, P, U32, P,
                                  U32, U32, I32, P]),
    "iac_probe_atomics": (I32, [P, P, U64, U64, I32, U64, C.POINTER(C.c_float)]),
    "iac_probe_copy": (I32, [P, P, P, C.c_size_t, C.POINTER(C.c_float)]),
    "iac_selfcheck_sqrt": (I32, [P, U64, U64, C.POINTER(U64)]),
    "iac_sim_postprocess": (I32, [P, P, U32, U32, C.c_double, I32, C.c_float, C.c_float, C.c_uint64, U32]),
}


class IACError(RuntimeError):
    def __init__(self, code: int, where: str, detail: str):
        super().__init__(f"{where}: error {code}: {detail}")
        self.code = code


_lib = None


def lib() -> C.CDLL:
    """Load libiac_hip.so once; raise if it is missing (build it first)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: run __graft_entry__.build() (or python "
                "incohautocorrelate_amd/build.py) before using the HIP path")
        # torch ships its own HIP runtime (libamdhip64.so.7, librocblas.so.5, ...).  Loading it
        # first makes libiac_hip.so bind to those same libraries by soname, so the process
        # holds one HIP runtime whichever of the two a caller touches first.
        import torch  # noqa: F401

        lib_ = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib_, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib_
    return _lib


def check(code: int, where: str) -> None:
    if code != 0:
        L = lib()
        detail = (L.iac_last_error() or b"").decode() or (L.iac_strerror(code) or b"").decode()
        raise IACError(code, where, detail)


def fptr(arr) -> F3:
    """Pointer to the data of a contiguous float32 numpy array."""
    return arr.ctypes.data_as(F3)
