"""MI355X-native pixel-pair autocorrelation path of IncohAutoCorrelate.

The compute lives in libiac_hip.so (hand-written gfx950 HIP kernels behind the
C ABI of include/iac_hip.h).  This module is the Python view of that ABI used
by the tests and the benchmark: device buffers are torch tensors on the GPU
(plumbing only), work is enqueued on torch's current HIP stream.
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from ._abi import IACError, MeshShape, Panel, TuneResult, check, fptr, lib

__all__ = [
    "IACError", "MeshShape", "Panel", "lib", "generate_pixel_map", "calc_kmap", "mesh_shape_big",
    "mesh_shape_small_cofq", "Context", "SimPlan", "NEAREST", "LINEAR",
]

NEAREST = 0
LINEAR = 1

# iac_ctx_set_option keys (include/iac_hip.h)
OPT_AC3D_KERNEL = 1
OPT_CELL_VOX = 2
OPT_FRAME_CHUNK = 3
OPT_SIM_CHUNK_ROWS = 4
OPT_PAIR_PREFETCH = 5
OPT_BOX_CAP = 6
OPT_CELL_VOX_V = 7
OPT_TASK_REGION = 8
OPT_PAIR_MODE = 9
OPT_LAP_GROUPS = 10
SIM_FP32_DIRECT, SIM_BF16X3, SIM_FP16X3, SIM_BF16, SIM_FP16 = 0, 1, 2, 3, 4
SIM_AMPLITUDE, SIM_INTENSITY = 0, 1


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


def cq_small_multiplicator(intensity: np.ndarray) -> float:
    """Detector::AutoCorrelate_CofQ_SmallMesh multiplicator (Detector.cpp:1215-1227)."""
    I = _f32(intensity)
    m = C.c_double()
    check(lib().iac_cq_small_multiplicator(fptr(I), I.size, C.byref(m)), "iac_cq_small_multiplicator")
    return m.value


def merge_prepare(rot9: np.ndarray, mean_intensity: np.ndarray, small_cq: np.ndarray):
    """Detector::Merge_smallCofQ host preparation (Detector.cpp:1339-1427): (rw10, mult)."""
    R = _f32(rot9).reshape(-1, 9)
    mi = _f32(mean_intensity)
    cq = np.ascontiguousarray(small_cq, dtype=np.float64)
    rw = np.empty((R.shape[0], 10), np.float32)
    m = C.c_double()
    check(lib().iac_merge_prepare(fptr(R), fptr(mi), R.shape[0], cq.ctypes.data_as(C.POINTER(C.c_double)),
                                  cq.size, fptr(rw), C.byref(m)), "iac_merge_prepare")
    return rw, m.value


def ac1d_shape(max_q, size: int, q_zoom: float = 1.0) -> tuple[float, float]:
    """AC1D::Initialize (AC1D.cpp:53-58): (Max_Q, dq_per_Step)."""
    mq = _f32(max_q, 3)
    a, b = np.zeros(1, np.float32), np.zeros(1, np.float32)
    check(lib().iac_ac1d_shape(fptr(mq), int(size), float(q_zoom), fptr(a), fptr(b)), "iac_ac1d_shape")
    return float(a[0]), float(b[0])


def ac1d_cq_multiplicator(intensity: np.ndarray) -> float:
    I = _f32(intensity)
    m = C.c_double()
    check(lib().iac_ac1d_cq_multiplicator(fptr(I), I.size, C.byref(m)), "iac_ac1d_cq_multiplicator")
    return m.value


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
        self._plans: weakref.WeakSet = weakref.WeakSet()
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
            for p in list(self._plans):  # plans hold device memory of this context
                p.destroy()
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

    def ac3d_records_batch(self, kmap, rec4, frame_off, n_total: int, shape: MeshShape, mesh, rots=None,
                           interp1: int = NEAREST, interp2: int = NEAREST, double_map: bool = True) -> None:
        nframes = frame_off.shape[0] - 1
        check(lib().iac_ac3d_records_batch(self._h, _ptr(kmap), kmap.shape[0], _ptr(rec4), _ptr(frame_off),
                                           nframes, int(n_total), _ptr(rots), C.byref(shape), interp1, interp2,
                                           int(bool(double_map)), _ptr(mesh)), "iac_ac3d_records_batch")

    def sparse_hit_list(self, kmap, frames, mask, threshold: float, step: float, rec4, frame_off) -> int:
        """CreateSparseHitList over dense frames [nframes, npix]; returns the record count."""
        n = C.c_uint64()
        check(lib().iac_sparse_hit_list(self._h, _ptr(kmap), kmap.shape[0], _ptr(frames), frames.shape[0],
                                        _ptr(mask), float(threshold), float(step), _ptr(rec4), _ptr(frame_off),
                                        C.byref(n)), "iac_sparse_hit_list")
        return int(n.value)

    # ---- AC1D --------------------------------------------------------------------
    def ac1d_cq(self, kmap, intensity, size: int, max_q: float, dq: float, interp: int, mult: float, hist):
        check(lib().iac_ac1d_cq(self._h, _ptr(kmap), kmap.shape[0], _ptr(intensity), int(size), float(max_q),
                                float(dq), interp, float(mult), _ptr(hist)), "iac_ac1d_cq")

    def ac1d_acuw(self, kmap, pix, cnt, frame_off, n_total: int, size: int, max_q: float, dq: float,
                  interp: int, hist, self_pair_factor: int = 2):
        check(lib().iac_ac1d_acuw(self._h, _ptr(kmap), kmap.shape[0], _ptr(pix), _ptr(cnt), _ptr(frame_off),
                                  frame_off.shape[0] - 1, int(n_total), int(size), float(max_q), float(dq), interp,
                                  int(self_pair_factor), _ptr(hist)), "iac_ac1d_acuw")

    def ac1d_calc_ac(self, ac_uw, cq, out):
        check(lib().iac_ac1d_calc_ac(self._h, _ptr(ac_uw), _ptr(cq), ac_uw.numel(), _ptr(out)), "iac_ac1d_calc_ac")

    # ---- C(q) ------------------------------------------------------------------
    def cq_small(self, kmap, intensity, shape: MeshShape, mesh, interp: int, mult: float) -> None:
        """AutoCorr_CQ_small over one intensity image into the uint64 small mesh."""
        check(lib().iac_cq_small(self._h, _ptr(kmap), kmap.shape[0], _ptr(intensity), C.byref(shape),
                                 interp, float(mult), _ptr(mesh)), "iac_cq_small")

    def cq_dense_frames(self, kmap, frames, group_off, shape: MeshShape, meshes, interp: int,
                        mult: float) -> None:
        """Dense frames [nframes, npix]; frames of group g accumulate into meshes[g]."""
        goff = np.ascontiguousarray(group_off, dtype=np.uint32)
        check(lib().iac_cq_dense_frames(self._h, _ptr(kmap), kmap.shape[0], _ptr(frames), frames.shape[0],
                                        goff.ctypes.data_as(C.POINTER(C.c_uint32)), goff.size - 1,
                                        C.byref(shape), interp, float(mult), _ptr(meshes)),
              "iac_cq_dense_frames")

    def u64_to_f64(self, src, mult: float, out) -> None:
        check(lib().iac_u64_to_f64(self._h, _ptr(src), src.numel(), float(mult), _ptr(out)), "iac_u64_to_f64")

    def merge_cq(self, small_cq, shape: MeshShape, rw10, interp: int, mult: float, big) -> None:
        """Merge_CQ: rotate and weight the small C(q) into the cubic big mesh."""
        check(lib().iac_merge_cq(self._h, _ptr(small_cq), C.byref(shape), _ptr(rw10), rw10.shape[0], interp,
                                 float(mult), _ptr(big)), "iac_merge_cq")

    def merge_cq_multi(self, small_meshes, shape: MeshShape, rw10, interp: int, mult: float, big) -> None:
        """Merge_CQ with one small mesh per event (meshes[i] rotated by event i)."""
        check(lib().iac_merge_cq_multi(self._h, _ptr(small_meshes), C.byref(shape), _ptr(rw10), rw10.shape[0],
                                       interp, float(mult), _ptr(big)), "iac_merge_cq_multi")

    def merge_ac_cq(self, ac, cq, out) -> None:
        check(lib().iac_merge_ac_cq(self._h, _ptr(ac), _ptr(cq), ac.numel(), _ptr(out)), "iac_merge_ac_cq")

    def ac1d_acuw_records(self, rec4, frame_off, n_total: int, size: int, max_q: float, dq: float, interp: int,
                          hist, self_pair_factor: int = 2) -> None:
        nframes = frame_off.numel() - 1
        check(lib().iac_ac1d_acuw_records(self._h, _ptr(rec4), _ptr(frame_off), nframes, n_total, size, max_q, dq,
                                          interp, self_pair_factor, _ptr(hist)), "iac_ac1d_acuw_records")

    def average_intensity(self, frames, mask, threshold: float, step: float, avg, mean_intensity=None,
                          photon_count=None) -> None:
        """Detector::LoadAndAverageIntensity over [nframes][npix] device frames."""
        nframes, npix = frames.shape
        check(lib().iac_average_intensity(self._h, _ptr(frames), nframes, npix, _ptr(mask), threshold, step,
                                          _ptr(avg), _ptr(mean_intensity), _ptr(photon_count)),
              "iac_average_intensity")

    def ppp_lap(self, frames, out, panels, adu_per_photon: float = 1.0, seed_threshold: float = 0.5,
                combined_threshold: float = 0.9, mask=None, offset=None, gain=None, gain_only: bool = False,
                photon_count=None, mean_intensity=None) -> None:
        """PPP::ProcessData_PF_LAP's per-frame chain (mask, gain correction, LAP photon finder)
        over [nframes][npix] device frames into `out` (which may be `frames`); panels is a
        list of (first_ind, fs, ss), PPP::DetectorPanel (IAC/PPP.h:14-18)."""
        nframes, npix = frames.shape
        tab = (Panel * max(1, len(panels)))(*[Panel(*map(int, q)) for q in panels])
        check(lib().iac_ppp_lap(self._h, _ptr(frames), _ptr(out), nframes, npix, _ptr(mask), _ptr(offset),
                                _ptr(gain), tab, len(panels), adu_per_photon, seed_threshold, combined_threshold,
                                int(gain_only), _ptr(photon_count), _ptr(mean_intensity)), "iac_ppp_lap")

    def ac3d_batch(self, kmap, pix, cnt, frame_off, n_total: int, shape: MeshShape, interp1: int = NEAREST,
                   interp2: int = NEAREST, double_map: bool = True) -> "Ac3dBatch":
        """A prepared, replayable frame batch (identity rotation) on the LDS engine."""
        h = C.c_void_p()
        nframes = frame_off.numel() - 1
        check(lib().iac_ac3d_batch_create(self._h, _ptr(kmap), kmap.numel() // 3, _ptr(pix), _ptr(cnt),
                                          _ptr(frame_off), nframes, n_total, C.byref(shape), interp1, interp2,
                                          int(double_map), C.byref(h)), "iac_ac3d_batch_create")
        b = Ac3dBatch(self, h, (kmap, pix, cnt, frame_off))
        self._plans.add(b)
        return b

    def ac3d_tune(self, kmap, pix, cnt, frame_off, shape: MeshShape, interp1: int = NEAREST, interp2: int = NEAREST,
                  double_map: bool = True, prefix_frames: int = 1000, force: bool = False) -> tuple[dict, list]:
        """The library's autotuner on a frame prefix; applies the winner to this context.
        Returns (winner, every candidate) as dicts (ms = -1 for a cached winner)."""
        best = TuneResult()
        allr = (TuneResult * 32)()
        n = C.c_uint32()
        check(lib().iac_ac3d_tune(self._h, _ptr(kmap), kmap.numel() // 3, _ptr(pix), _ptr(cnt), _ptr(frame_off),
                                  frame_off.numel() - 1, C.byref(shape), interp1, interp2, int(double_map),
                                  int(prefix_frames), int(force), C.byref(best), allr, 32, C.byref(n)),
              "iac_ac3d_tune")
        return best.as_dict(), [allr[i].as_dict() for i in range(n.value)]

    def ac3d_queue(self, kmap, shape: MeshShape, mesh, interp1: int = NEAREST, interp2: int = NEAREST,
                   double_map: bool = True, max_records: int = 0, max_frames: int = 0) -> "Ac3dQueue":
        """A frame queue: Detector::AutoCorrelateSparseList once per event, binned in batches."""
        h = C.c_void_p()
        check(lib().iac_ac3d_queue_create(self._h, _ptr(kmap), kmap.numel() // 3, C.byref(shape), interp1, interp2,
                                          int(double_map), _ptr(mesh), int(max_records), int(max_frames),
                                          C.byref(h)), "iac_ac3d_queue_create")
        q = Ac3dQueue(self, h, (kmap, mesh))
        self._plans.add(q)
        return q

    # ---- measurement probes ----
    def probe_atomics(self, mesh, n_updates: int, mode: int, hot: int = 0) -> float:
        ms = C.c_float()
        check(lib().iac_probe_atomics(self._h, _ptr(mesh), mesh.numel(), n_updates, mode, hot, C.byref(ms)),
              "iac_probe_atomics")
        return ms.value

    def probe_copy(self, src, dst) -> float:
        ms = C.c_float()
        nbytes = src.numel() * src.element_size()
        check(lib().iac_probe_copy(self._h, _ptr(src), _ptr(dst), nbytes, C.byref(ms)), "iac_probe_copy")
        return ms.value

    def selfcheck_sqrt(self, lo: int = 0, hi: int = 1 << 32) -> int:
        """Mismatches of the AC1D kernels' short square root against sqrtf over the
        float bit patterns [lo, hi)."""
        n = C.c_uint64()
        check(lib().iac_selfcheck_sqrt(self._h, int(lo), int(hi), C.byref(n)), "iac_selfcheck_sqrt")
        return n.value

    # ---- simulator (Simulator::SimulatePart, IAC/Simulator.cpp:662-944) ----
    def sim_plan(self, pixmap, emitters, wavelength: float, mode: int = SIM_BF16X3, subsampling: int = 0,
                 u_step=(0.0, 0.0, 0.0), v_step=(0.0, 0.0, 0.0)) -> "SimPlan":
        """Phase matrix of a detector (float [npix][3] device tensor) and an
        emitter set (double [nem][3] device tensor) for the GEMM simulator."""
        us = (C.c_double * 3)(*u_step)
        vs = (C.c_double * 3)(*v_step)
        h = C.c_void_p()
        npix = pixmap.numel() // 3
        nem = emitters.numel() // 3
        check(lib().iac_sim_plan_create(self._h, _ptr(pixmap), npix, subsampling, us, vs, wavelength,
                                        _ptr(emitters), nem, mode, C.byref(h)), "iac_sim_plan_create")
        plan = SimPlan(self, h, npix, nem)
        self._plans.add(plan)
        return plan

    def sim_run(self, plan: "SimPlan", phases, out, modes: int = 1, out_kind: int = SIM_AMPLITUDE) -> None:
        """phases: double [nframes * modes][nem]; out: float [nframes][npix]."""
        nframes = phases.numel() // (plan.nem * modes)
        check(lib().iac_sim_run(self._h, plan.handle, _ptr(phases), nframes, modes, out_kind, _ptr(out)),
              "iac_sim_run")

    def sim_direct_fp32(self, pixmap, emitters, phases, wavelength: float, out, modes: int = 1,
                        out_kind: int = SIM_AMPLITUDE, subsampling: int = 0, u_step=(0.0, 0.0, 0.0),
                        v_step=(0.0, 0.0, 0.0)) -> None:
        """The reference SimulateCrystal arithmetic in FP32: float emitters [nem][3],
        float phases [nframes * modes][nem] -> out float [nframes][npix]."""
        us = (C.c_float * 3)(*u_step)
        vs = (C.c_float * 3)(*v_step)
        npix = pixmap.numel() // 3
        nem = emitters.numel() // 3
        nframes = phases.numel() // (nem * modes)
        check(lib().iac_sim_direct_fp32(self._h, _ptr(pixmap), npix, subsampling, us, vs, wavelength,
                                        _ptr(emitters), nem, _ptr(phases), nframes, modes, out_kind, _ptr(out)),
              "iac_sim_direct_fp32")

    def sim_postprocess(self, frames, n_expected: float, poisson: bool = True, adu_noise: float = 0.0,
                        value_per_photon: float = 1.0, seed: int = 0, frame0: int = 0) -> None:
        """Rescale / Poisson / noise / value-per-photon of float [nframes][npix] in place."""
        nframes, npix = frames.shape
        check(lib().iac_sim_postprocess(self._h, _ptr(frames), npix, nframes, n_expected, int(poisson),
                                        adu_noise, value_per_photon, seed, frame0), "iac_sim_postprocess")


class Ac3dBatch:
    """Owns an iac_ac3d_batch: launch() enqueues one pass, capture() records it as a graph.
    The batch reads its input tensors by pointer at every launch, so it keeps them alive."""

    def __init__(self, ctx: Context, handle: C.c_void_p, inputs: tuple):
        self.ctx, self.handle, self._inputs = ctx, handle, inputs

    def launch(self, mesh, zero_mesh: bool = False) -> None:
        check(lib().iac_ac3d_batch_launch(self.ctx.handle, self.handle, _ptr(mesh), int(zero_mesh)),
              "iac_ac3d_batch_launch")

    def capture(self, mesh, zero_mesh: bool = False) -> None:
        check(lib().iac_ac3d_batch_capture(self.ctx.handle, self.handle, _ptr(mesh), int(zero_mesh)),
              "iac_ac3d_batch_capture")

    def destroy(self) -> None:
        if self.handle:
            lib().iac_ac3d_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Ac3dQueue:
    """Owns an iac_ac3d_queue; destroy() flushes, synchronises and frees it."""

    def __init__(self, ctx: Context, handle: C.c_void_p, keep: tuple):
        self.ctx, self.handle, self._keep = ctx, handle, keep

    def push(self, hitlist: np.ndarray, rot=None) -> None:
        """One frame's host hit list {kx,ky,kz,value} x n (float32) and rotation."""
        hl = np.ascontiguousarray(hitlist, np.float32).reshape(-1, 4)
        r = _f32(rot if rot is not None else np.eye(3, dtype=np.float32), 9)
        check(lib().iac_ac3d_queue_push(self.handle, hl.ctypes.data if len(hl) else None, len(hl), fptr(r)),
              "iac_ac3d_queue_push")

    def push_device(self, hitlist, rot=None) -> None:
        """One frame's device hit list ([n][4] float32 tensor) and rotation."""
        n = hitlist.shape[0] if hitlist is not None else 0
        r = _f32(rot if rot is not None else np.eye(3, dtype=np.float32), 9)
        check(lib().iac_ac3d_queue_push_device(self.handle, _ptr(hitlist) if n else None, n, fptr(r)),
              "iac_ac3d_queue_push_device")

    def flush(self) -> None:
        check(lib().iac_ac3d_queue_flush(self.handle), "iac_ac3d_queue_flush")

    def pending(self) -> tuple[int, int]:
        f, r = C.c_uint32(), C.c_uint64()
        check(lib().iac_ac3d_queue_pending(self.handle, C.byref(f), C.byref(r)), "iac_ac3d_queue_pending")
        return int(f.value), int(r.value)

    def destroy(self) -> None:
        if self.handle:
            h, self.handle = self.handle, None
            check(lib().iac_ac3d_queue_destroy(h), "iac_ac3d_queue_destroy")

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class SimPlan:
    """Owns an iac_sim_plan (device phase matrix); destroy() or garbage collection frees it."""

    def __init__(self, ctx: Context, handle: C.c_void_p, npix: int, nem: int):
        self.ctx, self.handle, self.npix, self.nem = ctx, handle, npix, nem

    def destroy(self) -> None:
        if self.handle:
            lib().iac_sim_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
