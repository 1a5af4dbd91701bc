"""ctypes view of liboracle.so -- the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, and only as the checker.  See iac_oracle.h for what each function
restates and where in the reference it comes from.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "liboracle.so"


class MeshShape(C.Structure):
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

    @classmethod
    def from_any(cls, s) -> "MeshShape":
        d = s.as_dict() if hasattr(s, "as_dict") else dict(s)
        o = cls()
        for k, v in d.items():
            if k == "center":
                o.center[:] = list(v)
            else:
                setattr(o, k, v)
        return o

    @property
    def nvox(self) -> int:
        return self.size_ab * self.size_ab * self.size_c


P = C.c_void_p
F = C.POINTER(C.c_float)
D = C.POINTER(C.c_double)
U64P = C.POINTER(C.c_uint64)
U32 = C.c_uint32
I32 = C.c_int
SP = C.POINTER(MeshShape)

_SIG = {
    "ora_float_to_int_f": (C.c_int64, [C.c_float]),
    "ora_float_to_int_d": (C.c_int64, [C.c_double]),
    "ora_int_to_float_u64": (C.c_float, [C.c_uint64]),
    "ora_pair_val": (C.c_uint32, [C.c_float]),
    "ora_generate_pixel_map": (None, [I32, I32, C.c_float, F, F, F, F]),
    "ora_calc_kmap": (None, [F, U32, F, F, F, F]),
    "ora_mesh_shape_big": (I32, [F, I32, C.c_float, SP]),
    "ora_mesh_shape_small_cofq": (I32, [F, I32, C.c_float, SP]),
    "ora_sparse_hit_list": (U32, [F, P, F, U32, C.c_float, C.c_float, F]),
    "ora_hit_list_from_photons": (None, [F, P, P, U32, F]),
    "ora_ac3d_sparse": (None, [F, U32, F, SP, I32, I32, I32, U64P, U64P, I32]),
    "ora_ac3d_sparse_frames": (None, [F, P, P, U64P, U32, P, SP, I32, I32, I32, U64P, U64P, I32]),
    "ora_ac3d_bruteforce": (None, [F, U32, F, SP, I32, I32, I32, U64P, U64P]),
    "ora_cq_small_multiplicator": (C.c_double, [F, U32]),
    "ora_cq_small": (None, [F, F, U32, SP, I32, C.c_double, U64P, U64P, I32]),
    "ora_merge_prepare": (C.c_double, [F, F, U32, D, C.c_size_t, F]),
    "ora_merge_cq": (None, [D, SP, F, U32, I32, C.c_double, U64P, U64P, I32]),
    "ora_ac1d_shape": (None, [F, U32, C.c_float, F, F]),
    "ora_ac1d_cq_multiplicator": (C.c_double, [F, U32]),
    "ora_ac1d_cq": (None, [F, F, U32, U32, C.c_float, C.c_float, I32, C.c_double, U64P, U64P, I32]),
    "ora_ac1d_acuw": (None, [F, U32, U32, C.c_float, C.c_float, I32, I32, D, U64P]),
    "ora_ac1d_calc_ac": (None, [D, D, U32, D]),
    "ora_merge_ac_cq": (None, [U64P, D, C.c_size_t, D]),
    "ora_cq_small_f64": (None, [F, F, U32, SP, D, I32]),
    "ora_merge_f64": (None, [D, SP, F, D]),
    "ora_simulate_fp64": (None, [F, U32, U32, U32, D, U32, D, U32, U32, C.c_double, I32, D, D, I32,
                                 D, I32]),
    "ora_philox4x32_10": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "ora_uniform2": (None, [C.c_uint64, U32, U32, U32, D]),
    "ora_exp": (C.c_double, [C.c_double]),
    "ora_log": (C.c_double, [C.c_double]),
    "ora_cos": (C.c_double, [C.c_double]),
    "ora_poisson_sample": (C.c_uint32, [C.c_double, D]),
    "ora_average_intensity": (None, [F, U32, U32, C.POINTER(C.c_int), C.c_float, C.c_float, F, F,
                                     C.POINTER(C.c_int)]),
    "ora_sim_postprocess": (None, [F, U32, C.c_double, I32, C.c_float, C.c_float, C.c_uint64, U32]),
    "ora_cvtt_i32": (C.c_int, [C.c_float]),
    "ora_ppp_lap_frame": (C.c_float, [F, U32, C.POINTER(C.c_int), F, F, P, U32, C.c_float, C.c_float, C.c_float,
                                      I32]),
}

_lib = None


def _open(path) -> C.CDLL:
    L = C.CDLL(str(path))
    for k, (r, a) in _SIG.items():
        fn = getattr(L, k)
        fn.restype = r
        fn.argtypes = a
    return L


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle` or __graft_entry__.build()")
        _lib = _open(LIB_PATH)
    return _lib


class use_library:
    """Context manager: every wrapper below calls the oracle built at `path`
    (a build with a planted mistake, -DORA_SLIP=n) until the block ends."""

    def __init__(self, path):
        self.path = path

    def __enter__(self):
        global _lib
        self.saved = lib()
        _lib = _open(self.path)
        return _lib

    def __exit__(self, *exc):
        global _lib
        _lib = self.saved
        return False


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def fp(a):
    return a.ctypes.data_as(F)


def dp(a):
    return a.ctypes.data_as(D)


def up(a):
    return a.ctypes.data_as(U64P)


IDENTITY = np.eye(3, dtype=np.float32).reshape(9)


# ---- geometry ----------------------------------------------------------------
def generate_pixel_map(size_a, size_b, pixel_size, center, vec_a, vec_b):
    out = np.empty((size_a * size_b, 3), np.float32)
    lib().ora_generate_pixel_map(size_a, size_b, pixel_size, fp(_f(center)), fp(_f(vec_a)),
                                 fp(_f(vec_b)), fp(out))
    return out


def calc_kmap(pixmap):
    pm = _f(pixmap).reshape(-1, 3)
    km = np.empty_like(pm)
    mx, mn, mq = (np.empty(3, np.float32) for _ in range(3))
    lib().ora_calc_kmap(fp(pm), pm.shape[0], fp(km), fp(mx), fp(mn), fp(mq))
    return km, mx, mn, mq


def mesh_shape_big(max_q, edge, zoom=1.0):
    s = MeshShape()
    r = lib().ora_mesh_shape_big(fp(_f(max_q)), edge, zoom, C.byref(s))
    if r != 0:
        raise ValueError("mesh size must be odd")
    return s


def mesh_shape_small_cofq(max_q, perp, zoom=1.0):
    s = MeshShape()
    r = lib().ora_mesh_shape_small_cofq(fp(_f(max_q)), perp, zoom, C.byref(s))
    if r != 0:
        raise ValueError("mesh size must be odd")
    return s


def detector(n, pixel_um=75.0, dist_um=500000.0):
    """Flat detector of the benchmark configs: (pixmap, kmap, max_q)."""
    pm = generate_pixel_map(n, n, pixel_um, [0, 0, dist_um], [1, 0, 0], [0, 1, 0])
    km, _, _, mq = calc_kmap(pm)
    return pm, km, mq


# ---- sparse hit list -----------------------------------------------------------
def sparse_hit_list(intensity, kmap, threshold=0.0, step=1.0, mask=None):
    I = _f(intensity).reshape(-1)
    km = _f(kmap).reshape(-1, 3)
    out = np.empty((I.size, 4), np.float32)
    m = None
    if mask is not None:
        m = np.ascontiguousarray(mask, dtype=np.int32)
    n = lib().ora_sparse_hit_list(fp(I), m.ctypes.data if m is not None else None, fp(km), I.size,
                                  threshold, step, fp(out))
    return out[:n].copy()


def hit_list_from_photons(kmap, pix, cnt):
    pix = np.ascontiguousarray(pix, np.uint32)
    cnt = np.ascontiguousarray(cnt, np.uint16)
    out = np.empty((pix.size, 4), np.float32)
    lib().ora_hit_list_from_photons(fp(_f(kmap)), pix.ctypes.data, cnt.ctypes.data, pix.size, fp(out))
    return out


# ---- 3D AC_uw ------------------------------------------------------------------
def ac3d_sparse(hl, shape, rot=None, interp1=0, interp2=0, double_map=True, mesh=None, nthreads=0):
    s = MeshShape.from_any(shape)
    hl = _f(hl).reshape(-1, 4)
    if mesh is None:
        mesh = np.zeros(s.nvox, np.uint64)
    dropped = np.zeros(1, np.uint64)
    R = _f(IDENTITY if rot is None else rot).reshape(9)
    lib().ora_ac3d_sparse(fp(hl), hl.shape[0], fp(R), C.byref(s), interp1, interp2, int(double_map),
                          up(mesh), up(dropped), nthreads)
    return mesh, int(dropped[0])


def ac3d_bruteforce(hl, shape, rot=None, interp1=0, interp2=0, double_map=True):
    s = MeshShape.from_any(shape)
    hl = _f(hl).reshape(-1, 4)
    mesh = np.zeros(s.nvox, np.uint64)
    dropped = np.zeros(1, np.uint64)
    R = _f(IDENTITY if rot is None else rot).reshape(9)
    lib().ora_ac3d_bruteforce(fp(hl), hl.shape[0], fp(R), C.byref(s), interp1, interp2,
                              int(double_map), up(mesh), up(dropped))
    return mesh, int(dropped[0])


def ac3d_frames(kmap, pix, cnt, frame_off, shape, rots=None, interp1=0, interp2=0, double_map=True,
                mesh=None, nthreads=0):
    s = MeshShape.from_any(shape)
    pix = np.ascontiguousarray(pix, np.uint32)
    cnt = np.ascontiguousarray(cnt, np.uint16)
    off = np.ascontiguousarray(frame_off, np.uint64)
    if mesh is None:
        mesh = np.zeros(s.nvox, np.uint64)
    dropped = np.zeros(1, np.uint64)
    r = None if rots is None else _f(rots).reshape(-1, 9)
    lib().ora_ac3d_sparse_frames(fp(_f(kmap)), pix.ctypes.data, cnt.ctypes.data, up(off), off.size - 1,
                                 r.ctypes.data if r is not None else None, C.byref(s), interp1,
                                 interp2, int(double_map), up(mesh), up(dropped), nthreads)
    return mesh, int(dropped[0])


def pair_val(p: float) -> int:
    return int(lib().ora_pair_val(float(p)))


# ---- C(q) ------------------------------------------------------------------------
def cq_small_multiplicator(intensity):
    return float(lib().ora_cq_small_multiplicator(fp(_f(intensity)), np.asarray(intensity).size))


def cq_small(intensity, kmap, shape, interp=0, mult=1.0, nthreads=0, mesh=None):
    s = MeshShape.from_any(shape)
    I = _f(intensity).reshape(-1)
    if mesh is None:
        mesh = np.zeros(s.nvox, np.uint64)
    dropped = np.zeros(1, np.uint64)
    lib().ora_cq_small(fp(I), fp(_f(kmap)), I.size, C.byref(s), interp, mult, up(mesh), up(dropped),
                       nthreads)
    return mesh, int(dropped[0])


def merge_prepare(rot9, mean_intensity, small_cq):
    R = _f(rot9).reshape(-1, 9)
    mi = _f(mean_intensity)
    cq = np.ascontiguousarray(small_cq, np.float64)
    rw = np.empty((R.shape[0], 10), np.float32)
    m = lib().ora_merge_prepare(fp(R), fp(mi), R.shape[0], dp(cq), cq.size, fp(rw))
    return rw, float(m)


def merge_cq(small_cq, small_shape, rw10, interp=0, mult=1.0, nthreads=0):
    s = MeshShape.from_any(small_shape)
    cq = np.ascontiguousarray(small_cq, np.float64)
    rw = _f(rw10).reshape(-1, 10)
    big = np.zeros(s.size_ab ** 3, np.uint64)
    dropped = np.zeros(1, np.uint64)
    lib().ora_merge_cq(dp(cq), C.byref(s), fp(rw), rw.shape[0], interp, mult, up(big), up(dropped), nthreads)
    return big, int(dropped[0])


def merge_ac_cq(ac, cq):
    ac = np.ascontiguousarray(ac, np.uint64)
    cq = np.ascontiguousarray(cq, np.float64)
    out = np.empty(ac.size, np.float64)
    lib().ora_merge_ac_cq(up(ac), dp(cq), ac.size, dp(out))
    return out


# ---- AC1D ------------------------------------------------------------------------
def ac1d_shape(max_q, size, zoom=1.0):
    mq, dq = np.zeros(1, np.float32), np.zeros(1, np.float32)
    lib().ora_ac1d_shape(fp(_f(max_q)), size, zoom, fp(mq), fp(dq))
    return float(mq[0]), float(dq[0])


def ac1d_cq_multiplicator(intensity):
    return float(lib().ora_ac1d_cq_multiplicator(fp(_f(intensity)), np.asarray(intensity).size))


def ac1d_cq(intensity, kmap, size, max_q, dq, interp=1, mult=1.0, nthreads=0):
    I = _f(intensity).reshape(-1)
    hist = np.zeros(size, np.uint64)
    dropped = np.zeros(1, np.uint64)
    lib().ora_ac1d_cq(fp(I), fp(_f(kmap)), I.size, size, max_q, dq, interp, mult, up(hist), up(dropped),
                      nthreads)
    return hist, int(dropped[0])


def ac1d_acuw(hl, size, max_q, dqdx, interp=1, self_pair_factor=2, ac=None):
    hl = _f(hl).reshape(-1, 4)
    if ac is None:
        ac = np.zeros(size, np.float64)
    dropped = np.zeros(1, np.uint64)
    lib().ora_ac1d_acuw(fp(hl), hl.shape[0], size, max_q, dqdx, interp, self_pair_factor, dp(ac), up(dropped))
    return ac, int(dropped[0])


def ac1d_calc_ac(ac_uw, cq):
    a = np.ascontiguousarray(ac_uw, np.float64)
    c = np.ascontiguousarray(cq, np.float64)
    out = np.empty(a.size, np.float64)
    lib().ora_ac1d_calc_ac(dp(a), dp(c), a.size, dp(out))
    return out


def cq_small_f64(intensity, kmap, shape, nthreads=0, mesh=None):
    s = MeshShape.from_any(shape)
    I = _f(intensity).reshape(-1)
    if mesh is None:
        mesh = np.zeros(s.nvox, np.float64)
    lib().ora_cq_small_f64(fp(I), fp(_f(kmap)), I.size, C.byref(s), dp(mesh), nthreads)
    return mesh


def merge_f64(small_cq, small_shape, rw10, big=None):
    s = MeshShape.from_any(small_shape)
    cq = np.ascontiguousarray(small_cq, np.float64)
    if big is None:
        big = np.zeros(s.size_ab ** 3, np.float64)
    lib().ora_merge_f64(dp(cq), C.byref(s), fp(_f(rw10).reshape(10)), dp(big))
    return big


# ---- simulator -----------------------------------------------------------------
def simulate_fp64(pixmap, emitters, phases, wavelength, subsampling=0, u_step=(0, 0, 0), v_step=(0, 0, 0),
                  out_kind=0, modes=1, pixels=None, frames=None, nthreads=0):
    """FP64 speckle sum.  phases is [nframes * modes][nem] (mode-minor); the
    result [nf][np] sums the subpixel-averaged |psi| (out_kind 0) or |psi|^2
    (out_kind 1) over the modes of each frame.  pixels / frames = (start, count)."""
    pm = _f(pixmap).reshape(-1, 3)
    em = np.ascontiguousarray(emitters, np.float64).reshape(-1, 3)
    ph = np.ascontiguousarray(phases, np.float64).reshape(-1, em.shape[0])
    nfr = ph.shape[0] // modes
    p0, npx = pixels if pixels is not None else (0, pm.shape[0])
    f0, nf = frames if frames is not None else (0, nfr)
    us = np.ascontiguousarray(u_step, np.float64)
    vs = np.ascontiguousarray(v_step, np.float64)
    out = np.empty((nf * modes, npx), np.float64)
    lib().ora_simulate_fp64(fp(pm), pm.shape[0], p0, npx, dp(em), em.shape[0], dp(ph), f0 * modes, nf * modes,
                            wavelength, subsampling, dp(us), dp(vs), out_kind, dp(out), nthreads)
    return out.reshape(nf, modes, npx).sum(axis=1)


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().ora_philox4x32_10(c, k, o)
    return list(o)


def uniform2(seed, frame, pixel, stream=0):
    u = np.empty(2, np.float64)
    lib().ora_uniform2(seed, frame, pixel, stream, dp(u))
    return u


def poisson_sample(mean, u):
    return int(lib().ora_poisson_sample(float(mean), dp(np.ascontiguousarray(u, np.float64))))


def sim_postprocess(intensity, n_expected, do_poisson=True, adu_noise=0.0, value_per_photon=1.0, seed=0,
                    frame0=0):
    """Rescale + Poisson + noise + value per photon on [nframes][npix] float
    frames (a copy); frame f uses counter frame0 + f."""
    I = np.array(intensity, np.float32, copy=True, order="C")
    I2 = I.reshape(-1, I.shape[-1])
    for f in range(I2.shape[0]):
        row = I2[f]
        lib().ora_sim_postprocess(fp(row), row.size, n_expected, int(do_poisson), adu_noise, value_per_photon,
                                  seed, frame0 + f)
    return I


# ---- averaged intensity --------------------------------------------------------
def average_intensity(frames, threshold, step, mask=None):
    """(avg [npix], mean_intensity [nframes], photon_count [nframes] or -1 where not set)."""
    fr = _f(frames)
    nframes, npix = fr.shape
    m = None if mask is None else np.ascontiguousarray(mask, np.int32)
    avg = np.empty(npix, np.float32)
    mi = np.empty(nframes, np.float32)
    pc = np.full(nframes, -1, np.int32)
    lib().ora_average_intensity(fp(fr), nframes, npix, None if m is None else m.ctypes.data_as(C.POINTER(C.c_int)),
                                threshold, step, fp(avg), fp(mi), pc.ctypes.data_as(C.POINTER(C.c_int)))
    return avg, mi, pc


# ---- pattern pre-processing: LAP photon finder -----------------------------------
class Panel(C.Structure):
    """PPP::DetectorPanel (IAC/PPP.h:14-18): first pixel, fs (row width), ss (rows)."""

    _fields_ = [("first_ind", C.c_uint32), ("fs", C.c_uint32), ("ss", C.c_uint32)]


def ppp_lap(frames, panels, adu=1.0, seed=0.5, comb=0.9, mask=None, offset=None, gain=None, gain_only=False):
    """PPP::ProcessData_PF_LAP's per-frame chain on [nframes][npix] float frames (a copy).
    panels: list of (first_ind, fs, ss).  Returns (frames, photon_count, mean_intensity)."""
    out = np.array(frames, np.float32, copy=True, order="C")
    out2 = out.reshape(-1, out.shape[-1])
    nframes, npix = out2.shape
    tab = (Panel * max(1, len(panels)))(*[Panel(*map(int, q)) for q in panels])
    m = None if mask is None else np.ascontiguousarray(mask, np.int32)
    o = None if offset is None else _f(offset)
    g = None if gain is None else _f(gain)
    pc = np.empty(nframes, np.int32)
    mi = np.empty(nframes, np.float32)
    for f in range(nframes):
        row = out2[f]
        s = lib().ora_ppp_lap_frame(fp(row), npix, None if m is None else m.ctypes.data_as(C.POINTER(C.c_int)),
                                    None if o is None else fp(o), None if g is None else fp(g),
                                    C.cast(tab, P), len(panels), adu, seed, comb, int(gain_only))
        pc[f] = lib().ora_cvtt_i32(s)
        mi[f] = np.float32(pc[f]) / np.float32(npix)
    return out, pc, mi


def cvtt_i32(x):
    """(int)x of the reference's x86-64 build (cvttss2si)."""
    return int(lib().ora_cvtt_i32(float(np.float32(x))))
