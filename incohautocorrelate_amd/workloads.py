"""Seeded synthetic inputs of the benchmark workloads (numpy only).

Geometry follows Simulator::GeneratePixelMap (Simulator.cpp:511-519): flat
detector, VecA=(1,0,0), VecB=(0,1,0), 75 um pixels, centre at z = 0.5 m
(coordinates in um).  Photon frames are stored as (pixel index uint32, count
uint16) lists sorted by pixel index -- the order Detector::CreateSparseHitList
(Detector.cpp:926-948) produces -- with frame f owning entries
[frame_off[f], frame_off[f+1]).
"""
from __future__ import annotations

import math

import numpy as np

PIXEL_UM = 75.0
DIST_UM = 500000.0
WAVELENGTH_A = 1.5


def detector(n: int):
    """(pixmap [n*n,3], kmap [n*n,3], max_q [3]) through the product's host geometry."""
    from . import calc_kmap, generate_pixel_map

    pm = generate_pixel_map(n, n, PIXEL_UM, [0.0, 0.0, DIST_UM], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0])
    km, _, _, mq = calc_kmap(pm)
    return pm, km, mq


def _distinct_positions(rng, npix: int, k: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Per frame f, k[f] distinct uniform pixel indices, sorted; returns (pix, frame_off)."""
    nf = k.size
    want = k.astype(np.int64)
    keys = np.empty(0, np.int64)
    have = np.zeros(nf, np.int64)
    for _ in range(64):
        deficit = want - have
        if not deficit.any():
            break
        fr = np.repeat(np.arange(nf, dtype=np.int64), deficit)
        new = fr * npix + rng.integers(0, npix, size=fr.size, dtype=np.int64)
        keys = np.unique(np.concatenate([keys, new]))
        have = np.bincount(keys // npix, minlength=nf).astype(np.int64)
    off = np.zeros(nf + 1, np.uint64)
    off[1:] = np.cumsum(have).astype(np.uint64)
    return (keys % npix).astype(np.uint32), off


def _zt_poisson(rng, lam: float, size: int) -> np.ndarray:
    """Zero-truncated Poisson(lam) counts by inversion of the CDF."""
    ks = np.arange(0, 64)
    logp = -lam + ks * math.log(lam) - np.array([math.lgamma(k + 1) for k in ks])
    cdf = np.cumsum(np.exp(logp))
    u = rng.uniform(cdf[0], 1.0, size=size)
    c = np.searchsorted(cdf, np.minimum(u, cdf[-1]))
    return np.clip(c, 1, 65535).astype(np.uint16)


def poisson_photon_frames(npix: int, nframes: int, mean: float, seed: int):
    """Frames of i.i.d. Poisson(mean) pixels, kept as their nonzero entries."""
    rng = np.random.default_rng(seed)
    k = rng.binomial(npix, -math.expm1(-mean), size=nframes)
    pix, off = _distinct_positions(rng, npix, k)
    cnt = _zt_poisson(rng, mean, pix.size)
    return {"pix": pix, "cnt": cnt, "frame_off": off}


def dense_poisson_frames(npix: int, nframes: int, mean: float, seed: int):
    """Dense uint32 Poisson frames [nframes, npix] and their photon lists."""
    rng = np.random.default_rng(seed)
    frames = rng.poisson(mean, size=(nframes, npix)).astype(np.uint32)
    f, p = np.nonzero(frames)
    off = np.zeros(nframes + 1, np.uint64)
    off[1:] = np.cumsum(np.bincount(f, minlength=nframes)).astype(np.uint64)
    cnt = np.minimum(frames[f, p], 65535).astype(np.uint16)
    return {"frames": frames, "pix": p.astype(np.uint32), "cnt": cnt, "frame_off": off}


def lognormal_photon_frames(npix: int, nframes: int, median: float, sigma: float, lo: int, hi: int,
                            seed: int):
    """Uneven frames: photons/frame ~ lognormal(median, sigma) clipped to [lo, hi],
    at distinct uniform pixels, count 1 each."""
    rng = np.random.default_rng(seed)
    n = np.clip(np.rint(rng.lognormal(math.log(median), sigma, size=nframes)), lo, hi).astype(np.int64)
    pix, off = _distinct_positions(rng, npix, n)
    return {"pix": pix, "cnt": np.ones(pix.size, np.uint16), "frame_off": off, "sigma": sigma}


def gamma_frames(npix: int, nframes: int, shape: float, mean: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.gamma(shape, mean / shape, size=(nframes, npix)).astype(np.float32)


def random_rotations(n: int, seed: int) -> np.ndarray:
    """n uniform random rotations (Shoemake's unit quaternions) as row-major 3x3 floats."""
    rng = np.random.default_rng(seed)
    u1, u2, u3 = rng.random((3, n))
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    x, y = a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2)
    z, w = b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R.reshape(n, 9).astype(np.float32)


def emitters_cube(n: int, edge_a: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.uniform(-edge_a / 2, edge_a / 2, size=(n, 3))


def emitter_phases(nframes: int, nem: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 2 * np.pi, size=(nframes, nem))
