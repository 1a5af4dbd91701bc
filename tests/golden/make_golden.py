#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from the oracle, after checking
each against an independent computation:
  ac3d_*   : the OpenMP restatement of the CPU path == the serial brute force
             over ordered pairs (every mode, a rotation);
  cq_*     : flat-field C(q) counts are integers (mult 1) and sum to the pair
             count within the cut; the random-intensity mesh is stored as is;
  merge_*  : an identity-rotation, weight-1 merge holds the small mesh's mass;
  ac1d_*   : the linear split keeps the nearest-neighbour mass up to the
             per-pair fixed-point rounding;
  sim_*    : FP64 field == numpy complex exponentials;
  post_*   : Poisson / noise post-processing (Philox stream, series functions).
The reference ships no fixtures or vectors for this path; these freeze the
oracle's outputs so the GPU tests can be checked without the oracle library.
Run: python tests/golden/make_golden.py   (writes golden.npz beside this file)
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

import oracle as O  # noqa: E402

from incohautocorrelate_amd import workloads as W  # noqa: E402


def sparse(m):
    nz = np.nonzero(m)[0].astype(np.uint32)
    return nz, m[nz]


def main() -> None:
    out: dict[str, np.ndarray] = {}
    rng = np.random.default_rng(2024)

    # ---- 3D AC_uw: 16x16 detector, three frames, 17^3 mesh
    n = 16
    _, km, mq = W.detector(n)
    shape = O.mesh_shape_big(mq, 15)
    data = W.poisson_photon_frames(n * n, 3, 0.06, 11)
    data["cnt"][::5] = 3
    rot = W.random_rotations(1, 3)[0]
    out["ac3d_kmap"] = km
    out["ac3d_pix"], out["ac3d_cnt"], out["ac3d_off"] = data["pix"], data["cnt"], data["frame_off"]
    out["ac3d_rot"] = rot
    out["ac3d_shape"] = np.array([shape.size_ab, shape.size_c], np.uint32)
    for name, (i1, i2, dm, R) in {"nn": (0, 0, 1, None), "lin": (1, 1, 1, None), "nn_lin_single": (0, 1, 0, None),
                                  "lin_nn": (1, 0, 1, None), "nn_rot": (0, 0, 1, rot)}.items():
        mesh, drop = O.ac3d_frames(km, data["pix"], data["cnt"], data["frame_off"], shape,
                                   None if R is None else np.tile(R, (3, 1)), i1, i2, bool(dm))
        brute = np.zeros_like(mesh)
        off = data["frame_off"]
        for f in range(3):
            hl = O.hit_list_from_photons(km, data["pix"][off[f]:off[f + 1]], data["cnt"][off[f]:off[f + 1]])
            b, _ = O.ac3d_bruteforce(hl, shape, R, i1, i2, bool(dm))
            brute += b
        assert np.array_equal(mesh, brute), name
        assert drop == 0
        out[f"ac3d_{name}_idx"], out[f"ac3d_{name}_val"] = sparse(mesh)

    # ---- C(q) small: flat field and random intensity, 14x14 detector
    n = 14
    _, km, mq = W.detector(n)
    ss = O.mesh_shape_small_cofq(mq, 21)
    flat = np.ones(n * n, np.float32)
    cqf, _ = O.cq_small(flat, km, ss, 0, 1.0)
    k = km.astype(np.float64)
    r = np.sqrt(((k[:, None, :] - k[None, :, :]) ** 2).sum(-1))
    # every pair well inside the cut counts once, none well outside (float cut at the surface)
    assert (r < ss.max_q * (1 - 1e-4)).sum() <= cqf.sum() <= (r < ss.max_q * (1 + 1e-4)).sum()
    I = rng.gamma(1.0, 0.5, n * n).astype(np.float32)
    mult = O.cq_small_multiplicator(I)
    cqr, _ = O.cq_small(I, km, ss, 1, mult)
    out["cq_kmap"], out["cq_I"] = km, I
    out["cq_shape"] = np.array([ss.size_ab, ss.size_c], np.uint32)
    out["cq_mult"] = np.array([mult])
    out["cq_flat_idx"], out["cq_flat_val"] = sparse(cqf)
    out["cq_lin_idx"], out["cq_lin_val"] = sparse(cqr)

    # ---- merge of the random small mesh with 5 events
    small = cqr.astype(np.float64) / mult
    rots = W.random_rotations(5, 17)
    mi = rng.gamma(2.0, 0.3, 5).astype(np.float32)
    rw, m2 = O.merge_prepare(rots, mi, small)
    big, _ = O.merge_cq(small, ss, rw, 0, m2)
    eye = np.zeros((1, 10), np.float32)
    eye[0, [0, 4, 8, 9]] = 1.0
    ident, _ = O.merge_cq(small, ss, eye, 0, 1e6)
    # identity rotation, weight 1: the big mesh holds the small one's mass
    assert abs(ident.astype(np.float64).sum() / 1e6 - small.sum()) <= 1e-6 * small.sum() + small.size * 1e-6
    out["merge_rots"], out["merge_mi"], out["merge_rw"] = rots, mi, rw
    out["merge_mult"] = np.array([m2])
    out["merge_idx"], out["merge_val"] = sparse(big)

    # ---- AC1D
    n = 20
    _, km, mq = W.detector(n)
    max_q, dq = O.ac1d_shape(mq, 64)
    I = rng.gamma(1.0, 1.0, n * n).astype(np.float32)
    m1 = O.ac1d_cq_multiplicator(I)
    hist, _ = O.ac1d_cq(I, km, 64, max_q, dq, 1, m1)
    hist_nn, _ = O.ac1d_cq(I, km, 64, max_q, dq, 0, m1)
    assert abs(int(hist.sum()) - int(hist_nn.sum())) <= 2 * (n * n) ** 2
    out["ac1d_kmap"], out["ac1d_I"] = km, I
    out["ac1d_par"] = np.array([max_q, dq, m1])
    out["ac1d_lin"], out["ac1d_nn"] = hist, hist_nn

    # ---- simulator FP64 field
    pm, _, _ = W.detector(8)
    em = W.emitters_cube(7, 200.0, 5)
    ph = W.emitter_phases(2, 7, 6)
    f64 = O.simulate_fp64(pm, em, ph, 1.5, out_kind=1)
    kk = (2 * np.pi / 1.5) * pm.astype(np.float64) / np.linalg.norm(pm.astype(np.float64), axis=1, keepdims=True)
    ref = np.abs(np.exp(1j * (kk @ em.T)[None] + 1j * ph[:, None, :]).sum(-1)) ** 2
    assert np.allclose(f64, ref, rtol=1e-12, atol=1e-10)
    out["sim_pm"], out["sim_em"], out["sim_ph"], out["sim_int"] = pm, em, ph, f64

    # ---- post-processing
    I = rng.gamma(1.0, 1.0, (2, 1500)).astype(np.float32)
    out["post_in"] = I
    out["post_poisson"] = O.sim_postprocess(I, 30.0, True, 0.0, 1.0, seed=99, frame0=4)
    out["post_noise"] = O.sim_postprocess(I, 30.0, True, 0.5, 2.0, seed=99, frame0=4)

    np.savez_compressed(HERE / "golden.npz", **out)
    print(f"wrote {HERE / 'golden.npz'} ({(HERE / 'golden.npz').stat().st_size} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()
