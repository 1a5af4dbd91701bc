"""Pins of the oracle functions that mass and symmetry invariants leave open.

Each pin is a plain function that raises AssertionError when the oracle breaks
it.  tests/test_oracle_pins.py runs them against the oracle, and
tests/test_oracle_slips.py runs them against builds of the oracle that each
carry one planted mistake (iac_oracle.c, ORA_SLIP).  A pin counts only if it
catches its mistake.

What each pin rests on (none of them re-types the oracle's C expression):
- two-photon frames whose q lies far from a rounding tie or bin edge, with the
  expected bins and weights computed in float64 from the geometry: the right bin
  (round vs floor), the self-pair factor (AC1D.cpp:275, 285-286), and which of the
  two (or eight) neighbours receives the fraction Sep (ACMesh.cpp:243-305,
  337-384; kernel.cl:110-183, 294-341, 735-760);
- an explicitly typed numpy float32 restatement of the AC1D CPU branch
  (AC1D.cpp:256-293), compared bit for bit: it fixes the float/double type of every
  intermediate (Sep, the split values) that a closed form within a tolerance cannot;
- active rotations: the row-major matrix of a +90 degree turn about z sends the
  +x voxel to +y (q_lab = M q_local, Settings.h:107), and a merged mesh rotated
  that way equals the identity merge turned by numpy's array operations.
"""
from __future__ import annotations

import numpy as np

import oracle as O

f32 = np.float32


def _hitlist(km, pix, vals):
    hl = np.empty((len(pix), 4), np.float32)
    hl[:, :3] = km[np.asarray(pix)]
    hl[:, 3] = vals
    return hl


def _find_pair(km, n, scale, want, start=(32, 32)):
    """Pixel b near a = start whose |k_a - k_b| * scale (float64) has a fractional part
    for which want(frac, value) holds."""
    a = start[0] * n + start[1]
    for dx in range(1, 12):
        for dy in range(0, 12):
            b = (start[0] + dx) * n + (start[1] + dy)
            q = (km[a].astype(np.float64) - km[b].astype(np.float64))
            if want(q * scale):
                return a, b
    raise AssertionError("no pixel pair with the wanted fractions")


# ---------------------------------------------------------------------------------
# AC1D AC_uw, CPU branch of Calculate_AC_UW_Mapped (AC1D.cpp:256-293)

def _ac1d_geom(size=200):
    _, km, mq = O.detector(64)
    max_q, dq = O.ac1d_shape(mq, size)
    return km, max_q, dq, size


def pin_ac1d_nn_two_photon(self_pair_factor=2):
    """Self pairs land in bin 0 with factor 2 (1 on request), the cross pair in
    bin floor(q/dq + 0.5) with factor 2 (the j<=k loop counts it once)."""
    km, max_q, dq, size = _ac1d_geom()

    def ok(q3):
        t = np.linalg.norm(q3) / dq
        fr = t - np.floor(t)
        return t > 4 and 0.6 < fr < 0.9

    a, b = _find_pair(km, 64, 1.0, ok)
    t = np.linalg.norm(km[a].astype(np.float64) - km[b]) / dq
    ca, cb = 3.0, 2.0
    ac, dropped = O.ac1d_acuw(_hitlist(km, [a, b], [ca, cb]), size, max_q, dq, interp=0,
                              self_pair_factor=self_pair_factor)
    expect = np.zeros(size)
    expect[0] = self_pair_factor * (ca * ca + cb * cb)
    expect[int(np.floor(t + 0.5))] += 2 * ca * cb
    assert dropped == 0
    assert np.array_equal(ac, expect), (np.nonzero(ac)[0], np.nonzero(expect)[0])


def pin_ac1d_linear_two_photon():
    """Linear: the lower bin floor(q/dq) gets (1 - Sep), the upper one Sep."""
    km, max_q, dq, size = _ac1d_geom()

    def ok(q3):
        t = np.linalg.norm(q3) / dq
        fr = t - np.floor(t)
        return t > 4 and 0.15 < fr < 0.35

    a, b = _find_pair(km, 64, 1.0, ok)
    t = np.linalg.norm(km[a].astype(np.float64) - km[b]) / dq
    n0, fr = int(np.floor(t)), t - np.floor(t)
    ca, cb = 3.0, 2.0
    ac, dropped = O.ac1d_acuw(_hitlist(km, [a, b], [ca, cb]), size, max_q, dq, interp=1)
    assert dropped == 0
    assert ac[0] == 2 * (ca * ca + cb * cb) and ac[1] == 0.0
    assert np.isclose(ac[n0], 2 * ca * cb * (1 - fr), rtol=1e-4)
    assert np.isclose(ac[n0 + 1], 2 * ca * cb * fr, rtol=1e-4)
    rest = np.delete(ac, [0, 1, n0, n0 + 1])
    assert not rest.any()


def ac1d_acuw_float32(hl, size, max_q, dq, interp, self_pair_factor=2):
    """AC1D.cpp:256-293 with every intermediate typed explicitly (numpy float32
    scalars; the histogram in float64, added in the loop's order)."""
    k = hl[:, :3].astype(f32)
    v = hl[:, 3].astype(f32)
    mq, d = f32(max_q), f32(dq)
    ac = np.zeros(size, np.float64)
    n = hl.shape[0]
    for j in range(n):
        for kk in range(j, n):
            dx, dy, dz = (f32(k[j, c] - k[kk, c]) for c in range(3))
            q = np.sqrt(f32(f32(f32(dx * dx) + f32(dy * dy)) + f32(dz * dz)))
            if q > mq:
                continue
            q = f32(q / d)
            two = f32(1) if (j == kk and self_pair_factor == 1) else f32(2)
            prod = f32(v[j] * v[kk])
            if interp == 0:
                sc = int(np.floor(np.float64(q) + 0.5))  # float q + double 0.5
                ac[sc] += np.float64(f32(two * prod))
            else:
                sc1 = int(np.floor(q))
                sep = f32(q - f32(sc1))
                ac[sc1] += np.float64(f32(two * f32(prod * f32(f32(1) - sep))))
                ac[sc1 + 1] += np.float64(f32(two * f32(prod * sep)))
    return ac


def pin_ac1d_float32_restatement():
    """Bit-exact against the typed restatement on random frames with float counts
    (close pairs, q/dq < 1, where float and double Sep round apart)."""
    _, km, mq = O.detector(48)
    size = 40
    max_q, dq = O.ac1d_shape(mq, size)
    rng = np.random.default_rng(5)
    for frame in range(3):
        n = 36
        pix = rng.choice(48 * 48, n, replace=False)
        # a few neighbours of the first pixels, so some pairs sit below one bin
        pix[-6:] = np.minimum(pix[:6] + 1, 48 * 48 - 1)
        pix = np.unique(pix)
        vals = rng.uniform(0.3, 3.0, pix.size).astype(np.float32)
        hl = _hitlist(km, pix, vals)
        for interp in (0, 1):
            got, _ = O.ac1d_acuw(hl, size, max_q, dq, interp=interp)
            want = ac1d_acuw_float32(hl, size, max_q, dq, interp)
            assert np.array_equal(got, want), (frame, interp, np.abs(got - want).max())


# ---------------------------------------------------------------------------------
# AC1D C(q), AutoCorr_CQ_AV (kernel.cl:656-764)

def _two_pixel_image(npix, a, b, ia, ib):
    I = np.zeros(npix, np.float32)
    I[a], I[b] = ia, ib
    return I


def pin_ac1d_cq_two_pixels():
    """Pairs (a,a), (b,b) land in bin 0; (a,b) and (b,a) in round(q/dq) (nearest)
    or split (1 - Sep) below / Sep above (linear); values truncate Val * M."""
    km, max_q, dq, size = _ac1d_geom()
    M = 1000.0
    for interp in (0, 1):
        if interp == 0:
            ok = lambda q3: (lambda t: t > 4 and 0.6 < t - np.floor(t) < 0.9)(np.linalg.norm(q3) / dq)
        else:
            ok = lambda q3: (lambda t: t > 4 and 0.15 < t - np.floor(t) < 0.35)(np.linalg.norm(q3) / dq)
        a, b = _find_pair(km, 64, 1.0, ok)
        t = np.linalg.norm(km[a].astype(np.float64) - km[b]) / dq
        I = _two_pixel_image(km.shape[0], a, b, 1.5, 2.0)
        h, dropped = O.ac1d_cq(I, km, size, max_q, dq, interp=interp, mult=M)
        assert dropped == 0
        h = h.astype(np.float64)
        assert h[0] == 2250 + 4000
        if interp == 0:
            assert h[int(np.floor(t + 0.5))] == 2 * 3000
            assert h.sum() == 6250 + 6000
        else:
            n0, fr = int(np.floor(t)), t - np.floor(t)
            assert abs(h[n0] - 2 * 3000 * (1 - fr)) <= 2
            assert abs(h[n0 + 1] - 2 * 3000 * fr) <= 2
            assert h[1] == 0


# ---------------------------------------------------------------------------------
# 3D linear corner weights: ACMesh.cpp:243-305 (first level), 337-384 (second level)

def _corner_mesh(shape, q, value, mesh=None, axes=(0, 1, 2)):
    """float64 expectation: the 8 corners floor(q) + {0,1}^3 of the voxel-space
    vector q (components along the mesh's fs, ms, ss axes), corner d gets
    value * prod(f if d else 1 - f)."""
    S, SC = shape.size_ab, shape.size_c
    cab, cc = (S - 1) // 2, (SC - 1) // 2
    if mesh is None:
        mesh = np.zeros(S * S * SC, np.float64)
    qq = np.array([q[axes[0]], q[axes[1]], q[axes[2]]], np.float64)
    base = np.floor(qq).astype(int)
    fr = qq - base
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = (fr[0] if dx else 1 - fr[0]) * (fr[1] if dy else 1 - fr[1]) * (fr[2] if dz else 1 - fr[2])
                fs, ms, ss = base[0] + dx + cab, base[1] + dy + cab, base[2] + dz + cc
                mesh[fs + S * ms + S * S * ss] += value * w
    return mesh


def _lin_pair(km, vox):
    """A pair whose scaled x and y fractions are far from 1/2 (a corner swap moves
    at least 40 % of the weight)."""
    def ok(q3):
        fr = q3 - np.floor(q3)
        return abs(q3[0]) > 1 and abs(q3[1]) > 0.5 and abs(fr[0] - 0.5) > 0.2 and abs(fr[1] - 0.5) > 0.2

    return _find_pair(km, 64, float(vox), ok)


def pin_ac3d_linear_two_photon():
    """All three modes with a linear level put 1000 * c_a * c_b * weight on the eight
    corners around +q and around -q, and 1000 * c^2 on the centre for each self pair."""
    _, km, mq = O.detector(64)
    s = O.mesh_shape_big(mq, 31)
    vox = s.voxel_per_dq
    a, b = _lin_pair(km, vox)
    ca, cb = 3.0, 2.0
    hl = _hitlist(km, [a, b], [ca, cb])
    q = (km[a].astype(np.float64) - km[b]) * vox
    expect = _corner_mesh(s, q, 1000 * ca * cb)
    _corner_mesh(s, -q, 1000 * ca * cb, expect)
    c = s.center[0] + s.size_ab * s.center[1] + s.size_ab ** 2 * s.center[2]
    expect[c] += 1000 * (ca * ca + cb * cb)
    for i1, i2, dm in ((0, 1, False), (1, 0, True), (1, 1, True)):
        got, dropped = O.ac3d_sparse(hl, s, None, i1, i2, dm)
        assert dropped == 0
        err = np.abs(got.astype(np.float64) - expect)
        assert err.max() <= 2, ((i1, i2, dm), err.max())


# ---------------------------------------------------------------------------------
# C(q) small linear (kernel.cl:110-183)

def pin_cq_small_linear_two_pixels():
    _, km, mq = O.detector(64)
    s = O.mesh_shape_small_cofq(mq, 31)
    vox = s.voxel_per_dq
    a, b = _lin_pair(km, vox)
    I = _two_pixel_image(km.shape[0], a, b, 3.0, 2.0)
    M = 1000.0
    got, dropped = O.cq_small(I, km, s, 1, M)
    assert dropped == 0
    q = (km[a].astype(np.float64) - km[b]) * vox
    axes = (s.k_a, s.k_b, s.k_c)
    expect = _corner_mesh(s, q, 6 * M, axes=axes)
    _corner_mesh(s, -q, 6 * M, expect, axes=axes)
    c = s.center[0] + s.size_ab * s.center[1] + s.size_ab ** 2 * s.center[2]
    expect[c] += (9 + 4) * M
    err = np.abs(got.astype(np.float64) - expect)
    assert err.max() <= 2, err.max()


# ---------------------------------------------------------------------------------
# Merge_CQ rotation (kernel.cl:258-262) and linear weights (kernel.cl:294-341)

def _rot_z(deg):
    t = np.deg2rad(deg)
    return np.array([[np.cos(t), -np.sin(t), 0], [np.sin(t), np.cos(t), 0], [0, 0, 1]])


def _rot_x(deg):
    t = np.deg2rad(deg)
    return np.array([[1, 0, 0], [0, np.cos(t), -np.sin(t)], [0, np.sin(t), np.cos(t)]])


def _rot_y(deg):
    t = np.deg2rad(deg)
    return np.array([[np.cos(t), 0, np.sin(t)], [0, 1, 0], [-np.sin(t), 0, np.cos(t)]])


def _rw(R, w=1.0):
    rw = np.zeros((1, 10), np.float32)
    rw[0, :9] = np.round(np.asarray(R), 12).reshape(9)
    rw[0, 9] = w
    return rw


def _small_shape():
    _, _, mq = O.detector(64)
    return O.mesh_shape_small_cofq(mq, 15)


def _small_index(s, q_lab):
    """Index of the small-mesh voxel holding the lab-frame voxel vector q_lab."""
    S, SC = s.size_ab, s.size_c
    cab, cc = (S - 1) // 2, (SC - 1) // 2
    fs, ms, ss = q_lab[s.k_a] + cab, q_lab[s.k_b] + cab, q_lab[s.k_c] + cc
    return fs + S * ms + S * S * ss


def pin_merge_quarter_turns():
    """+90 degrees about z sends +x to +y, about x sends +y to +z, about y sends
    +z to +x (active rotations, q_lab = M q_local)."""
    s = _small_shape()
    S = s.size_ab
    cab = (S - 1) // 2
    cases = [(_rot_z(90), (3, 0, 0), (0, 3, 0)), (_rot_z(90), (0, 2, 0), (-2, 0, 0)),
             (_rot_x(90), (0, 2, 0), (0, 0, 2)), (_rot_y(90), (0, 0, 1), (1, 0, 0)),
             (_rot_z(-90), (3, 1, 0), (1, -3, 0))]
    for R, qin, qout in cases:
        small = np.zeros(S * S * s.size_c)
        small[_small_index(s, qin)] = 5.0
        big, dropped = O.merge_cq(small, s, _rw(R), 0, 1.0)
        assert dropped == 0
        v = (qout[0] + cab) + S * (qout[1] + cab) + S * S * (qout[2] + cab)
        assert big[v] == 5 and big.sum() == 5, (qin, qout, np.nonzero(big)[0], v)


def pin_merge_rotation_permutes_mesh():
    """A random small mesh merged under +90 degrees about z equals the identity
    merge with the array turned: new[z, y, x] = old[z, c - (x - c), y]."""
    s = _small_shape()
    S = s.size_ab
    rng = np.random.default_rng(2)
    small = rng.integers(1, 100, S * S * s.size_c).astype(np.float64)
    eye, _ = O.merge_cq(small, s, _rw(np.eye(3)), 0, 1.0)
    rot, dropped = O.merge_cq(small, s, _rw(_rot_z(90)), 0, 1.0)
    assert dropped == 0
    old = eye.reshape(S, S, S)
    assert np.array_equal(rot.reshape(S, S, S), old[:, ::-1, :].transpose(0, 2, 1))
    rot_t, _ = O.merge_cq(small, s, _rw(_rot_z(90).T), 0, 1.0)
    assert not np.array_equal(rot_t, rot)


def pin_merge_linear_single_voxel():
    """A 20 degree turn about z of one small voxel: the four corners around R q_in
    in the q_z = 0 plane, floor corner weight (1 - f), weight w and M applied."""
    s = _small_shape()
    S = s.size_ab
    small = np.zeros(S * S * s.size_c)
    qin = np.array([4, 1, 0])
    small[_small_index(s, qin)] = 7.0
    R = _rot_z(20)
    rw = _rw(R, 0.5)
    M = 1000.0
    big, dropped = O.merge_cq(small, s, rw, 1, M)
    assert dropped == 0
    q = rw[0, :9].astype(np.float64).reshape(3, 3) @ qin
    cube = O.MeshShape.from_any(s)
    cube.size_c = S
    expect = _corner_mesh(cube, q, 7.0 * M * 0.5)
    err = np.abs(big.astype(np.float64) - expect)
    assert err.max() <= 2, err.max()


ALL_PINS = {
    "ac1d_nn_two_photon": pin_ac1d_nn_two_photon,
    "ac1d_nn_two_photon_factor1": lambda: pin_ac1d_nn_two_photon(1),
    "ac1d_linear_two_photon": pin_ac1d_linear_two_photon,
    "ac1d_float32_restatement": pin_ac1d_float32_restatement,
    "ac1d_cq_two_pixels": pin_ac1d_cq_two_pixels,
    "ac3d_linear_two_photon": pin_ac3d_linear_two_photon,
    "cq_small_linear_two_pixels": pin_cq_small_linear_two_pixels,
    "merge_quarter_turns": pin_merge_quarter_turns,
    "merge_rotation_permutes_mesh": pin_merge_rotation_permutes_mesh,
    "merge_linear_single_voxel": pin_merge_linear_single_voxel,
}

# planted mistake (iac_oracle.c ORA_SLIP) -> the pins that must catch it
SLIP_CAUGHT_BY = {
    1: ["ac1d_nn_two_photon", "ac1d_float32_restatement"],
    2: ["ac1d_nn_two_photon", "ac1d_float32_restatement"],
    3: ["ac1d_float32_restatement"],
    4: ["ac3d_linear_two_photon"],
    5: ["ac3d_linear_two_photon"],
    6: ["cq_small_linear_two_pixels"],
    7: ["merge_quarter_turns", "merge_rotation_permutes_mesh"],
    8: ["merge_linear_single_voxel"],
    9: ["ac1d_cq_two_pixels"],
    10: ["ac1d_cq_two_pixels"],
}
