"""Pins of the oracle's 3D AC_uw path (Detector.cpp:1010-1051, ACMesh.cpp:209-385).

Each check is chosen so that a plausible restatement mistake fails it:
- the OpenMP i<=j-plus-mirror loop equals a plain serial loop over all ordered
  pairs (a wrong mirror, a double-counted self pair, a lost pair);
- a two-photon frame has the closed-form mesh (wrong scale, wrong axis order);
- a 90 degree rotation about z permutes the identity mesh exactly the way the
  row-major convention q_lab = M q_local prescribes (a transposed matrix would
  rotate the other way);
- identity-R double mapping equals single nearest-neighbour binning;
- total mass equals the fixed-point value summed over the pairs inside Max_Q
  (counted independently with numpy float32), and NN meshes are point
  symmetric about the centre voxel.
"""
import numpy as np
import pytest

import oracle as O
from incohautocorrelate_amd import workloads as W

f32 = np.float32
MODES = [(i1, i2, dm) for i1 in (0, 1) for i2 in (0, 1) for dm in (True, False)]


def hitlist(km, pix, vals):
    hl = np.empty((len(pix), 4), np.float32)
    hl[:, :3] = km[np.asarray(pix)]
    hl[:, 3] = vals
    return hl


@pytest.mark.parametrize("i1,i2,dm", MODES)
def test_parallel_loop_equals_serial_ordered_pairs(i1, i2, dm):
    _, km, mq = O.detector(20)
    s = O.mesh_shape_big(mq, 15)
    rng = np.random.default_rng(10 + 4 * i1 + 2 * i2 + dm)
    rots = [None] + list(W.random_rotations(2, 1))
    for R in rots:
        n = int(rng.integers(1, 60))
        hl = hitlist(km, rng.choice(400, n, replace=False), rng.integers(1, 4, n))
        m1, d1 = O.ac3d_sparse(hl, s, R, i1, i2, dm)
        m2, d2 = O.ac3d_bruteforce(hl, s, R, i1, i2, dm)
        assert np.array_equal(m1, m2) and d1 == d2


def test_two_photon_closed_form():
    n = 64
    _, km, mq = O.detector(n)
    s = O.mesh_shape_big(mq, 31)
    S, C, vox = s.size_ab, s.center[0], s.voxel_per_dq
    a, b = 10 * n + 20, 13 * n + 21  # pixel (10,20) and (13,21): x = slow index, y = fast index
    ca, cb = 3, 2
    hl = hitlist(km, [a, b], [ca, cb])
    mesh, dropped = O.ac3d_sparse(hl, s)
    q = (km[a] - km[b]).astype(np.float64) * vox
    assert (np.abs(np.abs(q - np.floor(q)) - 0.5) > 1e-3).all()  # far from rounding ties
    d = np.round(q).astype(int)
    # small-angle geometry: q ~ (x_a - x_b, y_a - y_b, 0) / L with x = slow index * 75 um
    small_angle = np.round(np.array([-3 * 75.0, -1 * 75.0, 0.0]) / 500000.0 * vox).astype(int)
    assert np.array_equal(d, small_angle) and d[0] != 0
    expect = np.zeros(S ** 3, np.uint64)
    ctr = C + C * S + C * S * S
    expect[ctr] += 1000 * (ca * ca + cb * cb)
    for sgn in (1, -1):
        v = (C + sgn * d[0]) + (C + sgn * d[1]) * S + (C + sgn * d[2]) * S * S
        expect[v] += 1000 * ca * cb
    assert dropped == 0
    assert np.array_equal(mesh, expect)


def test_axis_order_fast_scan_is_x():
    """A displacement along the detector's x axis only moves the fast-scan (fs) index."""
    n = 64
    _, km, mq = O.detector(n)
    s = O.mesh_shape_big(mq, 31)
    S, C = s.size_ab, s.center[0]
    hl = hitlist(km, [30 * n + 30, 40 * n + 30], [1, 1])  # same y, 10 pixels apart in x
    mesh, _ = O.ac3d_sparse(hl, s)
    nz = np.nonzero(mesh)[0]
    fs, ms, ss = nz % S, (nz // S) % S, nz // (S * S)
    assert set(ms) == {C} and set(ss) == {C} and len(set(fs)) == 3


def test_rotation_row_major_90deg_about_z():
    _, km, mq = O.detector(48)
    s = O.mesh_shape_big(mq, 31)
    S, C = s.size_ab, s.center[0]
    rng = np.random.default_rng(3)
    hl = hitlist(km, rng.choice(48 * 48, 80, replace=False), rng.integers(1, 3, 80))
    m0, _ = O.ac3d_sparse(hl, s, None, 0, 0, True)
    R = np.array([0, -1, 0, 1, 0, 0, 0, 0, 1], np.float32)  # q_lab = (-q_y, q_x, q_z)
    m1, _ = O.ac3d_sparse(hl, s, R, 0, 0, True)
    g0 = m0.reshape(S, S, S)  # [ss][ms][fs]
    g1 = m1.reshape(S, S, S)
    # g1[z, y', x'] with x' = C - (y - C), y' = x
    idx = np.arange(S)
    expect = np.zeros_like(g0)
    for x in idx:
        for y in idx:
            xp, yp = 2 * C - y, x
            if 0 <= xp < S:
                expect[:, yp, xp] += g0[:, y, x]
    assert np.array_equal(g1, expect)


def test_rotation_row_major_90deg_about_x():
    """About x the mapping (q_x, -q_z, q_y) and its transpose (q_x, q_z, -q_y)
    differ on a mesh that is only point symmetric, so this pins row-major
    storage of RotMatrix (IAC/Settings.h:107)."""
    _, km, mq = O.detector(48)
    s = O.mesh_shape_big(mq, 31)
    S, C = s.size_ab, s.center[0]
    rng = np.random.default_rng(13)
    hl = hitlist(km, rng.choice(48 * 48, 80, replace=False), rng.integers(1, 3, 80))
    g0 = O.ac3d_sparse(hl, s, None, 0, 0, True)[0].reshape(S, S, S)  # [z][y][x]
    R = np.array([1, 0, 0, 0, 0, -1, 0, 1, 0], np.float32)
    g1 = O.ac3d_sparse(hl, s, R, 0, 0, True)[0].reshape(S, S, S)
    # new (x, y, z) = (x, -z, y): g1[C+y, C-z, x] = g0[C+z, C+y, x]
    expect = np.zeros_like(g0)
    for z in range(S):
        for y in range(S):
            expect[y, 2 * C - z, :] += g0[z, y, :]
    assert np.array_equal(g1, expect)
    g2 = O.ac3d_sparse(hl, s, R.reshape(3, 3).T.copy().reshape(9), 0, 0, True)[0].reshape(S, S, S)
    assert not np.array_equal(g2, g1)


def test_identity_double_map_equals_single_nn():
    _, km, mq = O.detector(40)
    s = O.mesh_shape_big(mq, 21)
    rng = np.random.default_rng(4)
    hl = hitlist(km, rng.choice(1600, 90, replace=False), rng.integers(1, 5, 90))
    assert np.array_equal(O.ac3d_sparse(hl, s, None, 0, 0, True)[0],
                          O.ac3d_sparse(hl, s, None, 0, 0, False)[0])


@pytest.mark.parametrize("rot_seed", [None, 5])
def test_mass_and_point_symmetry_nn(rot_seed):
    n = 96
    _, km, mq = O.detector(n)
    s = O.mesh_shape_big(mq, 41)
    rng = np.random.default_rng(6)
    pix = rng.choice(n * n, 150, replace=False)
    c = rng.integers(1, 4, 150)
    hl = hitlist(km, pix, c)
    R = None if rot_seed is None else W.random_rotations(1, rot_seed)[0]
    mesh, dropped = O.ac3d_sparse(hl, s, R, 0, 0, True)
    assert dropped == 0
    # independent count of ordered pairs inside Max_Q (numpy float32, correctly rounded sqrt)
    k = hl[:, :3]
    q = (k[:, None, :] - k[None, :, :]).astype(np.float32)
    s2 = (q[..., 0] * q[..., 0] + q[..., 1] * q[..., 1]) + q[..., 2] * q[..., 2]
    inside = ~(np.sqrt(s2.astype(np.float32)) > f32(s.max_q))
    vals = 1000 * np.outer(c, c)
    assert int(mesh.sum()) == int(vals[inside].sum())
    assert inside.sum() < inside.size  # some diagonal pairs are cut (Euclidean cut vs axis Max_Q)
    assert np.array_equal(mesh, mesh[::-1])  # mesh[C+d] == mesh[C-d]


def test_linear_modes_conserve_mass_within_rounding():
    _, km, mq = O.detector(40)
    s = O.mesh_shape_big(mq, 21)
    rng = np.random.default_rng(8)
    n = 70
    hl = hitlist(km, rng.choice(1600, n, replace=False), rng.integers(1, 4, n))
    k = hl[:, :3]
    q = (k[:, None, :] - k[None, :, :]).astype(np.float32)
    s2 = (q[..., 0] * q[..., 0] + q[..., 1] * q[..., 1]) + q[..., 2] * q[..., 2]
    inside = ~(np.sqrt(s2) > f32(s.max_q))
    total = int((1000 * np.outer(hl[:, 3], hl[:, 3]))[inside].sum())
    for i1, i2, dm in [(0, 1, True), (1, 0, True), (1, 1, True), (0, 1, False)]:
        mesh, dropped = O.ac3d_sparse(hl, s, None, i1, i2, dm)
        ncorner = (8 if i1 == 1 and dm else 1) * (8 if i2 == 1 else 1)
        assert dropped == 0
        assert abs(int(mesh.sum()) - total) <= n * n * ncorner  # each corner rounds by <= 1/2 unit


def test_out_of_mesh_updates_are_counted():
    """A non-orthonormal 'rotation' pushes pairs outside the mesh; the oracle
    drops and counts them (the reference would write out of bounds)."""
    _, km, mq = O.detector(32)
    s = O.mesh_shape_big(mq, 15)
    hl = hitlist(km, [0, 31, 31 * 32, 1023], [1, 1, 1, 1])
    R = np.diag([3.0, 3.0, 1.0]).astype(np.float32).reshape(9)
    mesh, dropped = O.ac3d_sparse(hl, s, R)
    assert dropped > 0
    assert int(mesh.sum()) // 1000 + dropped <= 16
