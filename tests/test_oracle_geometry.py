"""Geometry pins: pixel map, k-map, mesh shapes.

Closed forms (Simulator.cpp:511-519 puts pixel a at (a - SizeA/2)*PS, i.e.
half a pixel off centre), the unit-norm k-map (Detector.cpp:310-361), and the
mesh shapes at the benchmark geometries (ACMesh.cpp:74-206) whose values were
computed independently from the same formulas.  The product's host geometry
(libiac_hip.so, loadable without a GPU) must equal the oracle bit for bit.
"""
import numpy as np
import pytest

import oracle as O


def test_pixel_map_closed_form():
    n = 128
    pm = O.generate_pixel_map(n, n, 75.0, [0, 0, 500000], [1, 0, 0], [0, 1, 0]).reshape(n, n, 3)
    a = (np.arange(n) - n / 2) * 75.0
    assert np.array_equal(pm[:, :, 0], np.broadcast_to(a[:, None], (n, n)).astype(np.float32))
    assert np.array_equal(pm[:, :, 1], np.broadcast_to(a[None, :], (n, n)).astype(np.float32))
    assert (pm[:, :, 2] == 500000.0).all()
    # off-centre by half a pixel: no pixel at x = 0.5*PS*(n-1)
    assert pm[:, :, 0].min() == -64 * 75.0 and pm[:, :, 0].max() == 63 * 75.0


def test_pixel_map_tilted_vectors_normalised():
    pm = O.generate_pixel_map(4, 3, 100.0, [1, 2, 3], [2, 0, 0], [0, 0, 5]).reshape(4, 3, 3)
    # VecA/VecB are normalised: steps are exactly one pixel size along x and z
    assert np.allclose(np.diff(pm[:, 0, 0]), 100.0)
    assert np.allclose(np.diff(pm[0, :, 2]), 100.0)


def test_kmap_unit_and_direction():
    pm, km, mq = O.detector(64)
    nrm = np.linalg.norm(km.astype(np.float64), axis=1)
    assert np.abs(nrm - 1).max() < 3e-7
    ref = pm.astype(np.float64) / np.linalg.norm(pm.astype(np.float64), axis=1)[:, None]
    assert np.abs(km - ref).max() < 2e-7
    assert np.allclose(mq, np.abs(km.max(0) - km.min(0)))


@pytest.mark.parametrize("n,edge,S,vpdq", [(128, 63, 65, 1574), (1024, 255, 257, 823),
                                           (256, 127, 129, 1621)])
def test_big_mesh_shapes(n, edge, S, vpdq):
    _, _, mq = O.detector(n)
    s = O.mesh_shape_big(mq, edge)
    assert (s.size_ab, s.size_c, s.voxel_per_dq) == (S, S, vpdq)
    assert tuple(s.center) == ((S - 1) // 2,) * 3
    assert s.max_q == max(mq)
    # Voxel_per_dq = floor((C-2)/Max_Q): every |q| <= Max_Q maps within C-2 voxels of the
    # centre, so nearest-neighbour binning of cut-passing pairs stays inside the padding
    assert s.max_q * s.voxel_per_dq <= s.center[0] - 2
    assert s.max_q * (s.voxel_per_dq + 1) > s.center[0] - 2


@pytest.mark.parametrize("n,perp,sc", [(128, 63, 3), (1024, 255, 13), (256, 127, 3)])
def test_small_cofq_shapes(n, perp, sc):
    _, _, mq = O.detector(n)
    s = O.mesh_shape_small_cofq(mq, perp)
    assert s.size_c == sc and s.size_ab == perp + 2
    assert (s.k_a, s.k_c) == (0, 2)  # x/y extents tie, z is smallest


def test_even_mesh_rejected():
    _, _, mq = O.detector(16)
    with pytest.raises(ValueError):
        O.mesh_shape_big(mq, 64)


def test_product_host_geometry_matches_oracle():
    import incohautocorrelate_amd as iac

    for n in (16, 128, 1024):
        pm1 = iac.generate_pixel_map(n, n, 75.0, [0, 0, 500000], [1, 0, 0], [0, 1, 0])
        pm2 = O.generate_pixel_map(n, n, 75.0, [0, 0, 500000], [1, 0, 0], [0, 1, 0])
        assert pm1.tobytes() == pm2.tobytes()
        r1, r2 = iac.calc_kmap(pm1), O.calc_kmap(pm2)
        for a, b in zip(r1, r2):
            assert a.tobytes() == b.tobytes()
        for e in (15, 63, 255):
            assert iac.mesh_shape_big(r1[3], e).as_dict() == O.mesh_shape_big(r2[3], e).as_dict()
            assert (iac.mesh_shape_small_cofq(r1[3], e).as_dict()
                    == O.mesh_shape_small_cofq(r2[3], e).as_dict())
        for z in (0.5, 2.0):
            assert iac.mesh_shape_big(r1[3], 31, z).as_dict() == O.mesh_shape_big(r2[3], 31, z).as_dict()


def test_product_rejects_even_mesh():
    import incohautocorrelate_amd as iac

    _, _, mq = O.detector(16)
    with pytest.raises(iac.IACError) as e:
        iac.mesh_shape_big(mq, 64)
    assert e.value.code == -2
