"""CPU pins of the averaged-intensity restatement (Detector::LoadAndAverageIntensity)."""
import numpy as np

import oracle as O


def test_discretised_branch_counts_photons():
    """Integer photon frames with threshold 0.5 and step 1: ceil(v - 0.5) = v, so the
    average is the plain mean of the masked counts and PhotonCount the frame sum."""
    rng = np.random.default_rng(0)
    ph = rng.poisson(0.3, (9, 500)).astype(np.float32)
    mask = (rng.random(500) > 0.2).astype(np.int32)
    avg, mean_i, pc = O.average_intensity(ph, 0.5, 1.0, mask)
    masked = ph * mask
    assert np.array_equal(pc, masked.sum(axis=1).astype(np.int32))
    assert np.array_equal(avg, masked.sum(axis=0).astype(np.float32) * np.float32(1.0 / 9.0))
    assert np.allclose(mean_i, masked.sum(axis=1) / 500.0, rtol=1e-6)


def test_adu_frames_round_to_photons():
    """ADU values k * 6.4 + small noise with offset 3.2, step 6.4 give k photons."""
    rng = np.random.default_rng(1)
    k = rng.poisson(0.5, (6, 300))
    fr = (k * 6.4 + rng.uniform(-1, 1, k.shape) * (k > 0)).astype(np.float32)
    avg, _, pc = O.average_intensity(fr, 3.2, 6.4)
    assert np.array_equal(pc, k.sum(axis=1))
    assert np.array_equal(avg, k.sum(axis=0).astype(np.float32) * np.float32(1.0 / 6.0))


def test_threshold_branches():
    """No discretisation: values >= threshold enter the average; with a positive
    threshold the event mean sees the zeroed frame, with threshold <= 0 all values."""
    fr = np.array([[0.25, 2.0, -1.0], [4.0, 0.5, 3.0]], np.float32)
    avg, mi, pc = O.average_intensity(fr, 1.0, 0.0)
    assert np.array_equal(avg, np.array([4.0, 2.0, 3.0], np.float32) * np.float32(0.5))
    assert np.array_equal(mi, np.array([2.0 / 3, 7.0 / 3], np.float32))
    assert np.all(pc == -1)
    avg, mi, _ = O.average_intensity(fr, -0.5, 0.0)
    assert np.array_equal(avg, np.array([4.25, 2.5, 3.0], np.float32) * np.float32(0.5))
    assert np.array_equal(mi, np.array([1.25 / 3, 7.5 / 3], np.float32))
