"""Pins for the oracle's restatement of the LAP photon finder and the gain correction
(PPP::PhotonFinder_LargestAdjacentPixel, IAC/PPP.cpp:18-95; PPP::GainCorrection,
IAC/PPP.cpp:781-803; the frame chain of PPP::ProcessData_PF_LAP, IAC/PPP.cpp:469-495).

Every expected value below is worked out by hand from the reference's rules on inputs
that are exact in float32; each case is built so that one plausible slip (wrapping at
a row edge, the i - fs > 0 test on the upper neighbour, summing instead of taking the
largest neighbour, panels seeing each other, the gain branch) changes the answer."""
import numpy as np
import pytest

import oracle as O

P43 = [(0, 4, 3)]  # one panel, 4 pixels wide, 3 rows: pixel = 4*row + col


def lap(vals, panels=P43, npix=12, **kw):
    I = np.zeros((1, npix), np.float32)
    for j, v in vals.items():
        I[0, j] = v
    out, pc, mi = O.ppp_lap(I, panels, **kw)
    return out[0], pc[0], mi[0]


def expect(npix, ones=()):
    e = np.zeros(npix, np.float32)
    for j in ones:
        e[j] = 1
    return e


def test_right_neighbour_completes_a_photon():
    out, pc, mi = lap({0: 0.625, 1: 0.375})
    assert out.tolist() == expect(12, ones=[0]).tolist()
    assert pc == 1 and mi == np.float32(1) / np.float32(12)


def test_upper_neighbour_skipped_for_first_pixel_of_second_row():
    # i = fs: the reference tests i - fs > 0, so pixel 4 never looks at pixel 0
    out, _, _ = lap({4: 0.625, 0: 0.375})
    assert out.tolist() == [0.0] * 12
    # one pixel further the upper neighbour counts
    out, _, _ = lap({5: 0.625, 1: 0.375})
    assert out.tolist() == expect(12, ones=[5]).tolist()


def test_rows_do_not_wrap():
    out, _, _ = lap({8: 0.625, 7: 0.375})  # left of a row start is the previous row's end
    assert out.tolist() == [0.0] * 12
    out, _, _ = lap({3: 0.625, 4: 0.375})  # right of a row end is the next row's start
    assert out.tolist() == [0.0] * 12


def test_lower_neighbour_and_last_row():
    out, _, _ = lap({7: 0.625, 11: 0.375})
    assert out.tolist() == expect(12, ones=[7]).tolist()
    out, _, _ = lap({11: 0.625, 7: 0.375})  # the upper one of the last row
    assert out.tolist() == expect(12, ones=[11]).tolist()
    out, _, _ = lap({10: 0.625})  # nothing below the last row, nothing around
    assert out.tolist() == [0.0] * 12


def test_largest_neighbour_not_the_sum():
    base = {5: 0.5, 4: 0.25, 1: 0.25, 9: 0.3125}
    out, _, _ = lap({**base, 6: 0.375})  # 0.5 + 0.375 < 0.9 although the sum is 1.6875
    assert out.tolist() == [0.0] * 12
    out, _, _ = lap({**base, 6: 0.4375})  # 0.5 + 0.4375 >= 0.9
    assert out.tolist() == expect(12, ones=[5]).tolist()


def test_seed_threshold_is_inclusive_and_whole_photons_scale_with_adu():
    out, pc, _ = lap({0: 262.5, 1: 37.5, 2: 150.0}, adu=100.0)
    # 2.625 -> 2 + (0.625 + 0.375 >= 0.9) = 3; 0.375 < seed -> 0; 1.5: 0.5 + 0.375 < 0.9 -> 1
    assert out[:4].tolist() == [3.0, 0.0, 1.0, 0.0]
    assert pc == 4


def test_negative_and_outside_pixels():
    # pixels 12, 13 lie beyond the panels' 12 pixels and keep their values
    out, pc, _ = lap({0: -3.0, 5: 2.0, 12: -3.0, 13: 0.75}, npix=14)
    assert out.tolist() == [0.0] * 5 + [2.0] + [0.0] * 6 + [-3.0, 0.75]
    assert pc == int(np.float32(2.0) + np.float32(-3.0) + np.float32(0.75))  # (int)(-0.25) = 0


def test_panels_do_not_see_each_other():
    two = [(0, 4, 3), (12, 4, 3)]
    out, _, _ = lap({13: 0.625, 9: 0.375}, panels=two, npix=24)
    assert out.tolist() == [0.0] * 24  # pixel 13 is row 0 of the second panel
    out, _, _ = lap({13: 0.625, 9: 0.375}, panels=[(0, 4, 6)], npix=24)
    assert out.tolist() == expect(24, ones=[13]).tolist()
    # list order does not matter for disjoint panels
    a, _, _ = lap({13: 0.625, 17: 0.375, 2: 0.75, 3: 0.25}, panels=two, npix=24)
    b, _, _ = lap({13: 0.625, 17: 0.375, 2: 0.75, 3: 0.25}, panels=two[::-1], npix=24)
    assert a.tolist() == b.tolist() == expect(24, ones=[2, 13]).tolist()


def test_gain_correction_branches():
    I = np.array([[5.0, 0.5, 3.0, 7.0]], np.float32)
    off = np.ones(4, np.float32)
    gain = np.array([2.0, 2.0, 0.0, 2.0], np.float32)
    mask = np.array([1, 1, 1, 0], np.int32)
    out, _, _ = O.ppp_lap(I, [], mask=mask, offset=off, gain=gain)
    assert out[0].tolist() == [2.0, 0.0, 0.0, 0.0]  # (5-1)/2; negative -> 0; gain 0 -> 0; masked -> 0
    out, pc, _ = O.ppp_lap(I, [], mask=mask, offset=off, gain=gain, gain_only=True)
    assert out[0].tolist() == [2.0, -0.25, 0.0, 0.0]  # gain only allows negative values
    assert pc[0] == 1  # (int)1.75


def test_gain_then_lap():
    I = np.array([[0.375, 0.25]], np.float32)
    off = np.full(2, 0.0625, np.float32)
    gain = np.full(2, 0.5, np.float32)
    out, _, _ = O.ppp_lap(I, [(0, 2, 1)], offset=off, gain=gain)
    assert out[0].tolist() == [1.0, 0.0]  # corrected 0.625 and 0.375: 1.0 >= 0.9


def test_mask_without_gain():
    I = np.array([[0.625, 0.375, 0.75, 0.25]], np.float32)
    out, _, _ = O.ppp_lap(I, [(0, 4, 1)], mask=np.array([1, 0, 1, 1], np.int32))
    assert out[0].tolist() == [0.0, 0.0, 1.0, 0.0]  # masked right neighbour no longer helps pixel 0


@pytest.mark.parametrize("adu", [1.0, 37.0])
def test_random_frames_stay_within_one_photon(adu):
    rng = np.random.default_rng(3)
    fs, ss = 16, 9
    I = (adu * (rng.poisson(0.4, (5, fs * ss + 7)) + rng.uniform(-0.3, 1.0, (5, fs * ss + 7)))).astype(np.float32)
    out, pc, _ = O.ppp_lap(I, [(0, fs, ss)], adu=adu)
    inside = I[:, :fs * ss]
    whole = np.where(inside > 0, np.floor(inside / np.float32(adu)), 0)
    d = out[:, :fs * ss] - whole
    assert set(np.unique(d)) <= {0.0, 1.0}
    assert (d[inside <= 0] == 0).all()
    assert np.array_equal(out[:, fs * ss:], I[:, fs * ss:])
    # a seed pixel whose own fraction is at least the combined threshold always gains the photon
    frac = np.where(inside > 0, inside / np.float32(adu) - whole, 0)
    assert (d[frac >= 0.9] == 1).all()
    assert (d[frac < 0.5] == 0).all()


def test_photon_count_cast_follows_x86():
    """PhotonCount = (int)Sum: the x86-64 conversion (cvttss2si) truncates and gives
    0x80000000 for NaN and out-of-range sums; numpy's float32 -> int32 cast on this x86 host
    executes that instruction, so it pins the oracle's rule."""
    import warnings

    vals = np.array([3e9, -3e9, np.nan, np.inf, -np.inf, 2147483520.0, -2147483648.0, 2147483648.0, -2.5, 2.9,
                     0.0, -0.0, 1e-30], np.float32)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        host = vals.astype(np.int32)
    assert [O.cvtt_i32(v) for v in vals] == host.tolist()
    # a gain-only frame whose float sum passes 2^31
    I = np.full((1, 4), 1.0e9, np.float32)
    _, pc, mi = O.ppp_lap(I, [], gain_only=True)
    assert pc[0] == -2 ** 31 and mi[0] == np.float32(-2 ** 31) / np.float32(4)
