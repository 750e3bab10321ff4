"""Pins the oracle's luma noise coring (PalColour::doYNR, palcolour.cpp:411-448,
with the f_nrpal FIR of deemp.h:620-633).

* The taps, read back from the oracle as its impulse response, are a
  symmetric 25-tap high-pass: stop band below 1 MHz, pass band from 4 MHz at
  4fSC (the filter's design; a mistyped tap breaks one of these).
* Unclipped, the coring subtracts numpy's centred convolution of the
  zero-padded active samples with those taps: this fixes the 12-sample delay,
  the orientation and the zeros outside [active_video_start, active_video_end).
* The clip level is level * ((white - black) // 100): 376 for the default
  IRE levels (37632 / 100 = 376.32 in real arithmetic).
* In a decode, coring changes only E'Y, line by line, after the demodulation.
"""
import numpy as np
import pytest

AVS, AVE = 185, 1107


def _params(orc, level, **kw):
    return orc.pal_params(first_line=100, last_line=139, avs=AVS, ave=AVE, luma_nr_level=level, **kw)


def _taps(orc):
    """The FIR taps as the oracle applies them: unit impulse, no clipping."""
    p = _params(orc, 1e9)
    y = np.zeros(1135)
    h0 = 600
    y[h0] = 1.0
    out = orc.ynr_line(p, y)
    # out = y - hp, hp[h0 + j] = b[12 + j] for an impulse at h0
    hp = y - out
    return hp[h0 - 12:h0 + 13], hp


def test_taps_are_the_pal_noise_highpass(orc):
    b, hp = _taps(orc)
    assert np.count_nonzero(hp) == 25       # 25 taps, nothing else
    assert np.allclose(b, b[::-1], rtol=0, atol=1e-15)   # to the table's last printed digit
    fs = 17734475.0
    f = np.fft.rfftfreq(4096, 1 / fs)
    H = np.abs(np.fft.rfft(b, 4096))
    assert H[f <= 1.0e6].max() < 5e-3
    assert np.abs(H[f >= 4.0e6] - 1.0).max() < 5e-3
    assert abs(b.sum()) < 2e-3               # DC gain


def test_unclipped_coring_is_a_centred_convolution(orc):
    b, _ = _taps(orc)
    rng = np.random.default_rng(5)
    y = np.zeros(1135)
    y[AVS:AVE] = 16384 + rng.uniform(0, 37632, AVE - AVS)
    y[:AVS] = 9999.0                          # outside the active line: never read
    y[AVE:] = -9999.0
    got = orc.ynr_line(_params(orc, 1e9), y)
    pad = np.zeros(1135)
    pad[AVS:AVE] = y[AVS:AVE]
    want = y.copy()
    want[AVS:AVE] -= np.convolve(pad, b, mode="same")[AVS:AVE]
    assert np.abs(got - want).max() < 1e-9
    assert np.array_equal(got[:AVS], y[:AVS]) and np.array_equal(got[AVE:], y[AVE:])


@pytest.mark.parametrize("level, white, clip", [(1.0, 0xD300, 376.0), (2.5, 0xD300, 940.0),
                                                (1.0, 16384 + 37699, 376.0), (0.5, 16384 + 3799, 18.5)])
def test_clip_level_uses_integer_ire_scale(orc, level, white, clip):
    """A square wave whose high-pass is far above the clip: every sample moves
    by exactly the clip level, with the sign of its high-pass."""
    b, _ = _taps(orc)
    y = np.zeros(1135)
    y[AVS:AVE] = np.where((np.arange(AVE - AVS) // 2) % 2 == 0, 40000.0, 20000.0)
    p = _params(orc, level, white=white)
    got = orc.ynr_line(p, y)
    d = y - got
    pad = np.zeros(1135)
    pad[AVS:AVE] = y[AVS:AVE]
    hp = np.convolve(pad, b, mode="same")[AVS:AVE]
    big = np.abs(hp) > clip + 1e-6
    assert big.mean() > 0.9
    assert np.array_equal(d[AVS:AVE][big], np.sign(hp[big]) * clip)
    small = ~big
    assert not small.any() or np.abs(d[AVS:AVE][small] - hp[small]).max() < 1e-9


def test_decode_cores_luma_only(orc, synth):
    """Coring runs on each decoded E'Y line after demodulation
    (palcolour.cpp:643-645): U and V are unchanged, and each Y line equals the
    uncored line put through the coring step."""
    scene = synth.PatchScene(seed=4)
    fields = synth.encode_sequence([scene.frame(0)], synth.EncodeOptions(snr_db=30, rng=np.random.default_rng(4)))
    for kind in (6, 7):
        p0 = orc.pal_params()
        p1 = orc.pal_params(luma_nr_level=2.0)
        win, offs, s, e = orc.frame_window(fields, 0, kind, p0.black16b)
        thr = orc.uniform_thresholds(kind)
        _, yuv0 = orc.decode_window(p0, kind, win, offs, s, e, thr, 0, want_yuv=True)
        _, yuv1 = orc.decode_window(p1, kind, win, offs, s, e, thr, 0, want_yuv=True)
        assert np.array_equal(yuv0[1:], yuv1[1:])
        moved = 0
        for line in range(p0.first_active_frame_line, p0.last_active_frame_line + 1):
            want = orc.ynr_line(p1, yuv0[0, line])
            assert np.array_equal(yuv1[0, line], want), (kind, line)
            moved += int(np.any(want != yuv0[0, line]))
        assert moved == p0.last_active_frame_line - p0.first_active_frame_line + 1
