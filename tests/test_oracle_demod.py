"""Pins the oracle's demodulator and output stage end to end.

Reference: PalColour::buildLookUpTables / detectBurst / decodeLine
(palcolour.cpp:150-409, 456-646) and OutputWriter (output_writer.cpp:46-76,
402-486, 670-727).  The checks use what the mathematics fixes: a flat-colour
picture encoded by the synthetic PAL encoder must come back as the colour
that was encoded, gain/phase options act linearly, and the H.273
quantisation has closed forms.
"""
import numpy as np
import pytest


def test_filter_size_and_quadrature_reference(orc):
    p = orc.pal_params()
    s, c, fs, cf, yf = orc.pal_luts(p)
    assert fs == 7  # floor(0.5 * 4fSC / (1.1 MHz / 0.93)) = floor(7.4968)
    # fSC / sampleRate = 1/4 exactly: the reference carrier is in quadrature
    i = np.arange(1135)
    assert np.allclose(s, np.sin(np.pi * i / 2), atol=1e-12)
    assert np.allclose(c, np.cos(np.pi * i / 2), atol=1e-12)
    # the raised cosine reaches zero at the kernel half-width
    assert cf[7, 0] > 0 and np.all(cf >= 0)


@pytest.mark.parametrize("kind", [6, 7, 8])
@pytest.mark.parametrize("yuv", [(0.5, 0.2, -0.15), (0.3, -0.25, 0.3), (0.7, 0.0, 0.0)])
def test_flat_colour_round_trip(orc, synth, kind, yuv):
    """Encode a flat Y'UV picture, decode: the centre region returns the
    encoded colour (pins burst phase, V-switch, demod signs, the 2x gain,
    filter normalisation and the E' scaling).  Kind 6 (CHD_DEC_PAL_2D)
    filters the raw composite: its luma also pins the Y filter and the
    chroma re-synthesis (palcolour.cpp:525-630)."""
    y, u, v = yuv
    pic = np.zeros((synth.PIC_HEIGHT, synth.PIC_WIDTH))
    fields = []
    for f in range(3):
        for k in range(2):
            fields.append(synth.encode_field(2 * f + k, pic + y, pic + u, pic + v))
    fields = np.stack(fields)
    p = orc.pal_params(first_line=200, last_line=263, avs=500, ave=628)
    out = orc.decode_frame(p, kind, fields, 1, orc.uniform_thresholds(kind, 0.4), 0)
    ecb, ecr = orc.color_conversion()
    centre = out[:, 16:-16, 16:-16]
    # residual line-to-line ripple of the separation is ~0.5% of full
    # scale; a sign, gain, phase or U/V swap error is tens of percent
    assert np.abs(centre[0] - y).max() < 1e-2
    assert np.abs(centre[1] - u * ecb).max() < 1e-2
    assert np.abs(centre[2] - v * ecr).max() < 1e-2


def test_chroma_gain_scales_and_phase_rotates(orc, patch_capture):
    _, fields = patch_capture
    base = dict(first_line=300, last_line=331, avs=600, ave=664)
    p1 = orc.pal_params(**base)
    a = orc.decode_frame(p1, 7, fields, 2, orc.uniform_thresholds(7), 0)
    p2 = orc.pal_params(chroma_gain=2.0, **base)
    b = orc.decode_frame(p2, 7, fields, 2, orc.uniform_thresholds(7), 0)
    assert np.allclose(b[0], a[0])
    assert np.allclose(b[1:], 2 * a[1:], rtol=1e-6, atol=1e-7)
    # the phase option rotates the burst reference (palcolour.cpp:321-325):
    # 180 degrees negates both colour-difference signals, 360 is the identity
    p3 = orc.pal_params(chroma_phase_deg=180.0, **base)
    c = orc.decode_frame(p3, 7, fields, 2, orc.uniform_thresholds(7), 0)
    assert np.allclose(c[0], a[0])
    assert np.allclose(c[1:], -a[1:], rtol=1e-6, atol=1e-7)
    p4 = orc.pal_params(chroma_phase_deg=360.0, **base)
    d = orc.decode_frame(p4, 7, fields, 2, orc.uniform_thresholds(7), 0)
    assert np.allclose(d, a, rtol=1e-6, atol=1e-7)
    p5 = orc.pal_params(chroma_phase_deg=20.0, **base)
    e = orc.decode_frame(p5, 7, fields, 2, orc.uniform_thresholds(7), 0)
    assert not np.allclose(e[1:], a[1:], atol=1e-3)


def test_burst_detection_unit_vector(orc, patch_capture):
    _, fields = patch_capture
    p = orc.pal_params()
    vs = []
    for line in range(22, 310, 7):
        bp, bq, v = orc.detect_burst(p, fields[0], line)
        assert abs(np.hypot(bp, bq) - 1.0) < 1e-9
        vs.append(v)
    assert set(vs) == {-1.0, 1.0}


def _components(p, Y, U, V):
    h = 2 * p.field_height - 1
    full = [np.zeros((h, p.field_width)) for _ in range(3)]
    for arr, val in zip(full, (Y, U, V)):
        arr[:] = val
    return full


@pytest.mark.parametrize("clamp,ylo,yhi,clo,chi", [(0, 0, 65535, 0, 65535),
                                                   (1, 16 * 256, 235 * 256, 16 * 256, 240 * 256),
                                                   (3, 256, int(254.75 * 256), 256, int(254.75 * 256))])
def test_h273_quantisation(orc, clamp, ylo, yhi, clo, chi):
    import ctypes as C
    p = orc.pal_params(first_line=44, last_line=47, avs=185, ave=189, clamp_mode=clamp)
    black, white = p.black16b, p.white16b
    rng = white - black
    ecb, ecr = orc.color_conversion()
    L = orc.lib()
    cases = [(black, 0.0, 0.0, 16 * 256, 128 * 256, 128 * 256),
             (white, 0.0, 0.0, 235 * 256, 128 * 256, 128 * 256),
             (black + 0.5 * rng, 0.25 * rng / ecb, -0.25 * rng / ecr, 32128, 128 * 256 + 14336, 128 * 256 - 14336),
             (black + 1.3 * rng, 0.9 * rng / ecb, -0.9 * rng / ecr, None, None, None)]
    for Yv, Uv, Vv, ey, ecbv, ecrv in cases:
        Y, U, V = _components(p, Yv, Uv, Vv)
        out = np.zeros((3, p.out_height, p.out_width), np.uint16)
        L.orc_output_u16(C.byref(p), orc._dp(Y), orc._dp(U), orc._dp(V), out.ctypes.data_as(C.POINTER(C.c_uint16)))
        def q(x, scale, zero, lo, hi):
            v = np.sign(x) * np.floor(abs(scale * x + zero) + 0.5) if scale * x + zero >= 0 else -np.floor(abs(scale * x + zero) + 0.5)
            return int(min(max(v, lo), hi))
        if ey is None:
            ey = q((Yv - black) / rng, 219 * 256, 16 * 256, ylo, yhi)
            ecbv = q(Uv * ecb / rng, 224 * 256, 128 * 256, clo, chi)
            ecrv = q(Vv * ecr / rng, 224 * 256, 128 * 256, clo, chi)
        assert (out[0] == min(max(ey, ylo), yhi)).all()
        assert (out[1] == min(max(ecbv, clo), chi)).all()
        assert (out[2] == min(max(ecrv, clo), chi)).all()


def test_float_output_and_border(orc):
    import ctypes as C
    p = orc.pal_params(first_line=44, last_line=47, avs=185, ave=189, pad_left=1, pad_top=2, out_width=8, out_height=8)
    rng = p.white16b - p.black16b
    Y, U, V = _components(p, p.black16b + 0.25 * rng, 0.1 * rng, -0.2 * rng)
    out = np.full((3, 8, 8), 7.0, np.float32)
    orc.lib().orc_output_float(C.byref(p), orc._dp(Y), orc._dp(U), orc._dp(V), out.ctypes.data_as(C.POINTER(C.c_float)))
    ecb, ecr = orc.color_conversion()
    act = out[:, 2:6, 1:5]
    assert np.allclose(act[0], 0.25) and np.allclose(act[1], np.float32(0.1 * ecb)) and np.allclose(act[2], np.float32(-0.2 * ecr))
    border = out.copy()
    border[:, 2:6, 1:5] = 0
    assert (border == 0).all()
