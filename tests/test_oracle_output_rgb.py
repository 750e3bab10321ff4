"""Pins the oracle's R'G'B' output stage (OutputWriter RGB48 and
convertToFloatRGB, output_writer.cpp:431-452, 729-777).

* RGBS must be the BT.601 inverse of the oracle's own E'Y/E'Cb/E'Cr planes
  (which the flat-colour round trip pins): R' = E'Y + 2(1-kr) E'Cr,
  B' = E'Y + 2(1-kb) E'Cb, G' = (E'Y - kr R' - kb B') / kg.  This is a
  different route to the same numbers than the oracle's composite-U/V
  matrix, so a wrong coefficient, sign or swapped U/V fails it.
* RGB48 is that full-range signal times 65535, luma clamped first, each
  component clamped to [0, 65535] and truncated.
* The clamp boxes: SDR [0, 1], HDR [0, inf), and the BT.601 legal Y'CbCr
  box carried through the inverse matrix, whose values the reference states
  (output_writer.cpp:166-171).
"""
import numpy as np
import pytest

KR_KB = {"modern": (0.299, 0.114), "classic": (0.30, 0.11)}


def _component_frame(seed, spread=1.0):
    rng = np.random.default_rng(seed)
    yuv = np.zeros((3, 625, 1135))
    rng_y = rng.uniform(-0.2 * spread, 1.0 + 0.2 * spread, (625, 1135))
    yuv[0] = 16384 + rng_y * 37632
    yuv[1:] = rng.uniform(-0.6 * spread, 0.6 * spread, (2, 625, 1135)) * 37632 * 0.5
    return yuv


@pytest.mark.parametrize("precision", ["modern", "classic"])
def test_rgbs_is_the_bt601_inverse_of_ycbcr(orc, precision):
    p = orc.pal_params(first_line=100, last_line=139, avs=300, ave=364, color_difference=precision)
    yuv = _component_frame(1)
    rgb = orc.output_stage(p, yuv, orc.OUT_RGBS).astype(np.float64)
    ps = orc.output_stage(p, yuv, orc.OUT_PS).astype(np.float64)
    kr, kb = KR_KB[precision]
    kg = 1.0 - kr - kb
    r = ps[0] + 2 * (1 - kr) * ps[2]
    b = ps[0] + 2 * (1 - kb) * ps[1]
    g = (ps[0] - kr * r - kb * b) / kg
    # float32 planes on both sides
    assert np.abs(rgb[0] - r).max() < 1e-6
    assert np.abs(rgb[1] - g).max() < 1e-6
    assert np.abs(rgb[2] - b).max() < 1e-6


def test_rgb48_is_truncated_full_range(orc):
    p = orc.pal_params(first_line=100, last_line=139, avs=300, ave=364)
    yuv = _component_frame(2)
    rgb48 = orc.output_stage(p, yuv, orc.OUT_RGB48).astype(np.int64)   # [h, w, 3]
    f = orc.output_stage(p, yuv, orc.OUT_RGBS).astype(np.float64)      # unclamped E'R/E'G/E'B
    ps = orc.output_stage(p, yuv, orc.OUT_PS).astype(np.float64)
    ey = ps[0]
    ry = np.clip(ey * 65535.0, 0, 65535)
    want = np.floor(np.clip(ry[None] + (f - ey[None]) * 65535.0, 0, 65535)).transpose(1, 2, 0)
    d = np.abs(rgb48 - want)
    # float32 planes used to build `want` leave only .0 boundary ties
    assert d.max() <= 1 and (d == 0).mean() > 0.999
    # a pixel exactly at black is 0 in every component
    yuv0 = np.zeros_like(yuv)
    yuv0[0] = 16384
    assert not orc.output_stage(p, yuv0, orc.OUT_RGB48).any()


def test_rgbs_clamp_boxes(orc):
    yuv = _component_frame(3, spread=4.0)
    sdr = orc.output_stage(orc.pal_params(first_line=100, last_line=139, avs=300, ave=364, clamp_mode=1), yuv,
                           orc.OUT_RGBS)
    assert sdr.min() == 0.0 and sdr.max() == 1.0
    hdr = orc.output_stage(orc.pal_params(first_line=100, last_line=139, avs=300, ave=364, clamp_mode=2), yuv,
                           orc.OUT_RGBS)
    assert hdr.min() == 0.0 and hdr.max() > 2.0
    bt = orc.output_stage(orc.pal_params(first_line=100, last_line=139, avs=300, ave=364, clamp_mode=3), yuv,
                          orc.OUT_RGBS)
    # the box's R'G'B' extremes over the 8 corners of the BT.601 legal
    # Y'CbCr box, by the BT.601 definition
    kr, kb = KR_KB["modern"]
    kg = 1.0 - kr - kb
    corners = []
    for ey in (-15 / 219, 238.75 / 219):
        for cb in (-127 / 224, 126.75 / 224):
            for cr in (-127 / 224, 126.75 / 224):
                r, b = ey + 2 * (1 - kr) * cr, ey + 2 * (1 - kb) * cb
                corners.append((r, (ey - kr * r - kb * b) / kg, b))
    corners = np.array(corners)
    for c in range(3):
        assert abs(bt[c].min() - corners[:, c].min()) < 1e-6 and abs(bt[c].max() - corners[:, c].max()) < 1e-6, c
    # R' and B' also match the values the reference states; its stated G'
    # box (-0.667284, 1.690154) is 3e-5 from what its own formula gives
    for c, (lo, hi) in ((0, (-0.863377, 1.883502)), (2, (-1.073153, 2.092866))):
        assert abs(bt[c].min() - lo) < 1e-6 and abs(bt[c].max() - hi) < 1e-6, c


def test_rgb_border_is_black(orc):
    p = orc.pal_params(first_line=100, last_line=139, avs=300, ave=364, pad_left=3, pad_top=2, out_width=80,
                       out_height=48)
    yuv = _component_frame(4)
    rgb48 = orc.output_stage(p, yuv, orc.OUT_RGB48)
    rgbs = orc.output_stage(p, yuv, orc.OUT_RGBS)
    assert rgb48.shape == (48, 80, 3) and rgbs.shape == (3, 48, 80)
    inner = np.zeros((48, 80), bool)
    inner[2:42, 3:67] = True
    assert not rgb48[~inner].any() and not rgbs[:, ~inner].any()
    assert rgbs[:, inner].any()
