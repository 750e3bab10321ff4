"""Pins the oracle's carrier-mirror filter and tile lattice.

Reference: applyFilter in transform_pal_3d.cpp:249-333 and
transform_pal_2d.cpp (y_ref = (YTILE/2 + YTILE - y) % YTILE), the lattices of
filterFields (transform_pal_3d.cpp:123-165, transform_pal_2d.cpp:107-160) and
the overlap-add normalisation (:205-240).  Checks are behavioural: closed
forms at t = 0 and t > 1, single mirror pairs around the threshold, the
self-mirror (carrier) bins, and whole-field identities.
"""
import numpy as np
import pytest

SELF_3D = {(1, 4, 4), (1, 20, 4), (5, 4, 4), (5, 20, 4)}
SELF_2D = {(0, 4, 8), (0, 12, 8)}


def _spec(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


@pytest.mark.parametrize("kind,shape,xs", [(8, (8, 32, 9), range(2, 7)), (7, (1, 16, 17), range(4, 13))])
def test_zero_threshold_keeps_the_chroma_band(orc, kind, shape, xs):
    h = _spec(shape, 10)
    out = orc.apply_filter(h, np.zeros(768 if kind == 8 else 80), kind)
    want = np.zeros_like(h)
    want[:, :, list(xs)] = h[:, :, list(xs)]
    assert np.array_equal(out, want)


@pytest.mark.parametrize("kind,shape,selfset", [(8, (8, 32, 9), SELF_3D), (7, (1, 16, 17), SELF_2D)])
def test_threshold_above_one_keeps_only_carrier_bins(orc, kind, shape, selfset):
    h = _spec(shape, 11)
    out = orc.apply_filter(h, np.full(768 if kind == 8 else 80, 1.5 ** 2), kind)
    kept = {tuple(int(v) for v in i) for i in np.argwhere(out != 0)}
    assert kept == selfset
    for b in selfset:
        assert out[b] == h[b]


def _mirror3(z, y, x):
    return ((10 - z) % 8, (40 - y) % 32, 8 - x)


@pytest.mark.parametrize("bin_", [(0, 0, 2), (3, 17, 3), (7, 31, 2), (2, 9, 4), (6, 30, 4)])
def test_single_pair_decision_around_threshold(orc, bin_):
    t = 0.4
    m = _mirror3(*bin_)
    for r, keep in ((t * 1.001, True), (t * 0.999, False), (1 / (t * 1.001), True), (1 / (t * 0.999), False)):
        h = np.zeros((8, 32, 9), complex)
        h[bin_] = 1000.0 * np.exp(0.3j)
        h[m] = 1000.0 * r * np.exp(-1.1j)
        out = orc.apply_filter(h, np.full(768, t * t), 8)
        assert (out[bin_] != 0) == keep and (out[m] != 0) == keep, (bin_, r)
        if keep:
            assert out[bin_] == h[bin_] and out[m] == h[m]


def test_per_bin_threshold_order_and_or_semantics(orc):
    """Thresholds are indexed (z*32 + y)*3 + (x-2); a bin at x = 4 is visited
    from both sides of its pair and kept if either visit keeps it."""
    b = (3, 11, 2)
    m = _mirror3(*b)
    h = np.zeros((8, 32, 9), complex)
    h[b] = 100.0
    h[m] = 50.0  # ratio 0.5
    thr = np.zeros(768)
    thr[(3 * 32 + 11) * 3 + 0] = 0.6 ** 2  # only this visit's threshold discards
    out = orc.apply_filter(h, thr, 8)
    assert out[b] == 0 and out[m] == 0
    thr2 = np.zeros(768)
    thr2[(m[0] * 32 + m[1]) * 3 + 0] = 0.6 ** 2  # the partner (x = 6) is never a visit
    out2 = orc.apply_filter(h, thr2, 8)
    assert out2[b] == h[b] and out2[m] == h[m]
    # x = 4 pair: both bins are visits; discard needs both thresholds high
    b4 = (2, 9, 4)
    m4 = _mirror3(*b4)
    h4 = np.zeros((8, 32, 9), complex)
    h4[b4] = 100.0
    h4[m4] = 50.0
    i_b = (b4[0] * 32 + b4[1]) * 3 + 2
    i_m = (m4[0] * 32 + m4[1]) * 3 + 2
    thr3 = np.zeros(768)
    thr3[i_b] = 0.6 ** 2
    assert orc.apply_filter(h4, thr3, 8)[b4] != 0
    thr3[i_m] = 0.6 ** 2
    assert orc.apply_filter(h4, thr3, 8)[b4] == 0


def _window_fields(p, kind, make_line, n_fields):
    fw, fh = p.field_width, p.field_height
    x = np.arange(fw)
    field = np.zeros((fh, fw))
    field[:] = make_line(x)[None, :]
    return [field.copy() for _ in range(n_fields)]


@pytest.mark.parametrize("kind", [8, 7])
def test_subcarrier_tone_passes_unchanged_at_zero_threshold(orc, kind):
    """At t = 0 the transform is a horizontal band-pass over x bins 2..6 (3D)
    or 4..12 (2D).  A tone at fSC (4 samples/cycle) lies inside it, so after
    window, FFT, inverse and overlap-add every active sample must reproduce
    the tone exactly -- which pins the lattice, the window partition, the
    1/N normalisation and the active-area bounds."""
    p = orc.pal_params(first_line=100, last_line=163, avs=300, ave=428)
    black = p.black16b
    # phase 0: samples are exact integers (6000, 0, -6000, 0) -> a pure tone
    tone = lambda x: black + 8000.0 + 6000.0 * np.round(np.cos(2 * np.pi * x / 4))
    nf = 14 if kind == 8 else 2
    fields = [np.round(f).astype(np.uint16) for f in _window_fields(p, kind, tone, nf)]
    offs = [i % 2 for i in range(nf)]
    s, e = (4, 6) if kind == 8 else (0, 2)
    out = orc.filter_fields(p, kind, fields, offs, s, e, np.zeros(768 if kind == 8 else 80))
    x = np.arange(p.field_width)
    want_line = 6000.0 * np.round(np.cos(2 * np.pi * x / 4))
    for j in range(2):
        off = offs[s + j]
        first = (p.first_active_frame_line + 1 - off) // 2
        last = (p.last_active_frame_line - off) // 2 + 1
        act = out[j][first:last, p.active_video_start:p.active_video_end]
        assert np.abs(act - want_line[None, p.active_video_start:p.active_video_end]).max() < 1e-7
        # nothing outside the active area
        mask = np.ones(out[j].shape, bool)
        mask[first:last, p.active_video_start:p.active_video_end] = False
        assert np.abs(out[j][mask]).max() == 0.0


@pytest.mark.parametrize("kind", [8, 7])
def test_constant_input_gives_no_chroma_and_offset_invariance(orc, kind):
    """The window's spectrum occupies bins {0, +-1}, so a constant picture
    has no chroma, and adding a constant to every sample (and to black) does
    not change the chroma of any picture."""
    p = orc.pal_params(first_line=120, last_line=183, avs=500, ave=628)
    nf = 14 if kind == 8 else 2
    s, e = (4, 6) if kind == 8 else (0, 2)
    offs = [i % 2 for i in range(nf)]
    flat = [np.full((313, 1135), 30000, np.uint16) for _ in range(nf)]
    out = orc.filter_fields(p, kind, flat, offs, s, e, np.full(768 if kind == 8 else 80, 0.16))
    assert np.abs(out).max() < 1e-9
    rng = np.random.default_rng(5)
    pic = [rng.integers(17000, 40000, size=(313, 1135)).astype(np.uint16) for _ in range(nf)]
    a = orc.filter_fields(p, kind, pic, offs, s, e, np.full(768 if kind == 8 else 80, 0.16))
    p2 = orc.pal_params(first_line=120, last_line=183, avs=500, ave=628, black=p.black16b + 1234)
    pic2 = [(f + 1234).astype(np.uint16) for f in pic]
    b = orc.filter_fields(p2, kind, pic2, offs, s, e, np.full(768 if kind == 8 else 80, 0.16))
    assert np.abs(a - b).max() < 1e-6 * np.abs(a).max()


def test_3d_uses_only_the_two_lattice_slabs(orc):
    """The 3D lattice of a one-frame window starts at start-4 (transform_pal_3d.cpp:151):
    tiles at window fields 0 and 4 only, so fields 12 and 13 of a 14-field
    window never influence the output, while field 11 does."""
    p = orc.pal_params(first_line=120, last_line=183, avs=500, ave=628)
    rng = np.random.default_rng(6)
    pic = [rng.integers(17000, 40000, size=(313, 1135)).astype(np.uint16) for _ in range(14)]
    offs = [i % 2 for i in range(14)]
    thr = np.full(768, 0.16)
    a = orc.filter_fields(p, 8, pic, offs, 4, 6, thr)
    pic_b = list(pic)
    pic_b[12] = rng.integers(17000, 40000, size=(313, 1135)).astype(np.uint16)
    pic_b[13] = rng.integers(17000, 40000, size=(313, 1135)).astype(np.uint16)
    assert np.array_equal(a, orc.filter_fields(p, 8, pic_b, offs, 4, 6, thr))
    pic_c = list(pic)
    pic_c[11] = rng.integers(17000, 40000, size=(313, 1135)).astype(np.uint16)
    assert not np.array_equal(a, orc.filter_fields(p, 8, pic_c, offs, 4, 6, thr))
