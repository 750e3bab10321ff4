"""GPU parity: chd_decode_frame (HIP kernels) against the FP64 oracle.

Tolerances: float planes within 1e-4 of full scale
(E'Y 0..1, E'Cb/E'Cr +-0.5), yuv444p16 within 1 LSB.  Every call goes
through the C ABI in libchromadec_amd/lib/libchromadec.so.
"""
import os

import numpy as np
import pytest

from libchromadec_amd import abi

pytestmark = pytest.mark.gpu

F32_TOL = 1e-4
U16_TOL = 1


def _oracle(orc, kind, fields, frame, fmt, **geo):
    thr = geo.pop("thr", None)
    p = orc.pal_params(**geo)
    t = orc.uniform_thresholds(kind) if thr is None else thr
    return orc.decode_frame(p, kind, fields, frame, t, 0 if fmt == "yuv444ps" else 1)


def _pads(aw, ah, mult=16):
    """Left/top padding and padded size (OutputWriter::updateConfiguration,
    output_writer.cpp:214-238)."""
    pl = pt = 0
    w, h, pr, pb = aw, ah, 0, 0
    while w % mult:
        if w % 2 == 0: pr += 1
        else: pl += 1
        w = pl + aw + pr
    while h % mult:
        if h % 2 == 0: pb += 1
        else: pt += 1
        h = pt + ah + pb
    return pl, pt, w, h


def _check(got, want, fmt):
    d = np.abs(got.astype(np.float64) - want.astype(np.float64))
    tol = F32_TOL if fmt in ("yuv444ps", "grays") else U16_TOL
    assert d.max() <= tol, f"max {d.max()} at {np.unravel_index(d.argmax(), d.shape)}"
    return d.max()


@pytest.mark.parametrize("fmt", ["yuv444ps", "yuv444p16"])
def test_transform3d_full_frames(orc, patch_capture, fmt):
    path, fields = patch_capture
    v = abi.Video(path)
    dec = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_3D, output_format=fmt).commit()
    for f in (0, 2, 5):  # black look-behind, interior, black look-ahead
        got = dec.decode(f)
        assert got.shape == (3, 576, 922)
        _check(got, _oracle(orc, 8, fields, f, fmt), fmt)


@pytest.mark.parametrize("kind", [7, 8])
def test_black_capture_decodes_to_black(orc, tmp_path, kind):
    """Every tile of an all-black capture has all-zero windowed samples: its
    FP32 error bound is zero and each carrier-mirror pair is kept without an
    FP64 re-decision (0 < 0 is false).  The planes are exactly black, as the
    oracle's are."""
    from libchromadec_amd import synth
    fields = np.full((12, synth.FIELD_HEIGHT, synth.FIELD_WIDTH), synth.BLACK, dtype=np.uint16)
    path = str(tmp_path / "black.tbc")
    synth.write_tbc(path, fields)
    v = abi.Video(path)
    dec = abi.Decoder(v, kind, output_format="yuv444ps").commit()
    for f in (0, 3, 5):
        got = dec.decode(f)
        assert not got.any(), f
        _check(got, _oracle(orc, kind, fields, f, "yuv444ps"), "yuv444ps")


@pytest.mark.parametrize("kind", [6, 8])
def test_subcarrier_locked_second_field_shift(orc, patch_capture, tmp_path, kind):
    """Subcarrier-locked PAL: the second field of every frame is moved two
    samples earlier, with two black samples appended, before decoding
    (SourceField::shiftSecondField, source_field.cpp:51-66)."""
    import json
    from libchromadec_amd import synth
    _, fields = patch_capture
    path = str(tmp_path / "sclocked.tbc")
    side = synth.write_tbc(path, fields)
    with open(side) as fh:
        meta = json.load(fh)
    meta["videoParameters"]["isSubcarrierLocked"] = True
    with open(side, "w") as fh:
        json.dump(meta, fh)
    shifted = fields.copy()
    for f in range(1, fields.shape[0], 2):   # isFirstField is false for odd fields
        flat = fields[f].reshape(-1)
        shifted[f] = np.concatenate([flat[2:], np.full(2, synth.BLACK, flat.dtype)]).reshape(fields.shape[1:])
    v = abi.Video(path)
    dec = abi.Decoder(v, kind, output_format="yuv444ps").commit()
    for f in (0, 2):
        got = dec.decode(f)
        _check(got, _oracle(orc, kind, shifted, f, "yuv444ps"), "yuv444ps")
        # the shift is visible: the unshifted fields decode differently
        assert np.abs(got - _oracle(orc, kind, fields, f, "yuv444ps")).max() > 1e-2


@pytest.mark.parametrize("fmt", ["yuv444ps", "yuv444p16"])
@pytest.mark.parametrize("kind", [6, 7, 8])
def test_second_field_marked_first(orc, patch_capture, tmp_path, kind, fmt):
    """A second field whose metadata says isFirstField sits on the first
    field's lines (SourceField::getOffset, source_field.h:77-79), so the
    windows holding it no longer alternate line parity field by field and
    the 3D tiles take the full vertical and temporal transform.  Its frame
    carries the second field's lines only; the other lines are the zero
    Y/U/V the component frame starts with (component_frame.cpp:47-52)."""
    import json
    from libchromadec_amd import synth
    _, fields = patch_capture
    path = str(tmp_path / "order.tbc")
    side = synth.write_tbc(path, fields)
    with open(side) as fh:
        meta = json.load(fh)
    meta["fields"][3]["isFirstField"] = True   # second field of frame 1
    with open(side, "w") as fh:
        json.dump(meta, fh)
    v = abi.Video(path)
    dec = abi.Decoder(v, kind, output_format=fmt).commit()
    p = orc.pal_params()
    for f in (0, 1, 3):
        win, offs, s, e = orc.frame_window(fields, f, kind, p.black16b)
        offs = [0 if 2 * (f - (2 if kind == 8 else 0)) + w == 3 else o for w, o in enumerate(offs)]
        assert (offs[s:e] == [0, 0]) == (f == 1)
        want = orc.decode_window(p, kind, win, offs, s, e, orc.uniform_thresholds(kind), 0 if fmt == "yuv444ps" else 1)
        got = dec.decode(f)
        _check(got, want, fmt)
        if f == 1:
            # the offset matters: the regular order decodes differently
            reg = _oracle(orc, kind, fields, f, fmt).astype(np.float64)
            assert np.abs(got.astype(np.float64) - reg).max() > (1e-2 if fmt == "yuv444ps" else 100)


@pytest.mark.parametrize("kind", [6, 7, 8])
def test_second_field_first_pairing(orc, patch_capture, kind):
    """Second-field-first pairing (LdDecodeMetaData::getFieldNumber,
    core.cpp:1129-1150): frame k is the first field 2k+3 with the second
    field 2k+2 (1-based), one frame fewer than first-field-first
    (core.cpp:1092-1106)."""
    path, fields = patch_capture
    v = abi.Video(path, second_field_first=True)
    n = fields.shape[0] // 2 - 1
    assert v.info().num_frames == n
    paired = np.stack([fields[i] for k in range(n) for i in (2 * k + 2, 2 * k + 1)])
    dec = abi.Decoder(v, kind, output_format="yuv444ps").commit()
    for f in (0, 2, n - 1):
        got = dec.decode(f)
        _check(got, _oracle(orc, kind, paired, f, "yuv444ps"), "yuv444ps")
    with pytest.raises(abi.ChdError):
        dec.decode(n)


@pytest.mark.parametrize("kind", [6, 8])
def test_reverse_field_order(patch_capture, kind):
    """reverse_field_order flips the pairing order the metadata gives
    (isFirstFieldFirst != reverseFieldOrder, c_api_decoder.cpp:859-874): on a
    first-field-first capture it decodes the second-field-first frames, and
    on a video opened second-field-first it restores the plain ones."""
    path, fields = patch_capture
    n = fields.shape[0] // 2 - 1
    rev = abi.Decoder(abi.Video(path), kind, output_format="yuv444ps", reverse_field_order=True).commit()
    sff = abi.Decoder(abi.Video(path, second_field_first=True), kind, output_format="yuv444ps").commit()
    for f in (0, n - 1):
        assert np.array_equal(rev.decode(f), sff.decode(f)), f
    with pytest.raises(abi.ChdError):
        rev.decode(n)
    both = abi.Decoder(abi.Video(path, second_field_first=True), kind, output_format="yuv444ps",
                       reverse_field_order=True).commit()
    plain = abi.Decoder(abi.Video(path), kind, output_format="yuv444ps").commit()
    for f in (0, n):
        assert np.array_equal(both.decode(f), plain.decode(f)), f


@pytest.mark.parametrize("fmt", ["yuv444ps", "yuv444p16"])
@pytest.mark.parametrize("kind", [6, 7, 8])
def test_luma_nr(orc, patch_capture, kind, fmt):
    """luma_nr_level cores E'Y after demodulation (PalColour::doYNR,
    palcolour.cpp:411-448, 643-645); U and V are those of the uncored
    decode, bit for bit."""
    path, fields = patch_capture
    v = abi.Video(path)
    dec = abi.Decoder(v, kind, output_format=fmt, luma_nr_level=2.0).commit()
    for f in (0, 3):
        got = dec.decode(f)
        _check(got, _oracle(orc, kind, fields, f, fmt, luma_nr_level=2.0), fmt)
    plain = abi.Decoder(v, kind, output_format=fmt).commit().decode(3)
    assert np.array_equal(got[1:], plain[1:])
    d = np.abs(got[0].astype(np.float64) - plain[0].astype(np.float64))
    assert d.max() > (1e-3 if fmt == "yuv444ps" else 50)


@pytest.mark.parametrize("fmt", ["rgb48", "rgbs"])
def test_luma_nr_wide_line_rgb(orc, patch_capture, fmt):
    """Coring on a 1112-sample active line (the column-segmented 4fSC
    demodulator) with padding and R'G'B' output."""
    path, fields = patch_capture
    crop = dict(first_active_frame_line=40, last_active_frame_line=600, first_active_sample=9,
                last_active_sample=1120)
    v = abi.Video(path)
    dec = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_3D, output_format=fmt, padding_multiple=16, luma_nr_level=3.5,
                      **crop).commit()
    oi = dec.output_info()
    got = dec.decode(2)
    fl, ll = crop["first_active_frame_line"], crop["last_active_frame_line"]
    fs, ls = crop["first_active_sample"], crop["last_active_sample"]
    aw, ah = ls - fs + 1, ll - fl + 1
    pl, pt, w, h = _pads(aw, ah)
    assert (w, h) == (oi.width, oi.height)
    out_fmt = orc.OUT_RGB48 if fmt == "rgb48" else orc.OUT_RGBS
    want = orc.decode_frame(orc.pal_params(first_line=fl, last_line=ll, avs=fs, ave=ls + 1, pad_left=pl, pad_top=pt,
                                           out_width=oi.width, out_height=oi.height, luma_nr_level=3.5),
                            8, fields, 2, orc.uniform_thresholds(8), out_fmt)
    d = np.abs(got.astype(np.float64) - want.astype(np.float64))
    assert d.max() <= (1 if fmt == "rgb48" else F32_TOL), d.max()


@pytest.mark.parametrize("fmt", ["yuv444ps", "yuv444p16"])
@pytest.mark.parametrize("capture", ["bars_capture", "patch_capture"])
def test_pal2d_full_frames(orc, request, capture, fmt):
    """CHD_DEC_PAL_2D: the PALcolour 2D filter on the raw composite
    (decodeLine<uint16_t, false>), through the same demodulator kernels."""
    path, fields = request.getfixturevalue(capture)
    v = abi.Video(path)
    dec = abi.Decoder(v, abi.CHD_DEC_PAL_2D, output_format=fmt).commit()
    for f in (0, 2):
        _check(dec.decode(f), _oracle(orc, 6, fields, f, fmt), fmt)
    # CHD_DEC_AUTO resolves to CHD_DEC_PAL_2D for PAL (registry.cpp:96-107)
    auto = abi.Decoder(v, abi.CHD_DEC_AUTO, output_format=fmt).commit()
    assert np.array_equal(auto.decode(1), dec.decode(1))


@pytest.mark.parametrize("fmt", ["yuv444ps", "yuv444p16"])
def test_transform2d_full_frames(orc, bars_capture, fmt):
    path, fields = bars_capture
    v = abi.Video(path)
    dec = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_2D, output_format=fmt).commit()
    for f in range(3):
        _check(dec.decode(f), _oracle(orc, 7, fields, f, fmt), fmt)


@pytest.mark.parametrize("kind", [6, 7, 8])
@pytest.mark.parametrize("crop", [dict(first_active_frame_line=101, last_active_frame_line=170,
                                       first_active_sample=333, last_active_sample=470),
                                  dict(first_active_frame_line=0, last_active_frame_line=624,
                                       first_active_sample=0, last_active_sample=1134),
                                  # 4fSC demod of a line wider than one column-blocked workgroup
                                  dict(first_active_frame_line=40, last_active_frame_line=600,
                                       first_active_sample=9, last_active_sample=1120)])
def test_crop_padding_clamp(orc, patch_capture, kind, crop):
    path, fields = patch_capture
    v = abi.Video(path)
    dec = abi.Decoder(v, kind, output_format="yuv444ps", padding_multiple=16, output_clamp="legal_rgb_sdr",
                      **crop).commit()
    oi = dec.output_info()
    got = dec.decode(3)
    fl, ll = crop["first_active_frame_line"], crop["last_active_frame_line"]
    fs, ls = crop["first_active_sample"], crop["last_active_sample"]
    aw, ah = ls - fs + 1, ll - fl + 1
    assert oi.width % 16 == 0 and oi.height % 16 == 0 and oi.width >= aw and oi.height >= ah
    pl, pt, w, h = _pads(aw, ah)
    want = _oracle(orc, kind, fields, 3, "yuv444ps", first_line=fl, last_line=ll, avs=fs, ave=ls + 1,
                   clamp_mode=1, pad_left=pl, pad_top=pt, out_width=w, out_height=h)
    _check(got, want, "yuv444ps")


CLAMPS = {"none": 0, "legal_rgb_sdr": 1, "legal_rgb_hdr": 2, "legal_ycbcr_bt601": 3}


@pytest.mark.parametrize("kind", [6, 8])
@pytest.mark.parametrize("fmt", ["rgbs", "rgb48"])
@pytest.mark.parametrize("clamp", ["none", "legal_rgb_sdr", "legal_rgb_hdr", "legal_ycbcr_bt601"])
def test_rgb_outputs(orc, patch_capture, kind, fmt, clamp):
    """R'G'B' output formats (OutputWriter RGB48 / convertToFloatRGB): RGBS
    within 1e-4, RGB48 (truncated) within 1 LSB, black border, and the
    interleaved RGB48 planes borrowed at sample offsets 0/1/2."""
    import ctypes as C
    path, fields = patch_capture
    crop = dict(first_active_frame_line=101, last_active_frame_line=170, first_active_sample=333,
                last_active_sample=470)
    v = abi.Video(path)
    dec = abi.Decoder(v, kind, output_format=fmt, output_clamp=clamp, padding_multiple=16, **crop).commit()
    oi = dec.output_info()
    got = dec.decode(3)
    # active 138 x 70 padded to 144 x 80: odd remainders go left/top first
    # (OutputWriter::updateConfiguration, output_writer.cpp:214-238)
    assert (oi.width, oi.height) == (144, 80)
    p = orc.pal_params(first_line=101, last_line=170, avs=333, ave=471, clamp_mode=CLAMPS[clamp], pad_left=3,
                       pad_top=5, out_width=144, out_height=80)
    want = orc.decode_frame(p, kind, fields, 3, orc.uniform_thresholds(kind), orc.OUT_RGBS if fmt == "rgbs"
                            else orc.OUT_RGB48)
    assert got.shape == want.shape
    d = np.abs(got.astype(np.float64) - want.astype(np.float64))
    assert d.max() <= (F32_TOL if fmt == "rgbs" else 1), d.max()
    if fmt == "rgb48":
        f = C.c_void_p()
        abi.check(abi.lib().chd_decode_frame(dec.h, 3, C.byref(f)), "chd_decode_frame")
        ptrs = []
        for plane in (abi.CHD_PLANE_R, abi.CHD_PLANE_G, abi.CHD_PLANE_B):
            ptr, stride = C.c_void_p(), C.c_ssize_t()
            abi.check(abi.lib().chd_frame_get_plane(f, plane, C.byref(ptr), C.byref(stride)), "chd_frame_get_plane")
            assert stride.value == 144 * 6
            ptrs.append(ptr.value)
        assert ptrs[1] - ptrs[0] == 2 and ptrs[2] - ptrs[0] == 4
        abi.lib().chd_frame_free(f)


@pytest.mark.parametrize("kind", [7, 8])
@pytest.mark.parametrize("t", [0.0, 0.25, 1.5])
def test_threshold_values(orc, patch_capture, kind, t):
    path, fields = patch_capture
    geo = dict(first_line=260, last_line=331, avs=520, ave=680)
    v = abi.Video(path)
    dec = abi.Decoder(v, kind, output_format="yuv444ps", transform_threshold=t,
                      first_active_frame_line=260, last_active_frame_line=331,
                      first_active_sample=520, last_active_sample=679).commit()
    _check(dec.decode(2), _oracle(orc, kind, fields, 2, "yuv444ps", thr=orc.uniform_thresholds(kind, t), **geo),
           "yuv444ps")


@pytest.mark.parametrize("kind", [7, 8])
def test_thresholds_file(orc, patch_capture, tmp_path, kind):
    path, fields = patch_capture
    n = 768 if kind == 8 else 80
    vals = np.random.default_rng(9).uniform(0.1, 0.7, n)
    tf = tmp_path / "thr.txt"
    tf.write_text("\n".join(f"{x:.17g}" for x in vals))
    geo = dict(first_line=260, last_line=331, avs=520, ave=680)
    v = abi.Video(path)
    dec = abi.Decoder(v, kind, output_format="yuv444p16", transform_thresholds_file=str(tf),
                      first_active_frame_line=260, last_active_frame_line=331,
                      first_active_sample=520, last_active_sample=679).commit()
    _check(dec.decode(4), _oracle(orc, kind, fields, 4, "yuv444p16", thr=vals * vals, **geo), "yuv444p16")
    # a missing file leaves the thresholds uniform (registry.cpp:85-92)
    dec2 = abi.Decoder(v, kind, output_format="yuv444p16", transform_thresholds_file=str(tmp_path / "none.txt"),
                       first_active_frame_line=260, last_active_frame_line=331,
                       first_active_sample=520, last_active_sample=679).commit()
    _check(dec2.decode(4), _oracle(orc, kind, fields, 4, "yuv444p16", **geo), "yuv444p16")
    bad = tmp_path / "bad.txt"
    bad.write_text("0.4 0.4")
    dec3 = abi.Decoder(v, kind, transform_thresholds_file=str(bad))
    with pytest.raises(abi.ChdError) as e:
        dec3.commit()
    assert e.value.name == "CHD_E_INVALID_ARG"


def test_gain_phase_gray(orc, patch_capture):
    path, fields = patch_capture
    geo = dict(first_line=200, last_line=263, avs=400, ave=528)
    crop = dict(first_active_frame_line=200, last_active_frame_line=263, first_active_sample=400,
                last_active_sample=527)
    v = abi.Video(path)
    dec = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_3D, output_format="yuv444ps", chroma_gain=1.7,
                      chroma_phase_deg=-12.5, **crop).commit()
    _check(dec.decode(1), _oracle(orc, 8, fields, 1, "yuv444ps", chroma_gain=1.7, chroma_phase_deg=-12.5, **geo),
           "yuv444ps")
    g = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_3D, output_format="grays", **crop).commit().decode(1)
    assert g.shape == (1, 64, 128)
    _check(g, _oracle(orc, 8, fields, 1, "yuv444ps", **geo)[:1], "grays")
    g16 = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_2D, output_format="gray16", **crop).commit().decode(1)
    _check(g16, _oracle(orc, 7, fields, 1, "yuv444p16", **geo)[:1], "gray16")
    p2 = abi.Decoder(v, abi.CHD_DEC_PAL_2D, output_format="yuv444ps", chroma_gain=0.6, chroma_phase_deg=33.0,
                     **crop).commit()
    _check(p2.decode(1), _oracle(orc, 6, fields, 1, "yuv444ps", chroma_gain=0.6, chroma_phase_deg=33.0, **geo),
           "yuv444ps")


def test_out_of_range_and_reuse(patch_capture):
    import ctypes as C
    path, _ = patch_capture
    v = abi.Video(path)
    dec = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_3D, output_format="yuv444ps").commit()
    f = C.c_void_p()
    assert abi.lib().chd_decode_frame(dec.h, 6, C.byref(f)) == 13
    assert abi.lib().chd_decode_frame(dec.h, -1, C.byref(f)) == 13
    a = dec.decode(3)
    dec.decode(0)
    dec.decode(5)
    b = dec.decode(3)  # field cache reuse and eviction leave results unchanged
    assert np.array_equal(a, b)
    # a second commit is a no-op; options after commit are rejected
    dec.commit()
    with pytest.raises(abi.ChdError):
        dec.set_option("transform_threshold", 0.3)


def test_frame_outlives_its_decoder(patch_capture):
    """A frame's planes stay valid after chd_decoder_free (frames are
    caller-owned; their pinned buffers return to the pool on chd_frame_free)."""
    import ctypes as C
    path, _ = patch_capture
    v = abi.Video(path)
    want = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_3D, output_format="yuv444ps").commit().decode(2)
    dec = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_3D, output_format="yuv444ps").commit()
    f = C.c_void_p()
    abi.check(abi.lib().chd_decode_frame(dec.h, 2, C.byref(f)), "chd_decode_frame")
    dec.close()
    fr = abi.Frame(f)
    try:
        assert np.array_equal(fr.planes(), want)
    finally:
        fr.free()


@pytest.mark.parametrize("nr", [0.0, 1.5])
@pytest.mark.parametrize("kind", [6, 7, 8])
def test_async_matches_sync(patch_capture, kind, nr):
    """Async workers (each with its own plan, stream and, with luma coring,
    raw planes) give the synchronous frames bit for bit."""
    path, _ = patch_capture
    v = abi.Video(path)
    dec = abi.Decoder(v, kind, output_format="yuv444p16", thread_count=3, luma_nr_level=nr).commit()
    res = dec.decode_async([5, 0, 3, 1, 4, 2, 3])
    assert sorted(res) == [0, 1, 2, 3, 4, 5]
    for i in range(6):
        assert np.array_equal(res[i], dec.decode(i))
    import ctypes as C
    c = C.c_void_p()
    abi.lib().chd_cancel_create(C.byref(c))
    abi.lib().chd_cancel_request(c)
    res2 = dec.decode_async([0, 1], cancel=c)
    assert res2 == {0: "CHD_E_CANCELLED", 1: "CHD_E_CANCELLED"}
    abi.lib().chd_cancel_free(c)


def test_async_one_worker_two_in_flight_keeps_order(patch_capture):
    """One async worker keeps two frames in flight (one downloading while the
    next decodes, into alternating device buffers): every frame is still the
    synchronous one, delivered in list order, and a callback may call
    chd_decode_frame itself."""
    import ctypes as C
    import threading
    path, _ = patch_capture
    v = abi.Video(path)
    dec = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_3D, output_format="yuv444ps", thread_count=1).commit()
    want = {i: dec.decode(i) for i in range(6)}
    order, got, inner = [], {}, []
    mu = threading.Lock()

    def done(user, status, idx, frame):
        fr = abi.Frame(C.c_void_p(frame))
        with mu:
            order.append(idx)
            got.setdefault(idx, []).append(fr.planes())
        fr.free()
        if idx == 3 and not inner:
            inner.append(dec.decode(4))   # re-entrant synchronous decode

    cb = abi.FRAME_DONE_CB(done)
    idx = [0, 1, 2, 3, 4, 5, 5, 2, 0]
    arr = (C.c_int64 * len(idx))(*idx)
    abi.check(abi.lib().chd_decode_frames_async(dec.h, arr, len(idx), cb, None, None), "chd_decode_frames_async")
    assert order == idx
    for i, frames in got.items():
        for f in frames:
            assert np.array_equal(f, want[i])
    assert np.array_equal(inner[0], want[4])


@pytest.mark.parametrize("kind", [6, 7, 8])
def test_device_api_matches_abi_bitwise(patch_capture, kind):
    """chd_tpal_decode shares each 8-field slab between frames f and f+2;
    the result must be bit-identical to per-frame chd_decode_frame."""
    import torch
    path, fields = patch_capture
    dev = torch.device("cuda:0")
    plan = abi.TpalPlan(abi.pal_tpal_params(kind=kind, pixel_format=abi.CHD_PIXEL_YUV444PS, max_batch_frames=4))
    src = torch.from_numpy(fields.view(np.int16)).to(dev)
    packed = torch.empty((fields.shape[0], plan.packed_field_elems), dtype=torch.int16, device=dev)
    plan.pack(src.data_ptr(), fields.shape[0], packed.data_ptr())
    out = torch.zeros((6, 3, 576, 922), dtype=torch.float32, device=dev)
    # two runs: frames 0..3 and 4..5 (the first run also spans two chunks of 4)
    plan.decode([(packed.data_ptr(), 6, 0, 4), (packed.data_ptr(), 6, 4, 2)], out.data_ptr())
    torch.cuda.synchronize()
    v = abi.Video(path)
    dec = abi.Decoder(v, kind, output_format="yuv444ps").commit()
    host = out.cpu().numpy()
    for f in range(6):
        assert np.array_equal(host[f], dec.decode(f)), f


@pytest.mark.parametrize("kind", [7, 8])
def test_repeated_decodes_reuse_job_lists(patch_capture, kind):
    """A call that decodes the same frames into the same buffer as the plan's
    previous call skips the job-list upload; any other call uploads.  Decode
    A, A, B, A into one output buffer and compare each with a fresh plan."""
    import torch
    _, fields = patch_capture
    dev = torch.device("cuda:0")
    params = abi.pal_tpal_params(kind=kind, pixel_format=abi.CHD_PIXEL_YUV444PS, max_batch_frames=4)
    plan = abi.TpalPlan(params)
    src = torch.from_numpy(fields.view(np.int16)).to(dev)
    packed = torch.empty((fields.shape[0], plan.packed_field_elems), dtype=torch.int16, device=dev)
    plan.pack(src.data_ptr(), fields.shape[0], packed.data_ptr())
    out = torch.zeros((2, plan.frame_bytes // 4), dtype=torch.float32, device=dev)
    runs = {"A": [(packed.data_ptr(), 6, 1, 2)], "B": [(packed.data_ptr(), 6, 3, 2)]}

    def fresh(run):
        p = abi.TpalPlan(params)
        o = torch.zeros_like(out)
        p.decode(runs[run], o.data_ptr())
        torch.cuda.synchronize()
        p.close()
        return o.cpu().numpy()

    want = {r: fresh(r) for r in runs}
    assert not np.array_equal(want["A"], want["B"])
    for r in "AABA":
        out.zero_()
        plan.decode(runs[r], out.data_ptr())
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), want[r]), r
    plan.close()


def test_truncated_capture_decodes_black_neighbours(orc, patch_capture, tmp_path):
    """Fields the sidecar lists but the file lacks read as black frames
    (source_field.cpp:104-132)."""
    import shutil
    path, fields = patch_capture
    short = tmp_path / "short.tbc"
    with open(path, "rb") as a, open(short, "wb") as b:
        b.write(a.read(fields[0].nbytes * 10))  # 5 of 6 frames
    shutil.copy(path + ".json", str(short) + ".json")
    v = abi.Video(str(short))
    dec = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_3D, output_format="yuv444ps").commit()
    got = dec.decode(3)
    f2 = fields.copy()
    f2[10:] = 0x4000
    f2 = f2[:12]
    _check(got, _oracle(orc, 8, f2, 3, "yuv444ps"), "yuv444ps")


@pytest.mark.parametrize("kind", [7, 8])
def test_sqlite_sidecar_decodes_like_json(patch_capture, tmp_path, kind):
    """A capture whose sidecar is the `.tbc.db` form decodes bit for bit as
    with the `.tbc.json` form of the same metadata, dropout concealment
    included."""
    import json
    import shutil
    from libchromadec_amd import synth
    from test_sqlite_sidecar import _write_db
    _, fields = patch_capture
    drops = [[] for _ in range(fields.shape[0])]
    drops[4] = [(300, 340, 100)]
    jpath, dpath = str(tmp_path / "j.tbc"), str(tmp_path / "d.tbc")
    side = synth.write_tbc(jpath, fields, dropouts=drops)
    with open(side) as fh:
        meta = json.load(fh)
    shutil.copy(jpath, dpath)
    _write_db(dpath + ".db", meta)
    decs = []
    for p in (jpath, dpath):
        d = abi.Decoder(abi.Video(p), kind, output_format="yuv444ps")
        d.set_dropout(True)
        decs.append(d.commit())
    for f in (1, 2, 4):
        assert np.array_equal(decs[0].decode(f), decs[1].decode(f)), f
