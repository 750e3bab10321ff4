"""ctypes front end for the FP64 oracle (oracle/chd_oracle.c).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.  It is the checker, never the decoder.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")
_lib = None


class OrcParams(C.Structure):
    _fields_ = [("field_width", C.c_int32), ("field_height", C.c_int32),
                ("active_video_start", C.c_int32), ("active_video_end", C.c_int32),
                ("first_active_frame_line", C.c_int32), ("last_active_frame_line", C.c_int32),
                ("colour_burst_start", C.c_int32), ("colour_burst_end", C.c_int32),
                ("black16b", C.c_int32), ("white16b", C.c_int32),
                ("fsc", C.c_double), ("sample_rate", C.c_double),
                ("chroma_gain", C.c_double), ("chroma_phase_deg", C.c_double),
                ("ecb_from_u", C.c_double), ("ecr_from_v", C.c_double),
                ("clamp_mode", C.c_int32),
                ("pad_left", C.c_int32), ("pad_top", C.c_int32),
                ("out_width", C.c_int32), ("out_height", C.c_int32),
                ("kr", C.c_double), ("kb", C.c_double), ("luma_nr_level", C.c_double)]


# output formats of the oracle's OutputWriter: float Y'CbCr planes, u16
# Y'CbCr planes, RGB48 (interleaved u16), RGBS (float R'G'B' planes)
OUT_PS, OUT_P16, OUT_RGB48, OUT_RGBS = 0, 1, 2, 3


def _out_array(p, out_fmt, lead=()):
    dt = np.float32 if out_fmt in (OUT_PS, OUT_RGBS) else np.uint16
    shape = (p.out_height, p.out_width, 3) if out_fmt == OUT_RGB48 else (3, p.out_height, p.out_width)
    return np.zeros(tuple(lead) + shape, dtype=dt)


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B", "-C", _HERE], check=True, capture_output=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        P = C.POINTER
        d, i32 = C.c_double, C.c_int32
        _lib.orc_fft3d.argtypes = [P(d), P(d), C.c_int, C.c_int, C.c_int, C.c_int]
        _lib.orc_dft3d_naive.argtypes = [P(d), C.c_int, C.c_int, C.c_int, P(d), P(d)]
        _lib.orc_r2c3d.argtypes = [P(d), C.c_int, C.c_int, C.c_int, P(d), P(d), P(d), P(d)]
        _lib.orc_c2r3d.argtypes = [P(d), P(d), C.c_int, C.c_int, C.c_int, P(d), P(d), P(d)]
        _lib.orc_window.argtypes = [C.c_int, C.c_int]
        _lib.orc_window.restype = d
        _lib.orc_apply_filter3d.argtypes = [P(d), P(d), P(d), P(d), P(d)]
        _lib.orc_apply_filter2d.argtypes = [P(d), P(d), P(d), P(d), P(d)]
        _lib.orc_tpal3d_filter_fields.argtypes = [P(OrcParams), P(P(C.c_uint16)), P(i32), i32, i32, i32,
                                                  P(d), P(P(d))]
        _lib.orc_tpal2d_filter_fields.argtypes = _lib.orc_tpal3d_filter_fields.argtypes
        _lib.orc_pal_luts.argtypes = [P(OrcParams), P(d), P(d), P(i32), P(d), P(d)]
        _lib.orc_detect_burst.argtypes = [P(OrcParams), i32, P(C.c_uint16), P(d), P(d), P(d)]
        _lib.orc_palcolour_decode_frame.argtypes = [P(OrcParams), P(C.c_uint16), i32, P(C.c_uint16), i32,
                                                    P(d), P(d), P(d), P(d), P(d)]
        _lib.orc_output_float.argtypes = [P(OrcParams), P(d), P(d), P(d), P(C.c_float)]
        _lib.orc_output_u16.argtypes = [P(OrcParams), P(d), P(d), P(d), P(C.c_uint16)]
        _lib.orc_output_rgb48.argtypes = [P(OrcParams), P(d), P(d), P(d), P(C.c_uint16)]
        _lib.orc_output_rgbs.argtypes = [P(OrcParams), P(d), P(d), P(d), P(C.c_float)]
        _lib.orc_decode_window.argtypes = [P(OrcParams), i32, P(d), P(P(C.c_uint16)), P(i32), i32, i32, i32,
                                           i32, C.c_void_p, P(d), P(d)]
        _lib.orc_decode_window.restype = C.c_int
        _lib.orc_decode_window_masked.argtypes = [P(OrcParams), P(C.c_float), P(P(C.c_uint16)), P(i32), i32, i32,
                                                  i32, i32, C.c_void_p, P(d)]
        _lib.orc_decode_window_masked.restype = C.c_int
        _lib.orc_decode_sequence.argtypes = [P(OrcParams), i32, P(d), P(C.c_uint16), i32, P(i32), i32, i32,
                                             C.c_void_p, i32]
        _lib.orc_decode_sequence.restype = C.c_int
        _lib.orc_ynr_line.argtypes = [P(OrcParams), P(d)]
        _lib.orc_tile_spectra3d.argtypes = [P(OrcParams), P(P(d)), P(i32), d, P(d), P(d)]
        _lib.orc_tile_spectra3d.restype = C.c_int
    return _lib


def ynr_line(p: OrcParams, yline) -> np.ndarray:
    """PalColour::doYNR on one frame line of field_width Y samples (a copy)."""
    y = np.array(yline, dtype=np.float64)
    assert y.shape == (p.field_width,)
    lib().orc_ynr_line(C.byref(p), _dp(y))
    return y


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _u16p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint16))


# ---------------------------------------------------------------------------
# Colour conversion constants (color_conversion.cpp:16-67), default
# precisions "modern" colour difference and "scientific" broadcast scaling.
# ---------------------------------------------------------------------------

def color_conversion(color_difference: str = "modern", broadcast_scaling: str = "scientific"):
    kr, kb = (0.30, 0.11) if color_difference == "classic" else (0.299, 0.114)
    if broadcast_scaling == "classic":
        ur, vr = 0.493, 0.877
    elif broadcast_scaling == "modern":
        ur, vr = 0.492111, 0.877283
    else:
        ur, vr = 0.49211104112248356308804691718185, 0.87728321993817866838972487283129
    return 1.0 / (2.0 * (1.0 - kb) * ur), 1.0 / (2.0 * (1.0 - kr) * vr)


def pal_params(first_line=44, last_line=619, avs=185, ave=1107, black=0x4000, white=0xD300,
               clamp_mode=0, pad_left=0, pad_top=0, out_width=None, out_height=None,
               chroma_gain=1.0, chroma_phase_deg=0.0, color_difference="modern",
               broadcast_scaling="scientific", field_width=1135, field_height=313,
               burst=(98, 138), fsc=4433618.75, sample_rate=17734475.0, luma_nr_level=0.0) -> OrcParams:
    ecb, ecr = color_conversion(color_difference, broadcast_scaling)
    kr, kb = (0.30, 0.11) if color_difference == "classic" else (0.299, 0.114)
    aw, ah = ave - avs, last_line - first_line + 1
    return OrcParams(field_width, field_height, avs, ave, first_line, last_line, burst[0], burst[1],
                     black, white, fsc, sample_rate, chroma_gain, chroma_phase_deg, ecb, ecr,
                     clamp_mode, pad_left, pad_top, out_width or aw, out_height or ah, kr, kb, luma_nr_level)


def uniform_thresholds(kind: int, threshold: float = 0.4) -> np.ndarray:
    n = 768 if kind == 8 else 80
    return np.full(n, threshold * threshold, dtype=np.float64)


def _field_ptrs(fields):
    arr = (C.POINTER(C.c_uint16) * len(fields))()
    keep = []
    for i, f in enumerate(fields):
        f = np.ascontiguousarray(f, dtype=np.uint16)
        keep.append(f)
        arr[i] = _u16p(f)
    return arr, keep


def filter_fields(p: OrcParams, kind: int, fields, offsets, start, end, thr_sq):
    """TransformPal{2D,3D}::filterFields: returns (end-start, fh, fw) chroma."""
    L = lib()
    arr, keep = _field_ptrs(fields)
    offs = np.ascontiguousarray(offsets, dtype=np.int32)
    out = np.zeros((end - start, p.field_height, p.field_width))
    outp = (C.POINTER(C.c_double) * (end - start))(*[_dp(out[i]) for i in range(end - start)])
    thr = np.ascontiguousarray(thr_sq, dtype=np.float64)
    fn = L.orc_tpal3d_filter_fields if kind == 8 else L.orc_tpal2d_filter_fields
    fn(C.byref(p), arr, offs.ctypes.data_as(C.POINTER(C.c_int32)), len(fields), start, end, _dp(thr), outp)
    return out


def decode_window(p: OrcParams, kind: int, fields, offsets, start, end, thr_sq, out_fmt=0,
                  want_chroma=False, want_yuv=False):
    L = lib()
    arr, keep = _field_ptrs(fields)
    offs = np.ascontiguousarray(offsets, dtype=np.int32)
    thr = np.ascontiguousarray(thr_sq, dtype=np.float64)
    out = _out_array(p, out_fmt)
    chroma = np.zeros((2, p.field_height, p.field_width)) if want_chroma else None
    yuv = np.zeros((3, 2 * p.field_height - 1, p.field_width)) if want_yuv else None
    rc = L.orc_decode_window(C.byref(p), kind, _dp(thr), arr, offs.ctypes.data_as(C.POINTER(C.c_int32)),
                             len(fields), start, end, out_fmt, out.ctypes.data_as(C.c_void_p),
                             _dp(chroma) if chroma is not None else None,
                             _dp(yuv) if yuv is not None else None)
    if rc != 0:
        raise RuntimeError(f"orc_decode_window failed: {rc}")
    res = [out]
    if want_chroma:
        res.append(chroma)
    if want_yuv:
        res.append(yuv)
    return res[0] if len(res) == 1 else tuple(res)


def frame_window(seq_fields: np.ndarray, frame: int, kind: int, black: int = 0x4000):
    """The field window chd_decode_frame builds for 0-based `frame` of a
    regularly paired sequence (SourceField::loadFields, source_field.cpp:68-171):
    look-behind 2 / look-ahead 4 frames for 3D, none for 2D, black outside."""
    behind, ahead = (2, 4) if kind == 8 else (0, 0)
    nf = seq_fields.shape[0]
    blackf = np.full(seq_fields.shape[1:], black, dtype=np.uint16)
    fields, offs = [], []
    for w in range(2 * (behind + 1 + ahead)):
        fi = 2 * (frame - behind) + w
        fields.append(seq_fields[fi] if 0 <= fi < nf else blackf)
        offs.append(w % 2)
    return fields, offs, 2 * behind, 2 * behind + 2


def decode_frame(p: OrcParams, kind: int, seq_fields, frame, thr_sq, out_fmt=0, **kw):
    fields, offs, s, e = frame_window(seq_fields, frame, kind, p.black16b)
    return decode_window(p, kind, fields, offs, s, e, thr_sq, out_fmt, **kw)


def decode_frame_masked(p: OrcParams, seq_fields, frame, masks, out_fmt=0):
    """3D decode of `frame` with a soft mask per tile instead of the
    threshold: masks [2, n_tile_rows, n_tile_cols, 768] for the window's two
    Z tiles (slabs frame-2 and frame), visits v = (z*32 + y)*3 + (x-2)."""
    L = lib()
    fields, offs, s, e = frame_window(seq_fields, frame, 8, p.black16b)
    arr, keep = _field_ptrs(fields)
    o = np.ascontiguousarray(offs, dtype=np.int32)
    m = np.ascontiguousarray(masks, dtype=np.float32)
    out = _out_array(p, out_fmt)
    rc = L.orc_decode_window_masked(C.byref(p), m.ctypes.data_as(C.POINTER(C.c_float)), arr,
                                    o.ctypes.data_as(C.POINTER(C.c_int32)), len(fields), s, e, out_fmt,
                                    out.ctypes.data_as(C.c_void_p), None)
    if rc != 0:
        raise RuntimeError(f"orc_decode_window_masked failed: {rc}")
    return out


def tile_spectra3d(p: OrcParams, fields8, offsets8, fill: float):
    """Bins x = 2..6 of every 3D tile of one Z tile (8 fields as float
    arrays, their line offsets): complex [n_tile_rows * n_tile_cols, 5, 32, 8]
    (bin x = k + 2, ky, kz), tiles in lattice order; `fill` is the value of
    the lines the tile does not read from a field (black16bIre for a
    composite, 0 for chroma).  transform_pal_3d.cpp:123-202."""
    L = lib()
    fs = [np.ascontiguousarray(f, dtype=np.float64) for f in fields8]
    arr = (C.POINTER(C.c_double) * 8)(*[f.ctypes.data_as(C.POINTER(C.c_double)) for f in fs])
    n_ty = (p.last_active_frame_line - (p.first_active_frame_line - 16)) // 16 + 1
    n_tx = (p.active_video_end - p.active_video_start + 8 + 7) // 8
    re = np.zeros((n_ty * n_tx, 5, 32, 8))
    im = np.zeros_like(re)
    o = np.ascontiguousarray(offsets8, dtype=np.int32)
    L.orc_tile_spectra3d(C.byref(p), arr, o.ctypes.data_as(C.POINTER(C.c_int32)), fill,
                         re.ctypes.data_as(C.POINTER(C.c_double)), im.ctypes.data_as(C.POINTER(C.c_double)))
    return re + 1j * im


def decode_frame_from_chroma(p: OrcParams, field0, field1, chroma0, chroma1, out_fmt=0):
    """PalColour::decodeFrames with given chroma fields (palcolour.cpp:279-335)
    and OutputWriter: the demodulation half of a decode."""
    L = lib()
    fw, fh = p.field_width, p.field_height
    ca = fw * (2 * fh - 1)
    f0 = np.ascontiguousarray(field0, dtype=np.uint16)
    f1 = np.ascontiguousarray(field1, dtype=np.uint16)
    c0 = np.ascontiguousarray(chroma0, dtype=np.float64)
    c1 = np.ascontiguousarray(chroma1, dtype=np.float64)
    yuv = np.zeros(3 * ca)
    L.orc_palcolour_decode_frame(C.byref(p), _u16p(f0), 0, _u16p(f1), 1, _dp(c0), _dp(c1), _dp(yuv),
                                 _dp(yuv[ca:]), _dp(yuv[2 * ca:]))
    out = np.zeros((3, p.out_height, p.out_width), dtype=np.float32 if out_fmt == 0 else np.uint16)
    if out_fmt == 0:
        L.orc_output_float(C.byref(p), _dp(yuv), _dp(yuv[ca:]), _dp(yuv[2 * ca:]),
                           out.ctypes.data_as(C.POINTER(C.c_float)))
    else:
        L.orc_output_u16(C.byref(p), _dp(yuv), _dp(yuv[ca:]), _dp(yuv[2 * ca:]),
                         out.ctypes.data_as(C.POINTER(C.c_uint16)))
    return out


def decode_sequence(p: OrcParams, kind: int, seq_fields, frames, thr_sq, out_fmt=0, threads=0):
    L = lib()
    seq = np.ascontiguousarray(seq_fields, dtype=np.uint16)
    fr = np.ascontiguousarray(frames, dtype=np.int32)
    out = _out_array(p, out_fmt, (len(fr),))
    thr = np.ascontiguousarray(thr_sq, dtype=np.float64)
    rc = L.orc_decode_sequence(C.byref(p), kind, _dp(thr), _u16p(seq), seq.shape[0],
                               fr.ctypes.data_as(C.POINTER(C.c_int32)), len(fr), out_fmt,
                               out.ctypes.data_as(C.c_void_p), threads)
    if rc != 0:
        raise RuntimeError(f"orc_decode_sequence failed: {rc}")
    return out


# --- small helpers exposed for the oracle's own tests ----------------------

def fft3d(re, im, sign=-1):
    re = np.ascontiguousarray(re, dtype=np.float64).copy()
    im = np.ascontiguousarray(im, dtype=np.float64).copy()
    Z, Y, X = re.shape
    lib().orc_fft3d(_dp(re), _dp(im), Z, Y, X, sign)
    return re + 1j * im


def dft3d_naive(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    Z, Y, X = x.shape
    re = np.zeros_like(x); im = np.zeros_like(x)
    lib().orc_dft3d_naive(_dp(x), Z, Y, X, _dp(re), _dp(im))
    return re + 1j * im


def r2c3d(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    Z, Y, X = x.shape
    hr = np.zeros((Z, Y, X // 2 + 1)); hi = np.zeros_like(hr)
    sr = np.zeros_like(x); si = np.zeros_like(x)
    lib().orc_r2c3d(_dp(x), Z, Y, X, _dp(hr), _dp(hi), _dp(sr), _dp(si))
    return hr + 1j * hi


def c2r3d(h, shape):
    hr = np.ascontiguousarray(h.real, dtype=np.float64)
    hi = np.ascontiguousarray(h.imag, dtype=np.float64)
    Z, Y, X = shape
    out = np.zeros(shape); sr = np.zeros(shape); si = np.zeros(shape)
    lib().orc_c2r3d(_dp(hr), _dp(hi), Z, Y, X, _dp(out), _dp(sr), _dp(si))
    return out


def apply_filter(h, thr_sq, kind):
    hr = np.ascontiguousarray(h.real, dtype=np.float64)
    hi = np.ascontiguousarray(h.imag, dtype=np.float64)
    orr = np.zeros_like(hr); oi = np.zeros_like(hi)
    thr = np.ascontiguousarray(thr_sq, dtype=np.float64)
    fn = lib().orc_apply_filter3d if kind == 8 else lib().orc_apply_filter2d
    fn(_dp(hr), _dp(hi), _dp(orr), _dp(oi), _dp(thr))
    return orr + 1j * oi


def window(n, N):
    return lib().orc_window(n, N)


def pal_luts(p: OrcParams):
    s = np.zeros(1135); c = np.zeros(1135); fs = C.c_int32(0)
    cf = np.zeros((64, 4)); yf = np.zeros((64, 2))
    lib().orc_pal_luts(C.byref(p), _dp(s), _dp(c), C.byref(fs), _dp(cf), _dp(yf))
    return s, c, fs.value, cf[:fs.value + 1], yf[:fs.value + 1]


def detect_burst(p: OrcParams, field: np.ndarray, line: int):
    f = np.ascontiguousarray(field, dtype=np.uint16)
    bp, bq, v = C.c_double(), C.c_double(), C.c_double()
    lib().orc_detect_burst(C.byref(p), line, _u16p(f), C.byref(bp), C.byref(bq), C.byref(v))
    return bp.value, bq.value, v.value


def output_stage(p: OrcParams, yuv: np.ndarray, out_fmt: int) -> np.ndarray:
    """OutputWriter of a component frame yuv [3, 2*field_height-1,
    field_width] (composite-domain Y/U/V) in any of the OUT_* formats."""
    L = lib()
    yuv = np.ascontiguousarray(yuv, dtype=np.float64)
    ca = yuv.shape[1] * yuv.shape[2]
    flat = yuv.reshape(-1)
    out = _out_array(p, out_fmt)
    args = (C.byref(p), _dp(flat), _dp(flat[ca:]), _dp(flat[2 * ca:]))
    fn = {OUT_PS: L.orc_output_float, OUT_P16: L.orc_output_u16, OUT_RGB48: L.orc_output_rgb48,
          OUT_RGBS: L.orc_output_rgbs}[out_fmt]
    ptype = C.c_float if out.dtype == np.float32 else C.c_uint16
    fn(*args, out.ctypes.data_as(C.POINTER(ptype)))
    return out
