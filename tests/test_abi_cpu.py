"""The C-ABI library on the CPU: it loads, exports exactly what the headers
declare, and its host logic (sidecar parsing, option registry, commit
validation, error detail) behaves like the reference's ABI
(src/abi/c_api_decoder.cpp, c_api_video.cpp, c_api_frame.cpp).  No decode
happens here: commit needs a HIP device, and without one it must fail loudly.
"""
import ctypes as C
import json
import os
import re
import subprocess

import numpy as np
import pytest

from libchromadec_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include", "chromadec")


def declared_functions():
    names = set()
    for fn in os.listdir(INCLUDE):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(chd_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    # typedef'd callback types are not functions
    return {n for n in names if not n.endswith("_cb") and not n.endswith("_fn")}


def test_library_exports_every_declared_symbol():
    lib = abi.lib()
    declared = declared_functions()
    assert len(declared) > 40
    for name in sorted(declared):
        assert hasattr(lib, name), name
        assert name in abi.SIGNATURES, f"{name} has no ctypes binding"
    out = subprocess.run(["nm", "-D", "--defined-only", abi.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert exported == declared, (exported ^ declared)


def test_headers_compile_as_c11_and_cxx(tmp_path):
    src = tmp_path / "t.c"
    src.write_text("#include <chromadec/chromadec.h>\n#include <chromadec/tpal_device.h>\n"
                   "#include <chromadec/tpal_mask.h>\n"
                   "int main(void){return (int)sizeof(chd_video_info_t) > 0 ? 0 : 1;}\n")
    for cmd in (["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-c", str(src),
                 "-o", str(tmp_path / "a.o")],
                ["g++", "-std=c++17", "-Wall", "-Werror", "-x", "c++", "-I", os.path.join(ROOT, "include"), "-c",
                 str(src), "-o", str(tmp_path / "b.o")]):
        subprocess.run(cmd, check=True)


def test_struct_layouts_match_reference_abi():
    # field order/types of types.h (sizes on LP64)
    assert C.sizeof(abi.VideoInfo) == 104
    assert C.sizeof(abi.FrameInfo) == 24
    assert C.sizeof(abi.OutputInfo) == 24
    assert C.sizeof(abi.PlaneInfo) == 12


def test_version_status_and_log():
    L = abi.lib()
    maj, mn, pat = C.c_int(), C.c_int(), C.c_int()
    L.chd_version(C.byref(maj), C.byref(mn), C.byref(pat))
    assert (maj.value, mn.value, pat.value) == (0, 2, 1)
    assert L.chd_has_feature(b"rocm") == 1 and L.chd_has_feature(b"fftw") == 0
    assert L.chd_status_str(13) == b"CHD_E_OUT_OF_RANGE"
    assert L.chd_log_level_str(2) == b"WARN"
    assert L.chd_log_is_enabled(4) == 0
    old = L.chd_get_log_level()
    L.chd_set_log_level(9)
    assert L.chd_get_log_level() == 4
    L.chd_set_log_level(old)
    assert L.chd_init() == 0


def test_log_callback_receives_returned_errors(tmp_path):
    L = abi.lib()
    got = []
    CB = C.CFUNCTYPE(None, C.c_int32, C.c_uint, C.c_char_p, C.c_void_p)
    cb = CB(lambda lvl, flags, msg, user: got.append((lvl, flags, msg.decode())))
    L.chd_set_log_callback(C.cast(cb, C.c_void_p), None)
    try:
        with pytest.raises(abi.ChdError):
            abi.Video(str(tmp_path / "missing.tbc"))
    finally:
        L.chd_set_log_callback(None, None)
    assert any(flags == 1 and "missing.tbc" in msg for _, flags, msg in got)


def test_open_and_info(bars_capture):
    path, fields = bars_capture
    v = abi.Video(path)
    i = v.info()
    assert (i.standard, i.field_width, i.field_height) == (2, 1135, 313)
    assert (i.first_active_sample, i.last_active_sample) == (185, 1106)
    assert (i.first_active_frame_line, i.last_active_frame_line) == (44, 619)
    assert i.num_frames == fields.shape[0] // 2
    assert i.fsc_hz == (283.75 * 15625) + 25
    assert i.black_16b_ire == 0x4000 and i.white_16b_ire == 0xD300
    out = C.c_int32()
    L = abi.lib()
    assert L.chd_video_frame_line_to_signal_line(v.h, 44, C.byref(out)) == 0 and out.value == 23
    assert L.chd_video_frame_line_to_signal_line(v.h, 45, C.byref(out)) == 0 and out.value == 313 + 23
    assert L.chd_video_signal_line_to_frame_line(v.h, 336, C.byref(out)) == 0 and out.value == 45
    assert L.chd_video_frame_line_to_signal_line(v.h, 625, C.byref(out)) == 13
    for fl in range(0, 625):
        L.chd_video_frame_line_to_signal_line(v.h, fl, C.byref(out))
        back = C.c_int32()
        L.chd_video_signal_line_to_frame_line(v.h, out.value, C.byref(back))
        assert back.value == fl


def test_open_errors(tmp_path, bars_capture):
    with pytest.raises(abi.ChdError) as e:
        abi.Video(str(tmp_path / "nope.tbc"))
    assert e.value.name == "CHD_E_FILE_NOT_FOUND"
    raw = tmp_path / "raw.tbc"
    raw.write_bytes(b"\0" * 1000)
    with pytest.raises(abi.ChdError) as e:
        abi.Video(str(raw))
    assert e.value.name == "CHD_E_METADATA_MISSING"
    (tmp_path / "raw.tbc.json").write_text("{ not json")
    with pytest.raises(abi.ChdError) as e:
        abi.Video(str(raw))
    assert e.value.name == "CHD_E_METADATA_CORRUPT"
    path, _ = bars_capture
    meta = json.load(open(path + ".json"))
    meta["videoParameters"]["numberOfSequentialFields"] += 1
    bad = tmp_path / "bad.json"
    bad.write_text(json.dumps(meta))
    with pytest.raises(abi.ChdError) as e:
        abi.Video(path, str(bad))
    assert e.value.name == "CHD_E_METADATA_CORRUPT"
    (tmp_path / "x.tbc.db").write_bytes(b"SQLite format 3\0")
    (tmp_path / "x.tbc").write_bytes(b"\0" * 16)
    with pytest.raises(abi.ChdError) as e:
        abi.Video(str(tmp_path / "x.tbc"))
    # a truncated SQLite sidecar fails to read (c_api_video.cpp:526-530)
    assert e.value.name == "CHD_E_METADATA_CORRUPT"


def test_option_registry(bars_capture):
    path, _ = bars_capture
    v = abi.Video(path)
    L = abi.lib()
    d3 = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_3D)
    d7 = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_2D)
    dpal = abi.Decoder(v, 6)
    for name in ("transform_threshold", "output_format", "thread_count", "chroma_gain", "first_active_sample"):
        assert L.chd_decoder_has_option(d3.h, name.encode()) == 0
    assert L.chd_decoder_has_option(dpal.h, b"transform_threshold") == 1
    assert L.chd_decoder_has_option(d3.h, b"comb_adapt_threshold") == 1
    assert L.chd_decoder_has_option(d3.h, b"transform_mask") == 0
    assert L.chd_decoder_has_option(d7.h, b"transform_mask") == 1
    # wrong type for a known name
    assert L.chd_decoder_set_option_i32(d3.h, b"transform_threshold", 1) == 1
    assert b"not valid for this decoder kind" in L.chd_last_error()
    d3.set_option("transform_threshold", 0.5)
    d3.set_option("output_format", "yuv444ps")
    d3.set_option("padding_multiple", 8)


@pytest.mark.parametrize("opts,status", [
    (dict(output_format="bogus"), "CHD_E_INVALID_ARG"),
    (dict(output_clamp="bogus"), "CHD_E_INVALID_ARG"),
    (dict(padding_multiple=0), "CHD_E_INVALID_ARG"),
    (dict(padding_multiple=2000), "CHD_E_INVALID_ARG"),
    (dict(first_active_sample=1200), "CHD_E_INVALID_ARG"),
    (dict(last_active_frame_line=625), "CHD_E_INVALID_ARG"),
    (dict(output_format="yuv440p16"), "CHD_E_INVALID_ARG"),
    (dict(color_difference_precision="new"), "CHD_E_INVALID_ARG"),
    (dict(output_format="rgb48"), "CHD_E_UNSUPPORTED"),
    (dict(), "CHD_E_UNSUPPORTED"),  # valid options; no HIP device here
])
def test_commit_validation(bars_capture, opts, status):
    path, _ = bars_capture
    v = abi.Video(path)
    d = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_3D, **opts)
    with pytest.raises(abi.ChdError) as e:
        d.commit()
    assert e.value.name == status
    assert str(e.value).count("chd_decoder_commit") >= 1


def test_commit_kind_checks(bars_capture):
    path, _ = bars_capture
    v = abi.Video(path)
    with pytest.raises(abi.ChdError) as e:
        abi.Decoder(v, 4).commit()  # CHD_DEC_NTSC_3D: not built
    assert e.value.name == "CHD_E_DECODER_UNKNOWN"
    # CHD_DEC_PAL_2D, and CHD_DEC_AUTO (resolves to it for PAL), are built:
    # here they fail only for want of a HIP device
    for kind in (abi.CHD_DEC_PAL_2D, abi.CHD_DEC_AUTO):
        with pytest.raises(abi.ChdError) as e:
            abi.Decoder(v, kind).commit()
        assert e.value.name == "CHD_E_UNSUPPORTED"
    none = abi.Decoder(v, 13).commit()  # CHD_DEC_NONE: geometry only
    oi = none.output_info()
    assert (oi.width, oi.height, oi.num_planes, oi.num_frames) == (922, 576, 3, 3)
    f = C.c_void_p()
    assert abi.lib().chd_decode_frame(none.h, 0, C.byref(f)) == 9
    pad = abi.Decoder(v, 13, padding_multiple=16, output_format="grays").commit()
    oi = pad.output_info()
    assert (oi.width, oi.height, oi.num_planes) == (928, 576, 1)


def test_decode_before_commit_and_null_args(bars_capture):
    path, _ = bars_capture
    v = abi.Video(path)
    d = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_3D)
    f = C.c_void_p()
    L = abi.lib()
    assert L.chd_decode_frame(d.h, 0, C.byref(f)) == 1
    assert L.chd_decode_frame(None, 0, C.byref(f)) == 1
    assert L.chd_frame_get_info(None, None) == 1
    L.chd_frame_free(None)
    L.chd_video_free(None)
    L.chd_decoder_free(None)
    c = C.c_void_p()
    assert L.chd_cancel_create(C.byref(c)) == 0
    assert L.chd_cancel_is_requested(c) == 0
    L.chd_cancel_request(c)
    assert L.chd_cancel_is_requested(c) == 1
    L.chd_cancel_free(c)


def test_device_plan_without_gpu_fails_loudly():
    with pytest.raises(abi.ChdError) as e:
        abi.TpalPlan(abi.pal_tpal_params())
    assert e.value.name == "CHD_E_UNSUPPORTED"
    with pytest.raises(abi.ChdError) as e:
        abi.TpalPlan(abi.pal_tpal_params(kind=abi.CHD_DEC_PAL_2D))
    assert e.value.name == "CHD_E_UNSUPPORTED"
    bad = abi.pal_tpal_params(kind=4)   # CHD_DEC_NTSC_3D
    with pytest.raises(abi.ChdError) as e:
        abi.TpalPlan(bad)
    assert e.value.name == "CHD_E_DECODER_UNKNOWN"


def test_both_fields_marked_first_warns(tmp_path, bars_capture):
    """A frame whose second field is also marked isFirstField is paired as
    is, with a warning (LdDecodeMetaData::getFieldNumber, core.cpp:1173-1176);
    the frame count is unchanged (core.cpp:1092-1106)."""
    import json
    from libchromadec_amd import synth
    _, fields = bars_capture
    path = str(tmp_path / "order.tbc")
    side = synth.write_tbc(path, fields)
    with open(side) as fh:
        meta = json.load(fh)
    meta["fields"][3]["isFirstField"] = True
    with open(side, "w") as fh:
        json.dump(meta, fh)
    L = abi.lib()
    got = []
    CB = C.CFUNCTYPE(None, C.c_int32, C.c_uint, C.c_char_p, C.c_void_p)
    cb = CB(lambda lvl, flags, msg, user: got.append((lvl, msg.decode())))
    L.chd_set_log_callback(C.cast(cb, C.c_void_p), None)
    try:
        v = abi.Video(path)
        assert v.info().num_frames == fields.shape[0] // 2
        a, b = fields[0].copy(), fields[1].copy()
        v.correct_dropouts(0, a, b)
        assert not any("isFirstField" in m for _, m in got)
        a, b = fields[2].copy(), fields[3].copy()
        v.correct_dropouts(1, a, b)
    finally:
        L.chd_set_log_callback(None, None)
    warn = [m for lvl, m in got if lvl == 2 and "isFirstField" in m]
    assert warn and "fields 3 and 4" in warn[0], got
