"""ctypes bindings for libchromadec_amd/lib/libchromadec.so.

The shared library exports the libchromadec C ABI (include/chromadec/*.h);
this module binds it for Python callers, tests and bench.py, in the way a
host-language binding of the reference would (the reference ships a Rust
one, rust/chromadec-sys).  There is no Python fallback: if the library is
not built, importing the bindings raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libchromadec.so")

# enums (include/chromadec/*.h)
CHD_OK = 0
STATUS = {0: "CHD_OK", 1: "CHD_E_INVALID_ARG", 2: "CHD_E_FILE_NOT_FOUND", 3: "CHD_E_IO",
          4: "CHD_E_FORMAT_UNSUPPORTED", 5: "CHD_E_METADATA_MISSING", 6: "CHD_E_METADATA_CORRUPT",
          7: "CHD_E_PRESET_UNKNOWN", 8: "CHD_E_DECODER_UNKNOWN", 9: "CHD_E_DECODER_INCOMPATIBLE",
          10: "CHD_E_NN_MODEL_LOAD", 11: "CHD_E_NN_BACKEND_UNAVAILABLE", 12: "CHD_E_NN_INFERENCE",
          13: "CHD_E_OUT_OF_RANGE", 14: "CHD_E_CANCELLED", 15: "CHD_E_INTERNAL", 16: "CHD_E_OOM",
          17: "CHD_E_UNSUPPORTED"}
CHD_DEC_AUTO = 0
CHD_DEC_PAL_2D = 6
CHD_DEC_TRANSFORM_2D = 7
CHD_DEC_TRANSFORM_3D = 8
CHD_PIXEL_YUV444P16, CHD_PIXEL_YUV444PS, CHD_PIXEL_GRAY16, CHD_PIXEL_GRAYS = 0, 1, 4, 5
CHD_PIXEL_RGB48, CHD_PIXEL_RGBS = 2, 3
CHD_PLANE_Y, CHD_PLANE_CB, CHD_PLANE_CR = 0, 1, 2
CHD_PLANE_R, CHD_PLANE_G, CHD_PLANE_B = 3, 4, 5
FLOAT_FORMATS = (CHD_PIXEL_YUV444PS, CHD_PIXEL_GRAYS, CHD_PIXEL_RGBS)


class VideoInfo(C.Structure):
    _fields_ = [("standard", C.c_int32), ("encoding", C.c_int32), ("signal_state", C.c_int32),
                ("layout", C.c_int32), ("field_width", C.c_int32), ("field_height", C.c_int32),
                ("samples_per_frame", C.c_int32), ("sample_rate_hz", C.c_double), ("fsc_hz", C.c_double),
                ("first_active_sample", C.c_int32), ("last_active_sample", C.c_int32),
                ("first_active_frame_line", C.c_int32), ("last_active_frame_line", C.c_int32),
                ("black_16b_ire", C.c_int32), ("white_16b_ire", C.c_int32), ("blanking_16b_ire", C.c_int32),
                ("num_frames", C.c_int64), ("is_widescreen", C.c_int), ("is_subcarrier_locked", C.c_int),
                ("is_first_field_first", C.c_int)]


class VideoParamsOverride(C.Structure):
    _fields_ = [("standard", C.c_int32), ("encoding", C.c_int32), ("signal_state", C.c_int32),
                ("layout", C.c_int32), ("is_subcarrier_locked", C.c_int), ("is_second_field_first", C.c_int)]


class FrameInfo(C.Structure):
    _fields_ = [("format", C.c_int32), ("width", C.c_int32), ("height", C.c_int32),
                ("num_planes", C.c_int32), ("frame_index", C.c_int64)]


class PlaneInfo(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("first_frame_row", C.c_int32)]


class OutputInfo(C.Structure):
    _fields_ = [("format", C.c_int32), ("width", C.c_int32), ("height", C.c_int32),
                ("num_planes", C.c_int32), ("num_frames", C.c_int64)]


class TpalParams(C.Structure):
    _fields_ = [("kind", C.c_int32), ("field_width", C.c_int32), ("field_height", C.c_int32),
                ("active_video_start", C.c_int32), ("active_video_end", C.c_int32),
                ("first_active_frame_line", C.c_int32), ("last_active_frame_line", C.c_int32),
                ("colour_burst_start", C.c_int32), ("colour_burst_end", C.c_int32),
                ("black_16b_ire", C.c_int32), ("white_16b_ire", C.c_int32),
                ("fsc_hz", C.c_double), ("sample_rate_hz", C.c_double),
                ("threshold", C.c_double), ("thresholds", C.POINTER(C.c_double)),
                ("chroma_gain", C.c_double), ("chroma_phase_deg", C.c_double),
                ("ecb_from_u", C.c_double), ("ecr_from_v", C.c_double),
                ("pixel_format", C.c_int32), ("clamp", C.c_int32),
                ("out_width", C.c_int32), ("out_height", C.c_int32),
                ("pad_left", C.c_int32), ("pad_top", C.c_int32),
                ("max_batch_frames", C.c_int32), ("device", C.c_int32),
                ("kr", C.c_double), ("kb", C.c_double), ("luma_nr_level", C.c_double)]


class TpalSequence(C.Structure):
    _fields_ = [("d_fields", C.c_void_p), ("n_frames", C.c_int64),
                ("first_frame", C.c_int64), ("count", C.c_int64)]


class DropoutOpts(C.Structure):
    _fields_ = [("enabled", C.c_int), ("overcorrect", C.c_int), ("intra_field_only", C.c_int),
                ("reserved", C.c_void_p * 4)]


class DropoutStats(C.Structure):
    _fields_ = [("corrected", C.c_int32), ("failed", C.c_int32), ("total_distance", C.c_int64)]


FRAME_DONE_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_int32, C.c_int64, C.c_void_p)

_lib = None
_lock = threading.Lock()

# Every function include/chromadec/*.h declares: name -> (restype, argtypes).
_P = C.c_void_p
SIGNATURES = {
    "chd_init": (C.c_int32, []),
    "chd_shutdown": (None, []),
    "chd_version_string": (C.c_char_p, []),
    "chd_version": (None, [C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "chd_has_feature": (C.c_int, [C.c_char_p]),
    "chd_status_str": (C.c_char_p, [C.c_int32]),
    "chd_last_error": (C.c_char_p, []),
    "chd_clear_last_error": (None, []),
    "chd_set_log_callback": (None, [_P, _P]),
    "chd_log_to_stderr": (None, []),
    "chd_set_log_level": (None, [C.c_int32]),
    "chd_get_log_level": (C.c_int32, []),
    "chd_log_is_enabled": (C.c_int, [C.c_int32]),
    "chd_log_level_str": (C.c_char_p, [C.c_int32]),
    "chd_video_open_composite": (C.c_int32, [C.c_char_p, C.c_char_p, C.POINTER(VideoParamsOverride),
                                             C.POINTER(_P)]),
    "chd_video_free": (None, [_P]),
    "chd_video_get_info": (C.c_int32, [_P, C.POINTER(VideoInfo)]),
    "chd_video_frame_line_to_signal_line": (C.c_int32, [_P, C.c_int32, C.POINTER(C.c_int32)]),
    "chd_video_signal_line_to_frame_line": (C.c_int32, [_P, C.c_int32, C.POINTER(C.c_int32)]),
    "chd_decoder_create": (C.c_int32, [_P, C.c_int32, C.POINTER(_P)]),
    "chd_decoder_free": (None, [_P]),
    "chd_decoder_set_option_f64": (C.c_int32, [_P, C.c_char_p, C.c_double]),
    "chd_decoder_set_option_i32": (C.c_int32, [_P, C.c_char_p, C.c_int32]),
    "chd_decoder_set_option_bool": (C.c_int32, [_P, C.c_char_p, C.c_int]),
    "chd_decoder_set_option_str": (C.c_int32, [_P, C.c_char_p, C.c_char_p]),
    "chd_decoder_has_option": (C.c_int32, [_P, C.c_char_p]),
    "chd_decoder_commit": (C.c_int32, [_P]),
    "chd_decoder_get_output_info": (C.c_int32, [_P, C.POINTER(OutputInfo)]),
    "chd_decode_frame": (C.c_int32, [_P, C.c_int64, C.POINTER(_P)]),
    "chd_decode_frames_async": (C.c_int32, [_P, C.POINTER(C.c_int64), C.c_size_t, FRAME_DONE_CB, _P, _P]),
    "chd_cancel_create": (C.c_int32, [C.POINTER(_P)]),
    "chd_cancel_request": (None, [_P]),
    "chd_cancel_is_requested": (C.c_int, [_P]),
    "chd_cancel_free": (None, [_P]),
    "chd_frame_get_info": (C.c_int32, [_P, C.POINTER(FrameInfo)]),
    "chd_frame_get_plane_info": (C.c_int32, [_P, C.c_int32, C.POINTER(PlaneInfo)]),
    "chd_frame_get_plane": (C.c_int32, [_P, C.c_int32, C.POINTER(_P), C.POINTER(C.c_ssize_t)]),
    "chd_frame_get_plane_float": (C.c_int32, [_P, C.c_int32, C.POINTER(_P), C.POINTER(C.c_ssize_t)]),
    "chd_frame_free": (None, [_P]),
    "chd_tpal_plan_create": (C.c_int32, [C.POINTER(TpalParams), C.POINTER(_P)]),
    "chd_tpal_plan_free": (None, [_P]),
    "chd_tpal_packed_field_elems": (C.c_size_t, [_P]),
    "chd_tpal_frame_bytes": (C.c_size_t, [_P]),
    "chd_tpal_pack_fields": (C.c_int32, [_P, _P, C.c_int64, _P, _P]),
    "chd_tpal_decode": (C.c_int32, [_P, C.POINTER(TpalSequence), C.c_int32, _P, _P]),
    "chd_tpal_decode_static": (C.c_int32, [_P, C.POINTER(TpalSequence), C.c_int32, _P, _P]),
    "chd_tpal_set_timing": (C.c_int32, [_P, C.c_int]),
    "chd_tpal_decode_separated": (C.c_int32, [_P, _P, _P, C.c_int64, _P, _P]),
    "chd_tpal_batch_create": (C.c_int32, [_P, C.POINTER(TpalSequence), C.c_int32, _P, C.c_int32, C.POINTER(_P)]),
    "chd_tpal_batch_run": (C.c_int32, [_P, _P]),
    "chd_tpal_batch_config": (C.c_int32, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                          C.POINTER(C.c_double)]),
    "chd_tpal_batch_set_shape": (C.c_int32, [_P, C.c_int32, C.c_int32]),
    "chd_tpal_batch_free": (None, [_P]),
    "chd_tpal_last_timing": (C.c_int32, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                         C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    # dropout.h
    "chd_decoder_set_dropout": (C.c_int32, [_P, C.POINTER(DropoutOpts)]),
    "chd_decoder_get_last_dropout_stats": (C.c_int32, [_P, C.POINTER(DropoutStats)]),
    "chd_video_correct_dropouts": (C.c_int32, [_P, C.c_int64, C.POINTER(DropoutOpts), _P, _P,
                                               C.POINTER(DropoutStats)]),
    # tpal_mask.h
    "chd_mask_net_create": (C.c_int32, [C.c_int32, C.c_int64, C.POINTER(_P)]),
    "chd_mask_net_free": (None, [_P]),
    "chd_mask_net_set_params": (C.c_int32, [_P, C.POINTER(C.c_float), C.c_int64]),
    "chd_mask_net_get_params": (C.c_int32, [_P, C.POINTER(C.c_float), C.c_int64]),
    "chd_mask_net_load": (C.c_int32, [_P, C.c_char_p]),
    "chd_mask_net_save": (C.c_int32, [_P, C.c_char_p]),
    "chd_mask_net_forward": (C.c_int32, [_P, _P, C.c_int64, _P, _P]),
    "chd_mask_net_gradient": (C.c_int32, [_P, _P, _P, C.c_int64, C.c_float, _P, C.POINTER(C.c_double), _P]),
    "chd_mask_net_adam_step": (C.c_int32, [_P, _P, C.c_float, _P]),
    "chd_tpal_tiles_per_slab": (C.c_int64, [_P]),
    "chd_tpal_slab_spectra": (C.c_int32, [_P, _P, C.c_int64, C.c_int64, C.c_int32, _P, _P]),
    "chd_tpal_set_mask_net": (C.c_int32, [_P, _P]),
}

MASK_NET_PARAMS = 32033
TILE_BINS = 1280
TILE_VISITS = 768


def _share_hip_runtime_with_torch():
    """PyTorch-ROCm wheels carry their own libamdhip64, loaded by the
    unversioned name "libamdhip64.so".  If this library (which needs
    "libamdhip64.so.7") loaded /opt/rocm's copy first, torch would load a
    second HIP runtime and find no GPU.  Importing torch first makes the
    dynamic linker reuse torch's runtime (same SONAME) for this library."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            _share_hip_runtime_with_torch()
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(f"{LIB_PATH} is not built; run make -C libchromadec_amd")
            L = C.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


class ChdError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        self.name = STATUS.get(status, str(status))
        detail = lib().chd_last_error().decode(errors="replace")
        super().__init__(f"{where}: {self.name}: {detail}")


def check(status: int, where: str):
    if status != CHD_OK:
        raise ChdError(status, where)


class Video:
    def __init__(self, path: str, metadata_path: str | None = None, second_field_first: bool = False):
        self.h = C.c_void_p()
        ov = None
        if second_field_first:
            ov = VideoParamsOverride(0, 0, 0, 0, 0, 1)
        check(lib().chd_video_open_composite(path.encode(), metadata_path.encode() if metadata_path else None,
                                             C.byref(ov) if ov is not None else None, C.byref(self.h)),
              "chd_video_open_composite")

    def info(self) -> VideoInfo:
        i = VideoInfo()
        check(lib().chd_video_get_info(self.h, C.byref(i)), "chd_video_get_info")
        return i

    def correct_dropouts(self, frame_index: int, first: np.ndarray, second: np.ndarray, overcorrect=False,
                         intra_field_only=False):
        """Conceal the frame's dropouts in place in host copies of its fields."""
        assert first.dtype == np.uint16 and second.dtype == np.uint16
        assert first.flags.c_contiguous and second.flags.c_contiguous
        o = DropoutOpts(1, int(overcorrect), int(intra_field_only))
        st = DropoutStats()
        check(lib().chd_video_correct_dropouts(self.h, frame_index, C.byref(o), first.ctypes.data, second.ctypes.data,
                                               C.byref(st)), "chd_video_correct_dropouts")
        return st.corrected, st.failed, st.total_distance

    def close(self):
        if self.h:
            lib().chd_video_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Frame:
    """Owns a chd_frame; planes() copies the borrowed planes into numpy."""

    def __init__(self, handle):
        self.h = handle

    def info(self) -> FrameInfo:
        i = FrameInfo()
        check(lib().chd_frame_get_info(self.h, C.byref(i)), "chd_frame_get_info")
        return i

    def planes(self) -> np.ndarray:
        """[planes, height, width]; RGB48 (interleaved) as [height, width, 3]."""
        inf = self.info()
        is_float = inf.format in FLOAT_FORMATS
        if inf.format == CHD_PIXEL_RGB48:
            ptr, stride = C.c_void_p(), C.c_ssize_t()
            check(lib().chd_frame_get_plane(self.h, CHD_PLANE_R, C.byref(ptr), C.byref(stride)),
                  "chd_frame_get_plane")
            buf = (C.c_char * (stride.value * inf.height)).from_address(ptr.value)
            rows = np.frombuffer(buf, dtype=np.uint16).reshape(inf.height, stride.value // 2)
            return rows[:, :3 * inf.width].reshape(inf.height, inf.width, 3).copy()
        first = CHD_PLANE_R if inf.format == CHD_PIXEL_RGBS else CHD_PLANE_Y
        out = np.zeros((inf.num_planes, inf.height, inf.width), dtype=np.float32 if is_float else np.uint16)
        for p in range(inf.num_planes):
            ptr = C.c_void_p()
            stride = C.c_ssize_t()
            fn = lib().chd_frame_get_plane_float if is_float else lib().chd_frame_get_plane
            check(fn(self.h, first + p, C.byref(ptr), C.byref(stride)), fn.__name__)
            es = 4 if is_float else 2
            buf = (C.c_char * (stride.value * inf.height)).from_address(ptr.value)
            arr = np.frombuffer(buf, dtype=np.float32 if is_float else np.uint16)
            out[p] = arr.reshape(inf.height, stride.value // es)[:, :inf.width]
        return out

    def free(self):
        if self.h:
            lib().chd_frame_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Decoder:
    def __init__(self, video: Video, kind: int, **options):
        self.video = video
        self.h = C.c_void_p()
        check(lib().chd_decoder_create(video.h, kind, C.byref(self.h)), "chd_decoder_create")
        for k, v in options.items():
            self.set_option(k, v)

    def set_option(self, name: str, value):
        L = lib()
        if isinstance(value, bool):
            st = L.chd_decoder_set_option_bool(self.h, name.encode(), int(value))
        elif isinstance(value, int):
            st = L.chd_decoder_set_option_i32(self.h, name.encode(), value)
        elif isinstance(value, float):
            st = L.chd_decoder_set_option_f64(self.h, name.encode(), value)
        else:
            st = L.chd_decoder_set_option_str(self.h, name.encode(), str(value).encode())
        check(st, f"chd_decoder_set_option({name})")

    def set_dropout(self, enabled=True, overcorrect=False, intra_field_only=False):
        o = DropoutOpts(int(enabled), int(overcorrect), int(intra_field_only))
        check(lib().chd_decoder_set_dropout(self.h, C.byref(o)), "chd_decoder_set_dropout")
        return self

    def last_dropout_stats(self):
        st = DropoutStats()
        check(lib().chd_decoder_get_last_dropout_stats(self.h, C.byref(st)), "chd_decoder_get_last_dropout_stats")
        return st.corrected, st.failed, st.total_distance

    def commit(self):
        check(lib().chd_decoder_commit(self.h), "chd_decoder_commit")
        return self

    def output_info(self) -> OutputInfo:
        i = OutputInfo()
        check(lib().chd_decoder_get_output_info(self.h, C.byref(i)), "chd_decoder_get_output_info")
        return i

    def decode(self, index: int) -> np.ndarray:
        f = C.c_void_p()
        check(lib().chd_decode_frame(self.h, index, C.byref(f)), "chd_decode_frame")
        fr = Frame(f)
        try:
            return fr.planes()
        finally:
            fr.free()

    def decode_async(self, indices, cancel=None, keep=True):
        """Returns {index: planes or status name}; with keep=False each frame
        is freed in the callback unread and the value is "CHD_OK"."""
        res = {}
        mu = threading.Lock()

        def done(user, status, idx, frame):
            if status == CHD_OK and not keep:
                lib().chd_frame_free(C.c_void_p(frame))
                with mu:
                    res[idx] = "CHD_OK"
                return
            with mu:
                if status == CHD_OK:
                    fr = Frame(C.c_void_p(frame))
                    res[idx] = fr.planes()
                    fr.free()
                else:
                    res[idx] = STATUS.get(status, str(status))

        cb = FRAME_DONE_CB(done)
        arr = (C.c_int64 * len(indices))(*indices)
        check(lib().chd_decode_frames_async(self.h, arr, len(indices), cb, None, cancel), "chd_decode_frames_async")
        return res

    def close(self):
        if self.h:
            lib().chd_decoder_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def color_scales(color_difference="modern", broadcast_scaling="scientific"):
    """(ecb_from_u, ecr_from_v) as the reference's color_conversion.cpp:33-67."""
    kr, kb = (0.30, 0.11) if color_difference == "classic" else (0.299, 0.114)
    ur, vr = {"classic": (0.493, 0.877), "modern": (0.492111, 0.877283)}.get(
        broadcast_scaling, (0.49211104112248356308804691718185, 0.87728321993817866838972487283129))
    return 1.0 / (2.0 * (1.0 - kb) * ur), 1.0 / (2.0 * (1.0 - kr) * vr)


def pal_tpal_params(kind=CHD_DEC_TRANSFORM_3D, pixel_format=CHD_PIXEL_YUV444PS, threshold=0.4,
                    max_batch_frames=64, first_line=44, last_line=619, avs=185, ave=1107, device=0,
                    black=0x4000, white=0xD300, clamp=0, luma_nr_level=0.0):
    """Device-plan parameters for the line-locked 4fSC PAL geometry."""
    ecb, ecr = color_scales()
    return TpalParams(kind, 1135, 313, avs, ave, first_line, last_line, 98, 138, black, white,
                      4433618.75, 17734475.0, threshold, None, 1.0, 0.0, ecb, ecr, pixel_format, clamp,
                      ave - avs, last_line - first_line + 1, 0, 0, max_batch_frames, device, 0.299, 0.114,
                      luma_nr_level)


class TpalPlan:
    """chd_tpal_* device decode plan (pointers are device addresses, e.g.
    torch.Tensor.data_ptr() on cuda:0)."""

    def __init__(self, params: TpalParams):
        self.params = params
        self.h = C.c_void_p()
        check(lib().chd_tpal_plan_create(C.byref(params), C.byref(self.h)), "chd_tpal_plan_create")

    @property
    def packed_field_elems(self) -> int:
        return lib().chd_tpal_packed_field_elems(self.h)

    @property
    def frame_bytes(self) -> int:
        return lib().chd_tpal_frame_bytes(self.h)

    def pack(self, d_src: int, n_fields: int, d_dst: int, stream: int = 0):
        check(lib().chd_tpal_pack_fields(self.h, d_src, n_fields, d_dst, stream), "chd_tpal_pack_fields")

    def decode(self, seqs, d_out: int, stream: int = 0):
        """seqs: list of (d_fields, n_frames, first_frame, count)."""
        arr = (TpalSequence * len(seqs))(*[TpalSequence(*s) for s in seqs])
        check(lib().chd_tpal_decode(self.h, arr, len(seqs), d_out, stream), "chd_tpal_decode")

    def decode_static(self, seqs, d_out: int, stream: int = 0):
        """chd_tpal_decode_static: the same decode with no device allocation
        (workspace in the library's static per-GPU scratch array)."""
        arr = (TpalSequence * len(seqs))(*[TpalSequence(*s) for s in seqs])
        check(lib().chd_tpal_decode_static(self.h, arr, len(seqs), d_out, stream), "chd_tpal_decode_static")

    def decode_separated(self, d_fields: int, d_chroma: int, n_frames: int, d_out: int, stream: int = 0):
        check(lib().chd_tpal_decode_separated(self.h, d_fields, d_chroma, n_frames, d_out, stream),
              "chd_tpal_decode_separated")

    def set_timing(self, on: bool):
        check(lib().chd_tpal_set_timing(self.h, int(on)), "chd_tpal_set_timing")

    def last_timing(self):
        a, b = C.c_double(), C.c_double()
        n1, n2 = C.c_int32(), C.c_int32()
        check(lib().chd_tpal_last_timing(self.h, C.byref(a), C.byref(b), C.byref(n1), C.byref(n2)),
              "chd_tpal_last_timing")
        return a.value, b.value, n1.value, n2.value

    @property
    def tiles_per_slab(self) -> int:
        return lib().chd_tpal_tiles_per_slab(self.h)

    def slab_spectra(self, d_fields: int, n_fields: int, first_slab: int, n_slabs: int, d_spec: int,
                     stream: int = 0):
        check(lib().chd_tpal_slab_spectra(self.h, d_fields, n_fields, first_slab, n_slabs, d_spec, stream),
              "chd_tpal_slab_spectra")

    def set_mask_net(self, net):
        check(lib().chd_tpal_set_mask_net(self.h, net.h if net is not None else None), "chd_tpal_set_mask_net")
        self._net = net  # keep the network alive while the plan uses it

    def close(self):
        if self.h:
            lib().chd_tpal_plan_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MaskNet:
    """chd_mask_net_* learned carrier-mirror mask (pointers are device addresses)."""

    def __init__(self, max_tiles: int, device: int = 0):
        self.h = C.c_void_p()
        check(lib().chd_mask_net_create(device, max_tiles, C.byref(self.h)), "chd_mask_net_create")

    def set_params(self, params: np.ndarray):
        p = np.ascontiguousarray(params, dtype=np.float32)
        check(lib().chd_mask_net_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float)), p.size),
              "chd_mask_net_set_params")

    def params(self) -> np.ndarray:
        p = np.zeros(MASK_NET_PARAMS, np.float32)
        check(lib().chd_mask_net_get_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float)), p.size),
              "chd_mask_net_get_params")
        return p

    def load(self, path: str):
        check(lib().chd_mask_net_load(self.h, path.encode()), "chd_mask_net_load")

    def save(self, path: str):
        check(lib().chd_mask_net_save(self.h, path.encode()), "chd_mask_net_save")

    def forward(self, d_spec: int, n_tiles: int, d_mask: int, stream: int = 0):
        check(lib().chd_mask_net_forward(self.h, d_spec, n_tiles, d_mask, stream), "chd_mask_net_forward")

    def gradient(self, d_spec: int, d_truth: int, n_tiles: int, scale: float, d_grad: int, stream: int = 0,
                 want_loss: bool = True):
        loss = C.c_double(0.0)
        check(lib().chd_mask_net_gradient(self.h, d_spec, d_truth, n_tiles, scale, d_grad,
                                          C.byref(loss) if want_loss else None, stream), "chd_mask_net_gradient")
        return loss.value

    def adam_step(self, d_grad: int, lr: float, stream: int = 0):
        check(lib().chd_mask_net_adam_step(self.h, d_grad, lr, stream), "chd_mask_net_adam_step")

    def close(self):
        if self.h:
            lib().chd_mask_net_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TpalBatch:
    """chd_tpal_batch_*: a fixed multi-sequence decode captured in a HIP graph."""

    def __init__(self, plan: TpalPlan, seqs, d_out: int, autotune: bool = True):
        arr = (TpalSequence * len(seqs))(*[TpalSequence(*q) for q in seqs])
        self.h = C.c_void_p()
        self._plan = plan
        check(lib().chd_tpal_batch_create(plan.h, arr, len(seqs), d_out, int(autotune), C.byref(self.h)),
              "chd_tpal_batch_create")

    def run(self, stream: int = 0):
        check(lib().chd_tpal_batch_run(self.h, stream), "chd_tpal_batch_run")

    def config(self):
        a, b, t = C.c_int32(), C.c_int32(), C.c_double()
        check(lib().chd_tpal_batch_config(self.h, C.byref(a), C.byref(b), C.byref(t)), "chd_tpal_batch_config")
        return a.value, b.value, t.value

    def set_shape(self, strip: int, groups: int):
        check(lib().chd_tpal_batch_set_shape(self.h, strip, groups), "chd_tpal_batch_set_shape")

    def close(self):
        if self.h:
            lib().chd_tpal_batch_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
