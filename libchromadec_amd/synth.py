"""Synthetic PAL composite captures (ld-decode ``.tbc`` + ``.tbc.json``).

A vectorised PAL encoder for line-locked 4fSC sampling.  It follows the
signal model of the reference's legacy encoder,
``src/legacy/ld-chroma-decoder/encoder/palencoder.cpp:42-324`` and
``encoder.cpp:116-198`` (read as a specification; no code is shared):

* 4fSC = 17 734 475 Hz, 1135 x 313 samples per field, line-locked (0H at
  sample 0 of every line), ``prevCycles = prevLines * 283.7516`` (the 25 Hz
  offset), V-switch alternating per line, +-135 degree burst of 3/7
  black-to-white peak-to-peak from 5.6 us for 10 cycles, a 13-tap Gaussian
  1.3 MHz U/V low-pass, raised-cosine luma/chroma gates and sync pulses;
* output ``(VBS + C) * (white - black) + black`` clamped to
  [0x0100, 0xFEFF] and truncated to uint16 (``encoder.cpp:175-179``).

On top of that model it can add white Gaussian noise at a given SNR, a
per-line subcarrier phase jitter and sample dropouts (recorded in the
sidecar's ``dropOuts`` lists), which the stress workloads use.

The band-limited U/V that is actually modulated is returned as the chroma
ground truth for PSNR measurements.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

FSC = 4433618.75
SAMPLE_RATE = 4 * FSC
FIELD_WIDTH = 1135
FIELD_HEIGHT = 313
BURST_START, BURST_END = 98, 138
ACTIVE_START, ACTIVE_END = 185, 1107
BLACK, WHITE = 0x4000, 0xD300
# Input picture placement (palencoder.cpp:114-117).
PIC_WIDTH = 928
PIC_LEFT = (ACTIVE_START + ACTIVE_END) // 2 - PIC_WIDTH // 2
PIC_TOP = 44
PIC_HEIGHT = 620 - PIC_TOP
# 13-tap Gaussian U/V low-pass (scipy.signal.gaussian(13, 1.52) normalised).
_g = np.exp(-0.5 * (np.arange(13) - 6) ** 2 / 1.52 ** 2)
UV_LPF = _g / _g.sum()

# R'G'B' -> Y'UV (palencoder.cpp:290-292).
RGB_TO_YUV = np.array([[0.299, 0.587, 0.114],
                       [-0.147141, -0.288869, 0.436010],
                       [0.614975, -0.514965, -0.100010]])


def _gate(t, start, end, half_rise):
    """raisedCosineGate (encoder.h:111-124), vectorised over t and/or start/end."""
    t, start, end = np.broadcast_arrays(t, start, end)
    g = np.zeros(t.shape)
    rise = (t >= start - half_rise) & (t < start + half_rise)
    g = np.where(rise, 0.5 + 0.5 * np.sin((math.pi / 2) * ((t - start) / half_rise)), g)
    flat = (t >= start + half_rise) & (t < end - half_rise)
    g = np.where(flat, 1.0, g)
    fall = (t >= end - half_rise) & (t < end + half_rise)
    g = np.where(fall, 0.5 - 0.5 * np.sin((math.pi / 2) * ((t - end) / half_rise)), g)
    return g


def _sync_lengths(frame_lines):
    """Left/right sync pulse lengths per frame line (palencoder.cpp:227-249)."""
    fl = np.asarray(frame_lines)
    normal, eq, broad = 4.7e-6, 4.7e-6 / 2, 64.0e-6 / 2 - 4.7e-6
    left = np.full(fl.shape, normal)
    left = np.where(fl < 5, broad, left)
    left = np.where((fl >= 5) & (fl < 10), eq, left)
    left = np.where(fl >= 620, eq, left)
    right = np.zeros(fl.shape)
    right = np.where(fl < 4, broad, right)
    right = np.where((fl >= 4) & (fl < 9), eq, right)
    right = np.where((fl >= 619) & (fl < 624), eq, right)
    right = np.where(fl == 624, broad, right)
    return left, right, (left != normal)


@dataclass
class EncodeOptions:
    snr_db: float | None = None          # AWGN relative to white-black
    phase_jitter_deg: float = 0.0        # per-line RMS subcarrier phase jitter
    rng: np.random.Generator | None = None


def encode_field(field_no: int, y_pic, u_pic, v_pic, opts: EncodeOptions | None = None,
                 return_truth: bool = False):
    """Encode one field.  y/u/v_pic are [PIC_HEIGHT, PIC_WIDTH] Y'UV planes
    (Y 0..1) for the frame the field belongs to.  Returns uint16 [313, 1135]
    and, optionally, the truth: the band-limited U and V [313, 1135] and the
    modulated chroma actually added to the signal, in 16-bit sample units
    without the black level [313, 1135]."""
    opts = opts or EncodeOptions()
    off = field_no % 2
    frame_lines = np.arange(off, 2 * FIELD_HEIGHT, 2)           # 313 lines
    field_id = field_no % 8
    prev_lines = (field_id // 2) * 625 + (field_id % 2) * 313 + frame_lines // 2
    prev_cycles = np.mod(prev_lines * 283.7516, 1.0)
    vsw = np.where(prev_lines % 2 == 0, 1.0, -1.0)
    burst_off = vsw * 135.0 * math.pi / 180.0
    left_len, right_len, not_normal = _sync_lengths(frame_lines)
    burst_amp = np.full(frame_lines.shape, 3.0 / 7.0)
    burst_amp = np.where(not_normal, 0.0, burst_amp)
    burst_amp = np.where(frame_lines == 619, 0.0, burst_amp)
    burst_amp = np.where((vsw < 0) & np.isin(frame_lines, (10, 11, 618)), 0.0, burst_amp)

    half_luma = 2.0 / (4.0 * FSC)
    half_chroma = 3.0 / (4.0 * FSC)
    act_start = np.full(frame_lines.shape, ACTIVE_START / SAMPLE_RATE - 2 * half_chroma)
    act_end = np.full(frame_lines.shape, ACTIVE_END / SAMPLE_RATE + 2 * half_chroma)
    act_start = np.where(frame_lines == 44, 42.5e-6, act_start)
    act_end = np.where(frame_lines == 619, 30.35e-6, act_end)

    Y = np.zeros((FIELD_HEIGHT, FIELD_WIDTH))
    U = np.zeros_like(Y)
    V = np.zeros_like(Y)
    pic_rows = frame_lines - PIC_TOP
    has_pic = (pic_rows >= 0) & (pic_rows < PIC_HEIGHT)
    rows = np.nonzero(has_pic)[0]
    Y[rows, PIC_LEFT:PIC_LEFT + PIC_WIDTH] = y_pic[pic_rows[rows]]
    U[rows, PIC_LEFT:PIC_LEFT + PIC_WIDTH] = u_pic[pic_rows[rows]]
    V[rows, PIC_LEFT:PIC_LEFT + PIC_WIDTH] = v_pic[pic_rows[rows]]
    # 13-tap low-pass along each line, zeros outside (firfilter.h apply).
    def lpf(a):
        p = np.pad(a, ((0, 0), (6, 6)))
        out = np.zeros_like(a)
        for j, c in enumerate(UV_LPF):
            out += c * p[:, j:j + FIELD_WIDTH]
        return out
    U = lpf(U)
    V = lpf(V)

    t = np.arange(FIELD_WIDTH) / SAMPLE_RATE
    a = 2.0 * math.pi * (FSC * t[None, :] + prev_cycles[:, None])
    if opts.phase_jitter_deg:
        rng = opts.rng or np.random.default_rng()
        a = a + np.deg2rad(rng.normal(0.0, opts.phase_jitter_deg, size=(FIELD_HEIGHT, 1)))
    burst = np.sin(a + burst_off[:, None]) * burst_amp[:, None] / 2.0
    chroma = U * np.sin(a) + V * np.cos(a) * vsw[:, None]
    burst_gate = _gate(t[None, :], 5.6e-6, 5.6e-6 + 10.0 / FSC, 300e-9 / 2)
    chroma_gate = _gate(t[None, :], act_start[:, None], act_end[:, None], half_chroma)
    luma_gate = _gate(t[None, :], act_start[:, None], act_end[:, None], half_luma)
    C = burst * burst_gate + np.clip(chroma, -chroma_gate, chroma_gate)
    sync_level = -0.3 / 0.7
    lsync = _gate(t[None, :], 0.0, left_len[:, None], 200e-9 / 2)
    rsync = np.where(right_len[:, None] > 0,
                     _gate(t[None, :], 32e-6, 32e-6 + right_len[:, None], 200e-9 / 2), 0.0)
    VBS = np.clip(Y, -luma_gate, luma_gate) + sync_level * (lsync + rsync)
    sig = VBS + C
    dummy = frame_lines == 625
    sig[dummy] = 0.0
    scale = WHITE - BLACK
    out = sig * scale + BLACK
    if opts.snr_db is not None:
        rng = opts.rng or np.random.default_rng()
        sigma = scale * 10.0 ** (-opts.snr_db / 20.0)
        out = out + rng.normal(0.0, sigma, size=out.shape)
    out = np.clip(out, 0x0100, 0xFEFF).astype(np.uint16)
    if return_truth:
        return out, U, V, np.clip(chroma, -chroma_gate, chroma_gate) * scale
    return out


def rgb_to_yuv(rgb):
    """[..., 3] R'G'B' -> Y, U, V planes."""
    yuv = rgb @ RGB_TO_YUV.T
    return yuv[..., 0], yuv[..., 1], yuv[..., 2]


# ---------------------------------------------------------------------------
# Content generators (all return R'G'B' [PIC_HEIGHT, PIC_WIDTH, 3] in [0, 1])
# ---------------------------------------------------------------------------

BARS = np.array([[1, 1, 1], [1, 1, 0], [0, 1, 1], [0, 1, 0],
                 [1, 0, 1], [1, 0, 0], [0, 0, 1], [0, 0, 0]], dtype=np.float64)


def bars_gradient(frame: int, speed: float = 2.0):
    """Eight full-amplitude colour bars on the top two thirds, a diagonal
    luma gradient moving `speed` px/frame below."""
    img = np.zeros((PIC_HEIGHT, PIC_WIDTH, 3))
    cols = np.minimum((np.arange(PIC_WIDTH) * 8) // PIC_WIDTH, 7)
    split = (PIC_HEIGHT * 2) // 3
    img[:split] = BARS[cols][None, :, :]
    yy, xx = np.mgrid[split:PIC_HEIGHT, 0:PIC_WIDTH]
    g = np.mod((xx + yy + speed * frame) / 256.0, 1.0)
    img[split:] = g[..., None]
    return img


@dataclass
class PatchScene:
    """Static bars background with moving Gaussian-blurred colour patches."""
    seed: int
    n_patches: int | None = None
    max_speed: float = 4.0
    zone_plate: bool = False
    patches: list = field(default_factory=list)

    def __post_init__(self):
        rng = np.random.default_rng(self.seed)
        n = self.n_patches or int(rng.integers(6, 11))
        for _ in range(n):
            speed = rng.uniform(0.0, self.max_speed)
            ang = rng.uniform(0, 2 * math.pi)
            self.patches.append(dict(
                x=rng.uniform(0, PIC_WIDTH), y=rng.uniform(0, PIC_HEIGHT),
                vx=speed * math.cos(ang), vy=speed * math.sin(ang),
                sigma=rng.uniform(8, 40), rgb=rng.uniform(0, 1, 3),
                alpha=rng.uniform(0.6, 1.0)))

    def frame(self, f: int):
        img = np.zeros((PIC_HEIGHT, PIC_WIDTH, 3))
        cols = np.minimum((np.arange(PIC_WIDTH) * 8) // PIC_WIDTH, 7)
        img[:] = 0.75 * BARS[cols][None, :, :]
        yy, xx = np.mgrid[0:PIC_HEIGHT, 0:PIC_WIDTH]
        for p in self.patches:
            cx = np.mod(p["x"] + p["vx"] * f, PIC_WIDTH)
            cy = np.mod(p["y"] + p["vy"] * f, PIC_HEIGHT)
            w = p["alpha"] * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * p["sigma"] ** 2))
            img = img * (1 - w[..., None]) + p["rgb"][None, None, :] * w[..., None]
        if self.zone_plate:
            # radial chirp reaching 0.282 cycles/sample (5 MHz) at the corners
            r2 = ((xx - PIC_WIDTH / 2) ** 2 + (yy - PIC_HEIGHT / 2) ** 2)
            kmax = 0.282 / math.hypot(PIC_WIDTH / 2, PIC_HEIGHT / 2)
            zp = 0.5 + 0.5 * np.cos(math.pi * kmax * r2)
            zmask = (xx > PIC_WIDTH // 2)
            img = np.where(zmask[..., None], 0.5 * img + 0.5 * zp[..., None], img)
        return np.clip(img, 0.0, 1.0)


def encode_sequence(frames_rgb, opts: EncodeOptions | None = None, return_truth: bool = False,
                    first_field_no: int = 0):
    """Encode an iterable of R'G'B' pictures; returns uint16 [2N, 313, 1135]
    (and the U, V and modulated-chroma truth, [2N, 313, 1135] float each, if
    requested)."""
    fields, tu, tv, tc = [], [], [], []
    for i, rgb in enumerate(frames_rgb):
        y, u, v = rgb_to_yuv(rgb)
        for k in range(2):
            r = encode_field(first_field_no + 2 * i + k, y, u, v, opts, return_truth)
            if return_truth:
                fields.append(r[0]); tu.append(r[1]); tv.append(r[2]); tc.append(r[3])
            else:
                fields.append(r)
    f = np.stack(fields)
    if return_truth:
        return f, np.stack(tu), np.stack(tv), np.stack(tc)
    return f


def add_dropouts(fields, rng: np.random.Generator, per_field=(1, 3), length=(20, 400)):
    """Corrupt runs of samples in the active area and return the dropout
    lists per field as (startx, endx, fieldLine 1-based) triples."""
    drops = []
    for fi in range(fields.shape[0]):
        n = int(rng.integers(per_field[0], per_field[1] + 1))
        lst = []
        for _ in range(n):
            ln = int(rng.integers(length[0], length[1] + 1))
            line = int(rng.integers(23, 309))
            sx = int(rng.integers(ACTIVE_START, max(ACTIVE_START + 1, ACTIVE_END - ln)))
            ex = min(sx + ln, FIELD_WIDTH)
            fields[fi, line, sx:ex] = rng.integers(0x0100, 0x2000, size=ex - sx)
            lst.append((sx, ex, line + 1))
        drops.append(lst)
    return drops


def write_tbc(path: str, fields: np.ndarray, dropouts=None) -> str:
    """Write fields to `path` (.tbc) and the ld-decode JSON sidecar next to
    it (`path`.json) with the keys ld_metadata_json.cpp:65-113,233-255 reads.
    Returns the sidecar path."""
    fields = np.ascontiguousarray(fields, dtype="<u2")
    fields.tofile(path)
    n = fields.shape[0]
    flist = []
    for i in range(n):
        rec = {"seqNo": i + 1, "isFirstField": (i % 2) == 0, "syncConf": 100,
               "medianBurstIRE": 100.0 * (3.0 / 7.0) / 2.0, "fieldPhaseID": 0,
               "audioSamples": 0}
        if dropouts is not None and dropouts[i]:
            rec["dropOuts"] = {"startx": [d[0] for d in dropouts[i]],
                               "endx": [d[1] for d in dropouts[i]],
                               "fieldLine": [d[2] for d in dropouts[i]]}
        flist.append(rec)
    meta = {"videoParameters": {
        "numberOfSequentialFields": n, "system": "PAL", "isSubcarrierLocked": False,
        "isWidescreen": False, "colourBurstStart": BURST_START, "colourBurstEnd": BURST_END,
        "activeVideoStart": ACTIVE_START, "activeVideoEnd": ACTIVE_END,
        "white16bIre": WHITE, "black16bIre": BLACK, "blanking16bIre": BLACK,
        "fieldWidth": FIELD_WIDTH, "fieldHeight": FIELD_HEIGHT,
        "sampleRate": SAMPLE_RATE, "isMapped": False, "tapeFormat": ""},
        "fields": flist}
    side = path + ".json"
    with open(side, "w") as fh:
        json.dump(meta, fh)
    return side


def read_tbc(path: str) -> np.ndarray:
    raw = np.fromfile(path, dtype="<u2")
    return raw.reshape(-1, FIELD_HEIGHT, FIELD_WIDTH)


def make_capture(directory: str, name: str, frames_rgb, opts: EncodeOptions | None = None,
                 dropout_rng: np.random.Generator | None = None):
    """Encode and write a capture; returns (tbc_path, fields, dropouts)."""
    fields = encode_sequence(frames_rgb, opts)
    drops = add_dropouts(fields, dropout_rng) if dropout_rng is not None else None
    path = os.path.join(directory, name + ".tbc")
    write_tbc(path, fields, drops)
    return path, fields, drops
