"""CPU oracle of dropout concealment (test infrastructure only).

A plain-Python restatement of DropoutCorrector for a single source
(src/dropout/dropout_corrector.cpp:41-530) and of the second-field shift
(src/decoders/source_field.cpp:51-66).  Only the tests may use it; the
library's own implementation is libchromadec_amd/csrc/dropout.cpp.

Fields are numpy uint16 arrays [field_height, field_width]; dropout runs
are (startx, endx, field_line) with endx exclusive and field_line 1-based.
`vp` is a dict with field_width, field_height, colour_burst_end,
active_video_end, first_active_frame_line, last_active_frame_line and
system ("PAL", "NTSC", "PAL-M" or "SECAM").
"""
from __future__ import annotations

import numpy as np

# luma_filters.cpp:36-52
PAL_LUMA = (0.03283437, 0.23959832, 0.45513461, 0.23959832, 0.03283437)
NTSC_LUMA = (0.03275786, 0.23955702, 0.45537024, 0.23955702, 0.03275786)


def shift_second_field(data, runs, black):
    """source_field.cpp:51-66: drop two samples from the start, append black."""
    flat = np.concatenate([data.ravel()[2:], np.full(2, black, np.uint16)]).reshape(data.shape)
    out = []
    for s, e, ln in runs:
        if e - 2 > 0:
            out.append((max(s - 2, 0), e - 2, ln))
    return flat, out


def _lowpass(vp, line):
    c = PAL_LUMA if vp["system"] in ("PAL", "SECAM") else NTSC_LUMA
    n = len(line)
    out = np.zeros(n, np.uint16)
    for i in range(n):
        v = 0.0
        for j in range(5):
            k = i - 2 + j
            if 0 <= k < n:
                v += c[j] * int(line[k])
        out[i] = int(v)   # truncation, as the uint16 assignment in firfilter.h
    return out


def _locations(vp, runs, overcorrect):
    out = []
    for s, e, ln in runs:
        if ln < 1 or ln > vp["field_height"]:
            continue
        if overcorrect:
            s = s - 24 if s > 24 else 0
            e = e + 24 if e < vp["field_width"] - 24 else vp["field_width"]
        out.append([ln, s, e, "unknown"])
    cbe, ave = vp["colour_burst_end"], vp["active_video_end"]
    while True:
        splits = 0
        n = len(out)
        for i in range(n):
            d = out[i]
            if d[1] <= cbe:
                d[3] = "burst"
                if d[2] > cbe:
                    out.append([d[0], cbe + 1, d[2], "burst"])
                    d[2] = cbe
                    splits += 1
            elif d[1] <= ave:
                d[3] = "visible"
                if d[2] > ave:
                    d[2] = ave
                    splits += 1
        if splits == 0:
            return out


def _find(vp, mine, other, idx, this_first, match_phase, burst, intra):
    if not match_phase:
        step, other_off = 1, -1
    elif vp["system"] in ("PAL", "PAL-M"):
        step, other_off = 4, (-3 if this_first else -1)
    else:
        step, other_off = 2, -1
    first = (vp["first_active_frame_line"] + 1) // 2
    last = vp["last_active_frame_line"] // 2
    t = mine[idx]
    cands = []

    def scan(src, same, offset, stp):
        line = t[0] + offset
        if line - 1 < first or line - 1 > last:
            return
        while line - 1 >= first and line <= last:
            if any(s[0] == line and t[2] - s[1] >= 0 and s[2] - t[1] >= 0 for s in src):
                line += stp
                continue
            cands.append((same, line))
            return

    scan(mine, True, 0, -step)
    scan(mine, True, step, step)
    if not burst and not intra:
        scan(other, False, other_off, -step)
        scan(other, False, other_off + step, step)
    best, best_d = None, None
    dl = 2 * t[0] + (0 if this_first else 1)
    for same, line in cands:
        sl = 2 * line + ((0 if this_first else 1) if same else (1 if this_first else 0))
        d = abs(dl - sl)
        if best is None or d < best_d:      # equal quality for one source: first wins
            best, best_d = (same, line), d
    return best, best_d


def correct_frame(vp, first, second, runs1, runs2, overcorrect=False, intra=False):
    """Returns corrected copies of both fields and (corrected, failed, distance)."""
    f1, f2 = first.copy(), second.copy()
    stats = [0, 0, 0]
    if not runs1 and not runs2:
        return f1, f2, tuple(stats)
    l1 = _locations(vp, runs1, overcorrect)
    l2 = _locations(vp, runs2, overcorrect)
    for mine, other, data, odata, this_first in ((l1, l2, f1, f2, True), (l2, l1, f2, f1, False)):
        for i, d in enumerate(mine):
            rep, crep, dist = None, None, 0
            if d[3] == "burst":
                rep, dist = _find(vp, mine, other, i, this_first, True, True, intra)
            elif d[3] == "visible":
                rep, dist = _find(vp, mine, other, i, this_first, False, False, intra)
                crep, _ = _find(vp, mine, other, i, this_first, True, False, intra)
            if rep is None:
                stats[1] += 1
                continue
            stats[0] += 1
            stats[2] += dist
            src = (data if rep[0] else odata)[rep[1] - 1]
            dst = data[d[0] - 1]
            s, e = d[1], d[2]
            if crep is None or (d[0] == rep[1] and d[0] == crep[1]):
                dst[s:e] = src[s:e]
                continue
            lf = _lowpass(vp, src)
            dst[s:e] = lf[s:e]
            cl = (data if crep[0] else odata)[crep[1] - 1]
            lfc = _lowpass(vp, cl)
            for x in range(s, e):
                dst[x] = (int(dst[x]) + int(cl[x]) - int(lfc[x])) & 0xFFFF
    return f1, f2, tuple(stats)
