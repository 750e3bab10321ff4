"""Per-configuration error of the GPU decode against the FP64 oracle.

Run on a machine with the GPU (`python tests/parity_report.py`); prints a
markdown table with, for every benchmark configuration, the maximum
|GPU - oracle| of E'Y and of E'Cb/E'Cr (U/V) and the U/V RMS error, in
float planes (full scale: E'Y 0..1, E'Cb/E'Cr +-0.5) and in yuv444p16 LSB:

  2D    one synthetic PAL frame of bars plus a moving diagonal gradient,
        CHD_DEC_TRANSFORM_2D, threshold 0.4
  3D    the 64-frame moving-patch sequence at 40 dB (frames at both ends and
        inside), CHD_DEC_TRANSFORM_3D, threshold 0.4
  batch 16 sequences of 8-40 frames decoded as one captured graph (first,
        middle and last frame of every sequence)
  learned  the 64-frame sequence with the packaged learned mask; the oracle
        runs the same masks (the network's output) through its FP64 filter
  stress   24 frames of the stress capture (30 dB, phase jitter, fast
        motion, dropouts concealed by the library, whose concealment is also
        compared bit for bit with the oracle's), plain and learned
  PAL 2D   the 64-frame sequence with CHD_DEC_PAL_2D (the PALcolour 2D
        filter on the raw composite, no transform)
  luma coring  the 64-frame 3D decode with luma_nr_level 2.0

The oracle is the checker here; nothing it computes feeds the GPU path.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from libchromadec_amd import abi, synth  # noqa: E402
from oracle import dropout as od  # noqa: E402
from oracle import oracle as orc  # noqa: E402

H, W = 576, 922
N_TX3 = 117          # 3D tile columns of the PAL active line (922 samples from 177, hop 8)
WEIGHTS = os.path.join(ROOT, "libchromadec_amd", "data", "tpal3d_mask.bin")


def _torch():
    import torch
    return torch


def _stats(got, want, u16):
    d = got.astype(np.float64) - want.astype(np.float64)
    if d.ndim == 3:
        d = d[None]
    return {"y_max": float(np.abs(d[:, 0]).max()), "uv_max": float(np.abs(d[:, 1:]).max()),
            "uv_rms": float(np.sqrt(np.mean(d[:, 1:] ** 2))), "u16": u16}


def _merge(rows):
    return {"y_max": max(r["y_max"] for r in rows), "uv_max": max(r["uv_max"] for r in rows),
            "uv_rms": float(np.sqrt(np.mean([r["uv_rms"] ** 2 for r in rows]))), "u16": rows[0]["u16"]}


def _plan(kind, fmt, frames, **kw):
    pf = abi.CHD_PIXEL_YUV444PS if fmt == "ps" else abi.CHD_PIXEL_YUV444P16
    return abi.TpalPlan(abi.pal_tpal_params(kind=kind, pixel_format=pf, max_batch_frames=frames, **kw))


def _pack(plan, fields):
    torch = _torch()
    dev = torch.device("cuda:0")
    src = torch.from_numpy(np.ascontiguousarray(fields).view(np.int16)).to(dev)
    packed = torch.empty((fields.shape[0], plan.packed_field_elems), dtype=torch.int16, device=dev)
    plan.pack(src.data_ptr(), fields.shape[0], packed.data_ptr())
    return packed


def _decode(plan, packed, n_frames, fmt):
    torch = _torch()
    dt = torch.float32 if fmt == "ps" else torch.int16
    out = torch.zeros((n_frames, 3 * H * W), dtype=dt, device="cuda:0")
    plan.decode([(packed.data_ptr(), n_frames, 0, n_frames)], out.data_ptr())
    torch.cuda.synchronize()
    a = out.cpu().numpy().reshape(n_frames, 3, H, W)
    return a if fmt == "ps" else a.view(np.uint16)


def _oracle_plain(kind, fields, frames, fmt, **kw):
    return orc.decode_sequence(orc.pal_params(**kw), kind, fields, frames, orc.uniform_thresholds(kind),
                               0 if fmt == "ps" else 1, threads=min(16, os.cpu_count() or 1))


def _net():
    net = abi.MaskNet(max_tiles=1)
    net.load(WEIGHTS)
    return net


def _masks(plan, packed, n_fields, net):
    """The network's masks of every slab a decode of all frames reads."""
    torch = _torch()
    nf = n_fields // 2
    tps = plan.tiles_per_slab
    spec = torch.zeros(((nf + 2) * tps, abi.TILE_BINS, 2), dtype=torch.float32, device="cuda:0")
    plan.slab_spectra(packed.data_ptr(), n_fields, -2, nf + 2, spec.data_ptr())
    masks = torch.zeros((spec.shape[0], abi.TILE_VISITS), dtype=torch.float32, device="cuda:0")
    net.forward(spec.data_ptr(), spec.shape[0], masks.data_ptr())
    torch.cuda.synchronize()
    return masks.cpu().numpy().reshape(nf + 2, tps // N_TX3, N_TX3, abi.TILE_VISITS)


def _oracle_learned(fields, frames, m, fmt):
    p = orc.pal_params()
    return np.stack([orc.decode_frame_masked(p, fields, f, np.stack([m[f], m[f + 2]]), 0 if fmt == "ps" else 1)
                     for f in frames])


def config_2d():
    pic = synth.bars_gradient(0)
    fields = synth.encode_sequence([pic])
    res = {}
    for fmt in ("ps", "p16"):
        plan = _plan(abi.CHD_DEC_TRANSFORM_2D, fmt, 1)
        got = _decode(plan, _pack(plan, fields), 1, fmt)
        res[fmt] = _stats(got, _oracle_plain(7, fields, [0], fmt), fmt == "p16")
    return "2D, 1 frame (bars + gradient)", [0], res


def config_pal2d(fields):
    frames = [0, 1, 31, 62, 63]
    res = {}
    for fmt in ("ps", "p16"):
        plan = _plan(abi.CHD_DEC_PAL_2D, fmt, 64)
        got = _decode(plan, _pack(plan, fields), 64, fmt)
        res[fmt] = _stats(got[frames], _oracle_plain(6, fields, frames, fmt), fmt == "p16")
    return "PAL 2D (PALcolour filter, no transform), 64-frame sequence", frames, res


def config_3d(fields):
    frames = [0, 1, 2, 31, 32, 61, 62, 63]
    res = {}
    for fmt in ("ps", "p16"):
        plan = _plan(abi.CHD_DEC_TRANSFORM_3D, fmt, 64)
        got = _decode(plan, _pack(plan, fields), 64, fmt)
        res[fmt] = _stats(got[frames], _oracle_plain(8, fields, frames, fmt), fmt == "p16")
    return "3D, 64-frame sequence", frames, res


def config_ynr(fields):
    frames = [0, 1, 31, 62, 63]
    res = {}
    for fmt in ("ps", "p16"):
        plan = _plan(abi.CHD_DEC_TRANSFORM_3D, fmt, 64, luma_nr_level=2.0)
        got = _decode(plan, _pack(plan, fields), 64, fmt)
        res[fmt] = _stats(got[frames], _oracle_plain(8, fields, frames, fmt, luma_nr_level=2.0), fmt == "p16")
    return "3D with luma coring (luma_nr_level 2.0), 64-frame sequence", frames, res


def config_batch():
    import bench
    torch = _torch()
    seqs = bench.make_batch_workload()
    plan = _plan(abi.CHD_DEC_TRANSFORM_3D, "ps", 64)
    keep, runs = [], []
    for f in seqs:
        pk = _pack(plan, f)
        keep.append(pk)
        runs.append((pk.data_ptr(), f.shape[0] // 2, 0, f.shape[0] // 2))
    total = sum(r[3] for r in runs)
    out = torch.zeros((total, 3 * H * W), dtype=torch.float32, device="cuda:0")
    batch = abi.TpalBatch(plan, runs, out.data_ptr(), autotune=True)
    batch.run()
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(total, 3, H, W)
    rows, base, checked = [], 0, 0
    for f, r in zip(seqs, runs):
        n = r[3]
        idx = sorted({0, n // 2, n - 1})
        want = _oracle_plain(8, f, idx, "ps")
        rows.append(_stats(got[[base + i for i in idx]], want, False))
        base += n
        checked += len(idx)
    batch.close()
    return f"batch, 16 sequences ({total} frames, graph replay)", f"{checked} frames", {"ps": _merge(rows)}


def config_learned(fields):
    frames = [0, 1, 32, 62, 63]
    net = _net()
    res = {}
    for fmt in ("ps", "p16"):
        plan = _plan(abi.CHD_DEC_TRANSFORM_3D, fmt, 64)
        packed = _pack(plan, fields)
        m = _masks(plan, packed, fields.shape[0], net)
        plan.set_mask_net(net)
        got = _decode(plan, packed, 64, fmt)
        plan.set_mask_net(None)
        res[fmt] = _stats(got[frames], _oracle_learned(fields, frames, m, fmt), fmt == "p16")
    return "learned mask, 64-frame sequence (oracle runs the network's masks)", frames, res


def config_stress():
    from libchromadec_amd import stress
    n = 24
    fields, drops, _ = stress.make_capture(n_frames=n, chunk=12, workers=2)
    corrected, _, _ = stress.conceal(fields, drops)
    vp = dict(field_width=1135, field_height=313, colour_burst_end=138, active_video_end=1107,
              first_active_frame_line=44, last_active_frame_line=619, system="PAL")
    for f in range(n):
        a, b, _ = od.correct_frame(vp, fields[2 * f], fields[2 * f + 1], drops[2 * f], drops[2 * f + 1])
        assert np.array_equal(a, corrected[2 * f]) and np.array_equal(b, corrected[2 * f + 1]), f
    frames = [0, 5, 11, 12, 23]
    plan = _plan(abi.CHD_DEC_TRANSFORM_3D, "ps", n)
    packed = _pack(plan, corrected)
    res = {"ps": _stats(_decode(plan, packed, n, "ps")[frames], _oracle_plain(8, corrected, frames, "ps"), False)}
    net = _net()
    m = _masks(plan, packed, corrected.shape[0], net)
    plan.set_mask_net(net)
    got = _decode(plan, packed, n, "ps")
    plan.set_mask_net(None)
    res["ps learned"] = _stats(got[frames], _oracle_learned(corrected, frames, m, "ps"), False)
    return "stress, 24 frames (concealment bit-exact vs oracle)", frames, res


def main():
    import bench
    fields = bench.make_workload(64)[0]
    lines = ["| Config | Frames checked | Output | E'Y max | U/V max | U/V RMS |", "|---|---|---|---|---|---|"]
    for fn in (config_2d, lambda: config_3d(fields), config_batch, lambda: config_learned(fields), config_stress,
               lambda: config_pal2d(fields), lambda: config_ynr(fields)):
        name, frames, res = fn()
        for fmt, r in res.items():
            if r["u16"]:
                vals = f"{r['y_max']:.0f} LSB | {r['uv_max']:.0f} LSB | {r['uv_rms']:.3f} LSB"
            else:
                vals = f"{r['y_max']:.2e} | {r['uv_max']:.2e} | {r['uv_rms']:.2e}"
            lines.append(f"| {name} | {frames} | {fmt} | {vals} |")
        print(lines[-1], file=sys.stderr, flush=True)
    print("\n".join(lines))


if __name__ == "__main__":
    main()
