"""Stress workload: a long noisy capture with dropouts.

One continuous synthetic PAL capture (default 1000 frames): moving patches at
up to 16 px/frame over bars, 30 dB SNR, 1 degree RMS per-line subcarrier
phase jitter, 1-3 dropouts of 20-400 samples per field recorded in the
sidecar.  The dropouts are concealed by the library
(chd_video_correct_dropouts, the code chd_decode_frame runs with dropout
correction on), then every frame is decoded on the GPU with the plain
threshold and with the learned mask.  Reported: frames/s of each decode and
of the concealment, and the chroma PSNR of each decode against the ideal
separation (the same demodulator fed the true chroma).
"""
from __future__ import annotations

import os
import tempfile
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

from . import synth

TRUTH_BLACK = 32768


def _chunk(args):
    seed, f0, n, max_speed = args
    scene = synth.PatchScene(seed=seed, max_speed=max_speed)
    pics = [scene.frame(f) for f in range(f0, f0 + n)]
    opts = synth.EncodeOptions(snr_db=30.0, phase_jitter_deg=1.0, rng=np.random.default_rng((seed, f0)))
    fields, _, _, chroma = synth.encode_sequence(pics, opts, return_truth=True, first_field_no=2 * f0)
    return fields, np.clip(np.rint(chroma + TRUTH_BLACK), 0, 65535).astype(np.uint16)


def make_capture(n_frames: int = 1000, seed: int = 6, chunk: int = 20, workers: int = 0, max_speed: float = 16.0):
    """Composite fields with dropouts written into them, the dropout lists
    (per field: (startx, endx, field line 1-based)) and the true chroma."""
    workers = workers or min(16, os.cpu_count() or 1)
    jobs = [(seed, f0, min(chunk, n_frames - f0), max_speed) for f0 in range(0, n_frames, chunk)]
    with ProcessPoolExecutor(workers) as ex:
        parts = list(ex.map(_chunk, jobs))
    fields = np.concatenate([p[0] for p in parts])
    truth = np.concatenate([p[1] for p in parts])
    drops = synth.add_dropouts(fields, np.random.default_rng(seed + 1))
    return fields, drops, truth


def conceal(fields: np.ndarray, drops):
    """Library concealment of every frame (host), via a sidecar-only capture.
    Returns corrected fields, (corrected, failed, distance) totals, seconds."""
    from . import abi
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "stress.tbc")
        synth.write_tbc(path, fields[:0], None)   # empty sample file: only the sidecar is read
        # sidecar for all fields with their dropouts
        import json
        side = path + ".json"
        meta = json.load(open(side))
        meta["videoParameters"]["numberOfSequentialFields"] = int(fields.shape[0])
        meta["fields"] = []
        for i in range(fields.shape[0]):
            rec = {"seqNo": i + 1, "isFirstField": (i % 2) == 0, "syncConf": 100, "fieldPhaseID": 0}
            if drops[i]:
                rec["dropOuts"] = {"startx": [r[0] for r in drops[i]], "endx": [r[1] for r in drops[i]],
                                   "fieldLine": [r[2] for r in drops[i]]}
            meta["fields"].append(rec)
        json.dump(meta, open(side, "w"))
        v = abi.Video(path)
        out = fields.copy()
        tot = np.zeros(3, np.int64)
        t0 = time.perf_counter()
        for f in range(fields.shape[0] // 2):
            a, b = out[2 * f], out[2 * f + 1]
            tot += v.correct_dropouts(f, a, b)
        dt = time.perf_counter() - t0
        v.close()
    return out, tuple(int(x) for x in tot), dt


def _sep_psnr_device(out, ideal, margin_x=8, margin_y=4):
    import torch
    a = out[:, 1:, margin_y:576 - margin_y, margin_x:922 - margin_x].double()
    b = ideal[:, 1:, margin_y:576 - margin_y, margin_x:922 - margin_x].double()
    mse = torch.mean((a - b) ** 2).item()
    return 10.0 * np.log10(1.0 / mse)


def run(n_frames: int = 1000, weights: str | None = None, log=print):
    import torch
    from . import abi
    t0 = time.perf_counter()
    fields, drops, truth = make_capture(n_frames)
    log(f"stress capture: {n_frames} frames encoded in {time.perf_counter() - t0:.0f} s")
    corrected, dstats, conceal_s = conceal(fields, drops)
    dev = torch.device("cuda:0")
    plan = abi.TpalPlan(abi.pal_tpal_params(kind=abi.CHD_DEC_TRANSFORM_3D, max_batch_frames=64))
    pk = torch.empty((fields.shape[0], plan.packed_field_elems), dtype=torch.int16, device=dev)
    src = torch.from_numpy(corrected.view(np.int16)).to(dev)
    plan.pack(src.data_ptr(), fields.shape[0], pk.data_ptr())
    del src
    n = fields.shape[0] // 2
    seqs = [(pk.data_ptr(), n, 0, n)]
    res = {"frames": n, "snr_db": 30, "phase_jitter_deg_rms": 1.0, "max_motion_px_per_frame": 16,
           "dropouts": {"runs": int(sum(len(d) for d in drops)), "corrected": dstats[0], "failed": dstats[1],
                        "host_concealment_frames_per_s": round(n / conceal_s, 1)}}
    ideal = torch.empty((n, 3, 576, 922), dtype=torch.float32, device=dev)
    ch = torch.from_numpy(truth.astype(np.float32) - float(TRUTH_BLACK)).to(dev)
    plan.decode_separated(pk.data_ptr(), ch.data_ptr(), n, ideal.data_ptr())
    del ch
    out = torch.empty_like(ideal)
    net = None
    for name in ("plain", "learned"):
        if name == "learned":
            net = abi.MaskNet(max_tiles=1)
            net.load(weights or os.path.join(os.path.dirname(__file__), "data", "tpal3d_mask.bin"))
            plan.set_mask_net(net)
        plan.decode(seqs, out.data_ptr())
        torch.cuda.synchronize()
        t = time.perf_counter()
        plan.decode(seqs, out.data_ptr())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        res[name] = {"frames_per_s": round(n / dt, 1), "psnr_vs_ideal_separation_db": round(_sep_psnr_device(out, ideal), 3)}
        log(f"{name}: {res[name]}")
    plan.set_mask_net(None)
    return res
