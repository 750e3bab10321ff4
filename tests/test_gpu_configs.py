"""GPU: every benchmark configuration at its stated size, against the FP64 oracle.

  C1  one synthetic PAL frame of bars plus a moving gradient, Transform-2D
  C2  the 64-frame moving-patch sequence at 40 dB, Transform-3D, frames at
      both ends and inside; the same sequence with PAL 2D and with luma coring
  C3  16 sequences of seeded lengths 8-40 (391 frames, bars, patches and zone
      plates) decoded as one autotuned, graph-captured batch; the first,
      middle and last frame of every sequence
  C4  the learned mask: its frames against the oracle running the network's
      own masks through its FP64 soft-mask filter, and its gain over the
      plain threshold on the 200 held-out frames of seed 5
  C5  a stress capture (30 dB, 1 degree RMS phase jitter, motion up to
      16 px/frame, dropouts concealed by the library) decoded through
      chd_decode_frame with the learned mask and dropout concealment on

Tolerances: float planes within 1e-4 of full scale, yuv444p16 within 1 LSB.
Each test records its worst errors; with CHD_PARITY_TABLE set, the session
writes them as a markdown table (tests/conftest.py).  The oracle is the
checker; nothing it computes feeds the GPU path."""
import os

import numpy as np
import pytest

from libchromadec_amd import abi, synth
from oracle import dropout as od

pytestmark = pytest.mark.gpu

H, W = 576, 922
F32_TOL = 1e-4
U16_TOL = 1
N_TX3 = 117          # 3D tile columns of the PAL active line
WEIGHTS = os.path.join(os.path.dirname(abi.LIB_PATH), "..", "data", "tpal3d_mask.bin")
PAL_VP = dict(field_width=1135, field_height=313, colour_burst_end=138, active_video_end=1107,
              first_active_frame_line=44, last_active_frame_line=619, system="PAL")
THREADS = 16


def _torch():
    import torch
    return torch


def _stats(got, want):
    d = got.astype(np.float64) - want.astype(np.float64)
    if d.ndim == 3:
        d = d[None]
    return {"y_max": float(np.abs(d[:, 0]).max()), "uv_max": float(np.abs(d[:, 1:]).max()),
            "uv_rms": float(np.sqrt(np.mean(d[:, 1:] ** 2)))}


def _check(parity_rows, name, frames, fmt, got, want):
    s = _stats(got, want)
    parity_rows.append((name, frames, fmt, s))
    tol = F32_TOL if fmt == "ps" else U16_TOL
    assert s["y_max"] <= tol and s["uv_max"] <= tol, (name, fmt, s)


def _plan(kind, fmt, frames, **kw):
    pf = abi.CHD_PIXEL_YUV444PS if fmt == "ps" else abi.CHD_PIXEL_YUV444P16
    return abi.TpalPlan(abi.pal_tpal_params(kind=kind, pixel_format=pf, max_batch_frames=frames, **kw))


def _pack(plan, fields):
    torch = _torch()
    src = torch.from_numpy(np.ascontiguousarray(fields).view(np.int16)).to("cuda:0")
    packed = torch.empty((fields.shape[0], plan.packed_field_elems), dtype=torch.int16, device="cuda:0")
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


def _oracle(orc, kind, fields, frames, fmt, **kw):
    return orc.decode_sequence(orc.pal_params(**kw), kind, fields, frames, orc.uniform_thresholds(kind),
                               0 if fmt == "ps" else 1, threads=THREADS)


def _net():
    net = abi.MaskNet(max_tiles=1)
    net.load(WEIGHTS)
    return net


def _masks(plan, packed, n_fields, net):
    """The network's masks of every slab a decode of all frames reads,
    indexed by slab + 2 (slab -2 first)."""
    torch = _torch()
    nf = n_fields // 2
    tps = plan.tiles_per_slab
    spec = torch.zeros(((nf + 2) * tps, abi.TILE_BINS, 2), dtype=torch.float32, device="cuda:0")
    plan.slab_spectra(packed.data_ptr(), n_fields, -2, nf + 2, spec.data_ptr())
    masks = torch.zeros((spec.shape[0], abi.TILE_VISITS), dtype=torch.float32, device="cuda:0")
    net.forward(spec.data_ptr(), spec.shape[0], masks.data_ptr())
    torch.cuda.synchronize()
    return masks.cpu().numpy().reshape(nf + 2, tps // N_TX3, N_TX3, abi.TILE_VISITS)


def _oracle_learned(orc, fields, frames, m, fmt):
    p = orc.pal_params()
    # frame f's window anchors its slabs at fields 2f-4 and 2f: slabs f-2, f
    return np.stack([orc.decode_frame_masked(p, fields, f, np.stack([m[f], m[f + 2]]), 0 if fmt == "ps" else 1)
                     for f in frames])


@pytest.fixture(scope="module")
def c2_fields():
    import bench
    return bench.make_workload(64)[0]


def test_c1_transform2d_one_frame(orc, parity_rows):
    fields = synth.encode_sequence([synth.bars_gradient(0)])
    for fmt in ("ps", "p16"):
        plan = _plan(abi.CHD_DEC_TRANSFORM_2D, fmt, 1)
        got = _decode(plan, _pack(plan, fields), 1, fmt)
        _check(parity_rows, "C1 Transform-2D, 1 frame (bars + gradient)", "0", fmt, got,
               _oracle(orc, 7, fields, [0], fmt))
        plan.close()


def test_c2_transform3d_64_frames(orc, parity_rows, c2_fields):
    frames = [0, 1, 2, 31, 32, 61, 62, 63]
    for fmt in ("ps", "p16"):
        plan = _plan(abi.CHD_DEC_TRANSFORM_3D, fmt, 64)
        got = _decode(plan, _pack(plan, c2_fields), 64, fmt)
        _check(parity_rows, "C2 Transform-3D, 64-frame sequence", ", ".join(map(str, frames)), fmt, got[frames],
               _oracle(orc, 8, c2_fields, frames, fmt))
        plan.close()


@pytest.mark.parametrize("variant", ["pal2d", "coring"])
def test_c2_pal2d_and_luma_coring_64_frames(orc, parity_rows, c2_fields, variant):
    frames = [0, 1, 31, 62, 63]
    kind = abi.CHD_DEC_PAL_2D if variant == "pal2d" else abi.CHD_DEC_TRANSFORM_3D
    kw = {} if variant == "pal2d" else {"luma_nr_level": 2.0}
    name = ("C2 sequence, PAL 2D (PALcolour filter on the composite)" if variant == "pal2d"
            else "C2 sequence, Transform-3D with luma coring (level 2.0)")
    for fmt in ("ps", "p16"):
        plan = _plan(kind, fmt, 64, **kw)
        got = _decode(plan, _pack(plan, c2_fields), 64, fmt)
        _check(parity_rows, name, ", ".join(map(str, frames)), fmt, got[frames],
               _oracle(orc, kind, c2_fields, frames, fmt, **kw))
        plan.close()


def test_c3_batch_16_sequences(orc, parity_rows):
    """The full C3 workload as one autotuned, graph-captured batch, replayed
    twice with eager work between the replays; both replays are checked."""
    import bench
    torch = _torch()
    seqs = bench.make_batch_workload()
    assert len(seqs) == 16 and abs(sum(f.shape[0] // 2 for f in seqs) - 384) <= 8
    plan = _plan(abi.CHD_DEC_TRANSFORM_3D, "ps", 64)
    keep, runs = [], []
    for f in seqs:
        pk = _pack(plan, f)
        keep.append(pk)
        runs.append((pk.data_ptr(), f.shape[0] // 2, 0, f.shape[0] // 2))
    total = sum(r[3] for r in runs)
    out = torch.zeros((total, 3 * H * W), dtype=torch.float32, device="cuda:0")
    batch = abi.TpalBatch(plan, runs, out.data_ptr(), autotune=True)
    picks = []
    base = 0
    for f, r in zip(seqs, runs):
        n = r[3]
        idx = sorted({0, n // 2, n - 1})
        picks.append((f, idx, [base + i for i in idx]))
        base += n
    want = None
    for rep in range(2):
        out.fill_(float("nan"))                      # eager work between replays
        batch.run()
        torch.cuda.synchronize()
        got = out.cpu().numpy().reshape(total, 3, H, W)
        if want is None:
            want = [_oracle(orc, 8, f, idx, "ps") for f, idx, _ in picks]
        rows = [_stats(got[gi], w) for (_, _, gi), w in zip(picks, want)]
        s = {"y_max": max(r["y_max"] for r in rows), "uv_max": max(r["uv_max"] for r in rows),
             "uv_rms": float(np.sqrt(np.mean([r["uv_rms"] ** 2 for r in rows])))}
        assert s["y_max"] <= F32_TOL and s["uv_max"] <= F32_TOL, (rep, s, batch.config())
    n_checked = sum(len(idx) for _, idx, _ in picks)
    parity_rows.append((f"C3 batch, 16 sequences ({total} frames, graph replay)",
                        f"first, middle, last of each ({n_checked})", "ps", s))
    batch.close()
    plan.close()


def test_c4_learned_matches_oracle_with_the_same_masks(orc, parity_rows, c2_fields):
    frames = [0, 1, 32, 62, 63]
    net = _net()
    for fmt in ("ps", "p16"):
        plan = _plan(abi.CHD_DEC_TRANSFORM_3D, fmt, 64)
        packed = _pack(plan, c2_fields)
        m = _masks(plan, packed, c2_fields.shape[0], net)
        plan.set_mask_net(net)
        got = _decode(plan, packed, 64, fmt)
        plan.set_mask_net(None)
        _check(parity_rows, "C4 learned mask, C2 sequence (oracle runs the network's masks)",
               ", ".join(map(str, frames)), fmt, got[frames], _oracle_learned(orc, c2_fields, frames, m, fmt))
        plan.close()
    net.close()


def test_c4_learned_gain_on_200_held_out_frames():
    """The shipped weights against the plain 0.4 threshold on the 200
    held-out frames of seed 5 (never trained on): chroma PSNR against the
    ideal separation (the same demodulator fed the true modulated chroma)
    rises by at least 1 dB, and PSNR against the encoder's U/V does not
    fall."""
    from libchromadec_amd import mask_training as mt
    _, data = mt.make_dataset(200, seed=5, seq_len=20, workers=16, log=lambda m: None)
    assert sum(f.shape[0] // 2 for f, _, _, _ in data) == 200
    net = abi.MaskNet(max_tiles=8192)
    net.load(WEIGHTS)
    r = mt.evaluate(net, data)
    print("C4 held out (200 frames):", {k: round(v, 3) for k, v in r.items()})
    assert r["sep_learned"] - r["sep_plain"] >= 1.0, r
    assert r["uv_learned"] >= r["uv_plain"] - 0.05, r
    net.close()


@pytest.fixture(scope="module")
def stress_capture(tmp_path_factory):
    """12 frames in the C5 style: 30 dB, 1 degree RMS per-line phase jitter,
    motion up to 16 px/frame, 1-3 dropouts of 20-400 samples per field."""
    rng = np.random.default_rng(66)
    scene = synth.PatchScene(seed=66, max_speed=16.0)
    fields = synth.encode_sequence([scene.frame(i) for i in range(12)],
                                   synth.EncodeOptions(snr_db=30, phase_jitter_deg=1.0, rng=rng))
    drops = synth.add_dropouts(fields, rng)
    path = str(tmp_path_factory.mktemp("c5") / "stress.tbc")
    synth.write_tbc(path, fields, drops)
    return path, fields, drops


def test_c5_stress_learned_with_concealment(orc, parity_rows, stress_capture):
    """chd_decode_frame with transform_mask "learned" and dropout concealment:
    the library conceals every frame of the look-around window before the
    decode (c_api_decoder.cpp:212-321), so the oracle decodes its own
    concealment of the whole capture with the network's masks of those
    corrected slabs."""
    path, fields, drops = stress_capture
    corrected = fields.copy()
    for f in range(fields.shape[0] // 2):
        a, b, _ = od.correct_frame(PAL_VP, fields[2 * f], fields[2 * f + 1], drops[2 * f], drops[2 * f + 1])
        corrected[2 * f], corrected[2 * f + 1] = a, b
    assert not np.array_equal(corrected, fields)
    net = _net()
    plan = _plan(abi.CHD_DEC_TRANSFORM_3D, "ps", 12)
    m = _masks(plan, _pack(plan, corrected), corrected.shape[0], net)
    plan.close()
    net.close()
    v = abi.Video(path)
    dec = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_3D, output_format="yuv444ps",
                      transform_mask="learned").set_dropout(True).commit()
    frames = [0, 5, 11]
    got = np.stack([dec.decode(f) for f in frames])
    _check(parity_rows, "C5 stress, learned mask, dropouts concealed (chd_decode_frame)",
           ", ".join(map(str, frames)), "ps", got, _oracle_learned(orc, corrected, frames, m, "ps"))
    # the plain threshold leg of the same capture and window
    plain = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_3D, output_format="yuv444ps").set_dropout(True).commit()
    gotp = np.stack([plain.decode(f) for f in frames])
    _check(parity_rows, "C5 stress, plain threshold, dropouts concealed (chd_decode_frame)",
           ", ".join(map(str, frames)), "ps", gotp, _oracle(orc, 8, corrected, frames, "ps"))
    dec.close()
    plain.close()
    v.close()
