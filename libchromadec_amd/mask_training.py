"""Training and evaluation of the learned Transform-PAL 3D mask.

Training data are synthetic PAL sequences encoded from known component
pictures (libchromadec_amd.synth), so the true chroma of every tile is known:
the composite and the modulated chroma are both run through the library's
3D tile transform (chd_tpal_slab_spectra) and the network is fitted by its
own BF16 forward/backward kernels (chd_mask_net_gradient) and Adam
(chd_mask_net_adam_step).  torch supplies device memory and random batch
gathers only.

    python -m libchromadec_amd.mask_training --frames 2000 --steps 4000 \\
        --out libchromadec_amd/data/tpal3d_mask.bin

Evaluation decodes held-out frames with the plain threshold and with the
learned mask and reports two chroma PSNRs over the active picture less an
8-sample / 4-line margin:
  * separation PSNR: E'Cb/E'Cr against the ideal separation, i.e. the same
    PalColour demodulation applied to the true modulated chroma
    (chd_tpal_decode_separated).  The mask only changes the Y/C separation,
    and this isolates it;
  * PSNR against the band-limited U/V the encoder modulated, which also
    carries the demodulator's own chroma bandwidth limit (the same for every
    mask).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

from . import synth

N_PARAMS = 32033
TRUTH_BLACK = 32768       # truth chroma is stored as uint16 around this level
ACTIVE = dict(first=44, last=619, avs=185, ave=1107)


# ---------------------------------------------------------------------------
# Synthetic sequences
# ---------------------------------------------------------------------------

def sequence_spec(seed: int):
    """Content and channel parameters of one training/evaluation sequence."""
    rng = np.random.default_rng(seed)
    return dict(seed=seed, zone_plate=bool(rng.uniform() < 0.35), max_speed=float(rng.choice([2.0, 4.0, 8.0, 16.0])),
                snr_db=float(rng.choice([30.0, 35.0, 40.0, 45.0])),
                jitter=float(rng.choice([0.0, 0.0, 1.0])), bars=bool(rng.uniform() < 0.15))


def encode_sequence(spec: dict, n_frames: int, with_uv: bool = True):
    """Returns composite fields [2n, 313, 1135] uint16, truth chroma as
    uint16 around TRUTH_BLACK, and the band-limited U, V (float32; None
    unless with_uv)."""
    if spec.get("bars"):
        pics = [synth.bars_gradient(f, speed=min(spec["max_speed"], 4.0)) for f in range(n_frames)]
    else:
        scene = synth.PatchScene(seed=spec["seed"], max_speed=spec["max_speed"], zone_plate=spec["zone_plate"])
        pics = [scene.frame(f) for f in range(n_frames)]
    opts = synth.EncodeOptions(snr_db=spec["snr_db"], phase_jitter_deg=spec["jitter"],
                               rng=np.random.default_rng(spec["seed"] + 7919))
    fields, u, v, chroma = synth.encode_sequence(pics, opts, return_truth=True)
    truth = np.clip(np.rint(chroma + TRUTH_BLACK), 0, 65535).astype(np.uint16)
    if not with_uv:
        return fields, truth, None, None
    return fields, truth, u.astype(np.float32), v.astype(np.float32)


def _encode_job(args):
    spec, n, with_uv = args
    return encode_sequence(spec, n, with_uv)


def make_dataset(n_frames: int, seed: int, seq_len: int = 20, workers: int = 0, log=print, with_uv: bool = True):
    """Sequences of `seq_len` frames totalling n_frames, encoded in parallel."""
    n_seq = max(1, (n_frames + seq_len - 1) // seq_len)
    specs = [sequence_spec(seed * 100003 + i) for i in range(n_seq)]
    lens = [min(seq_len, n_frames - i * seq_len) for i in range(n_seq)]
    workers = workers or min(16, os.cpu_count() or 1)
    out = []
    t0 = time.time()
    with ProcessPoolExecutor(workers) as ex:
        for i, r in enumerate(ex.map(_encode_job, [(sp, n, with_uv) for sp, n in zip(specs, lens)])):
            out.append(r)
            if (i + 1) % max(1, n_seq // 10) == 0 or i + 1 == n_seq:
                log(f"encoded {i + 1}/{n_seq} sequences ({sum(lens[:i + 1])} frames) in {time.time() - t0:.0f} s")
    return specs, out


# ---------------------------------------------------------------------------
# Training
# ---------------------------------------------------------------------------

def init_params(seed: int) -> np.ndarray:
    """He initialisation; the output bias starts the mask near the threshold's
    typical keep rate."""
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, math.sqrt(2.0 / (27 * 4)), 27 * 4 * 32), np.zeros(32),
             rng.normal(0, math.sqrt(2.0 / (27 * 32)), 27 * 32 * 32), np.zeros(32),
             rng.normal(0, math.sqrt(1.0 / (27 * 32)), 27 * 32), np.zeros(1)]
    return np.concatenate(parts).astype(np.float32)


def train(data, steps: int = 4000, batch: int = 256, lr: float = 2e-3, slabs_per_round: int = 8,
          steps_per_round: int = 16, seed: int = 4, log=print, device: int = 0):
    import torch
    from . import abi

    dev = torch.device(f"cuda:{device}")
    plan = abi.TpalPlan(abi.pal_tpal_params(kind=abi.CHD_DEC_TRANSFORM_3D, max_batch_frames=slabs_per_round,
                                            device=device))
    tplan = abi.TpalPlan(abi.pal_tpal_params(kind=abi.CHD_DEC_TRANSFORM_3D, max_batch_frames=slabs_per_round,
                                             device=device, black=TRUTH_BLACK))
    pe = plan.packed_field_elems
    seqs = []
    for fields, truth, _, _ in data:
        pk = torch.empty((fields.shape[0], pe), dtype=torch.int16, device=dev)
        tk = torch.empty_like(pk)
        src = torch.from_numpy(fields.view(np.int16)).to(dev)
        plan.pack(src.data_ptr(), fields.shape[0], pk.data_ptr())
        src = torch.from_numpy(truth.view(np.int16)).to(dev)
        tplan.pack(src.data_ptr(), truth.shape[0], tk.data_ptr())
        seqs.append((pk, tk, fields.shape[0]))
    torch.cuda.synchronize()
    tps = plan.tiles_per_slab
    spec = torch.empty((slabs_per_round * tps, abi.TILE_BINS, 2), dtype=torch.float32, device=dev)
    tspec = torch.empty_like(spec)
    net = abi.MaskNet(max_tiles=batch, device=device)
    net.set_params(init_params(seed))
    grad = torch.zeros(N_PARAMS, dtype=torch.float32, device=dev)
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    scale = None
    hist = []
    t0 = time.time()
    step = 0
    run = 0.0
    while step < steps:
        # a round: spectra of a few random slabs, then batches drawn from them
        for j in range(slabs_per_round):
            pk, tk, nf = seqs[int(rng.integers(0, len(seqs)))]
            k = int(rng.integers(-2, nf // 2))
            plan.slab_spectra(pk.data_ptr(), nf, k, 1, spec[j * tps].data_ptr())
            tplan.slab_spectra(tk.data_ptr(), nf, k, 1, tspec[j * tps].data_ptr())
        for _ in range(steps_per_round):
            if step >= steps:
                break
            idx = torch.randint(0, slabs_per_round * tps, (batch,), device=dev, generator=gen)
            s, ts = spec[idx].contiguous(), tspec[idx].contiguous()
            if scale is None:
                # fixed loss scale: 1 / (sum of |F|^2 over the first batch)
                a = s.double()
                scale = float(1.0 / (a * a).sum().item())
            loss = net.gradient(s.data_ptr(), ts.data_ptr(), batch, scale, grad.data_ptr(),
                                want_loss=(step % 50 == 0))
            cur_lr = lr * (0.05 + 0.95 * 0.5 * (1.0 + math.cos(math.pi * step / steps)))
            net.adam_step(grad.data_ptr(), cur_lr)
            if step % 50 == 0:
                run = loss if step == 0 else 0.9 * run + 0.1 * loss
                hist.append((step, loss))
            if step % 250 == 0:
                log(f"step {step}/{steps} loss {loss:.5f} (avg {run:.5f}) lr {cur_lr:.2e} {time.time() - t0:.0f} s")
            step += 1
    torch.cuda.synchronize()
    log(f"trained {steps} steps of {batch} tiles in {time.time() - t0:.0f} s")
    return net, hist


# ---------------------------------------------------------------------------
# Evaluation
# ---------------------------------------------------------------------------

def separation_psnr(frames_out: np.ndarray, ideal: np.ndarray, margin_x: int = 8, margin_y: int = 4) -> float:
    """PSNR (dB) of E'Cb/E'Cr against the ideal-separation decode."""
    a = frames_out[:, 1:, margin_y:frames_out.shape[2] - margin_y, margin_x:frames_out.shape[3] - margin_x]
    b = ideal[:, 1:, margin_y:ideal.shape[2] - margin_y, margin_x:ideal.shape[3] - margin_x]
    d = a.astype(np.float64) - b
    return 10.0 * math.log10(1.0 / float(np.mean(d * d)))


def chroma_psnr(frames_out: np.ndarray, u: np.ndarray, v: np.ndarray, frame_ids, ecb: float, ecr: float,
                margin_x: int = 8, margin_y: int = 4) -> float:
    """PSNR (dB) of decoded E'Cb/E'Cr (yuv444ps frames, full active picture
    922 x 576) against ecb*U and ecr*V of the encoder's truth fields."""
    first, last, avs, ave = ACTIVE["first"], ACTIVE["last"], ACTIVE["avs"], ACTIVE["ave"]
    lines = np.arange(first, last + 1)
    sel = lines[margin_y:len(lines) - margin_y]
    cols = np.arange(avs + margin_x, ave - margin_x)
    se, n = 0.0, 0
    for out, f in zip(frames_out, frame_ids):
        fi = 2 * f + (sel % 2)
        fl = sel // 2
        tu = u[fi[:, None], fl[:, None], cols[None, :]] * ecb
        tv = v[fi[:, None], fl[:, None], cols[None, :]] * ecr
        du = out[1][sel - first][:, cols - avs] - tu
        dv = out[2][sel - first][:, cols - avs] - tv
        se += float((du.astype(np.float64) ** 2).sum() + (dv.astype(np.float64) ** 2).sum())
        n += du.size + dv.size
    return 10.0 * math.log10(1.0 / (se / n))


def decode_frames(plan, fields: np.ndarray, device: int = 0) -> np.ndarray:
    """All frames of a sequence through chd_tpal_decode (yuv444ps)."""
    import torch
    dev = torch.device(f"cuda:{device}")
    n = fields.shape[0] // 2
    src = torch.from_numpy(fields.view(np.int16)).to(dev)
    pk = torch.empty((fields.shape[0], plan.packed_field_elems), dtype=torch.int16, device=dev)
    plan.pack(src.data_ptr(), fields.shape[0], pk.data_ptr())
    out = torch.empty((n, plan.frame_bytes // 4), dtype=torch.float32, device=dev)
    plan.decode([(pk.data_ptr(), n, 0, n)], out.data_ptr())
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(n, 3, 576, 922)


def ideal_frames(plan, fields: np.ndarray, truth: np.ndarray, device: int = 0) -> np.ndarray:
    """Frames decoded from the true chroma (uint16 around TRUTH_BLACK)."""
    import torch
    dev = torch.device(f"cuda:{device}")
    n = fields.shape[0] // 2
    src = torch.from_numpy(fields.view(np.int16)).to(dev)
    pk = torch.empty((fields.shape[0], plan.packed_field_elems), dtype=torch.int16, device=dev)
    plan.pack(src.data_ptr(), fields.shape[0], pk.data_ptr())
    ch = torch.from_numpy(truth.astype(np.float32) - float(TRUTH_BLACK)).to(dev)
    out = torch.empty((n, plan.frame_bytes // 4), dtype=torch.float32, device=dev)
    plan.decode_separated(pk.data_ptr(), ch.data_ptr(), n, out.data_ptr())
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(n, 3, 576, 922)


def evaluate(net, data, device: int = 0, threshold: float = 0.4):
    """Chroma PSNRs over every frame of every sequence: dict with
    sep_plain / sep_learned (against the ideal separation) and uv_plain /
    uv_learned (against the encoder's U/V)."""
    from . import abi
    ecb, ecr = abi.color_scales()
    plain = abi.TpalPlan(abi.pal_tpal_params(kind=abi.CHD_DEC_TRANSFORM_3D, threshold=threshold, device=device,
                                             max_batch_frames=16))
    learned = abi.TpalPlan(abi.pal_tpal_params(kind=abi.CHD_DEC_TRANSFORM_3D, device=device, max_batch_frames=16))
    learned.set_mask_net(net)
    acc = {k: [0.0, 0] for k in ("sep_plain", "sep_learned", "uv_plain", "uv_learned")}

    def add(key, psnr, n):
        acc[key][0] += n * 10.0 ** (-psnr / 10.0)
        acc[key][1] += n

    for fields, truth, u, v in data:
        n = fields.shape[0] // 2
        ideal = ideal_frames(plain, fields, truth, device)
        for name, plan in (("plain", plain), ("learned", learned)):
            out = decode_frames(plan, fields, device)
            add("sep_" + name, separation_psnr(out, ideal), n)
            add("uv_" + name, chroma_psnr(out, u, v, range(n), ecb, ecr), n)
    learned.set_mask_net(None)
    return {k: 10.0 * math.log10(a[1] / a[0]) for k, a in acc.items()}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--frames", type=int, default=2000)
    ap.add_argument("--eval-frames", type=int, default=200)
    ap.add_argument("--seq-len", type=int, default=20)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--eval-seed", type=int, default=5)
    ap.add_argument("--out", required=True)
    ap.add_argument("--summary", default=None)
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    _, data = make_dataset(args.frames, args.seed, args.seq_len, log=log, with_uv=False)
    net, hist = train(data, steps=args.steps, batch=args.batch, lr=args.lr, seed=args.seed, log=log)
    del data
    net.save(args.out)
    log(f"weights written to {args.out}")
    summary = dict(frames=args.frames, steps=args.steps, batch=args.batch, lr=args.lr, seed=args.seed,
                   loss_history=hist)
    if args.eval_frames > 0:
        _, evdata = make_dataset(args.eval_frames, args.eval_seed, args.seq_len, log=log)
        r = evaluate(net, evdata)
        log(f"held-out separation PSNR: plain {r['sep_plain']:.2f} dB, learned {r['sep_learned']:.2f} dB "
            f"({r['sep_learned'] - r['sep_plain']:+.2f} dB); against the encoder's U/V: plain "
            f"{r['uv_plain']:.2f} dB, learned {r['uv_learned']:.2f} dB")
        summary.update(eval_frames=args.eval_frames, eval_seed=args.eval_seed, **{"psnr_" + k: v for k, v in r.items()})
    if args.summary:
        with open(args.summary, "w") as fh:
            json.dump(summary, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
