"""Transform-PAL 3D throughput on one MI355X.

Workload (the GPU configuration the headline metric is quoted on): a 64-frame
synthetic PAL sequence, 4fSC composite stored as 1135x313 uint16 fields,
seeded moving Gaussian-blurred colour patches over a static bar background
with 40 dB SNR noise; CHD_DEC_TRANSFORM_3D, threshold 0.4, yuv444ps output.
One step decodes all 64 frames with the device API (chd_tpal_decode), fields
already resident in HBM, output planes written to HBM.

Prints one JSON line.  `roofline` is for one decode, the step's two kernels
(the 3D tile transform and the demodulator, each timed with HIP events on
the decode stream over the K timed steps), against the algorithmic bytes of
the step (composite in, planes out); `traffic` is their measured HBM bytes
(profiles/tpal3d_tiles_c2_traffic.json).  `cpu_baseline` times the FP64
oracle (a port of the reference algorithm) on the host.

Beside the headline (plain threshold) the line carries:
  * `learned`: the same 64 frames decoded with the packaged learned mask
    (frames/s), and the chroma PSNR of both decodes against the ideal
    separation (the same demodulator fed the true chroma the encoder
    modulated) and against the encoder's U/V;
  * `batched`: 16 sequences of seeded lengths 8-40 frames (about 384 frames;
    bars, moving patches and luma zone plates up to 5 MHz) decoded as one
    fixed batch, one persistent tile launch replayed from a HIP graph, with
    the strip length and grid picked by the library's autotuner;
  * `transform_2d`: the same 64 frames decoded with CHD_DEC_TRANSFORM_2D.
These three legs run before the headline's warm-up and timed steps, so that
the headline is timed at the GPU's sustained clocks; `abi_end_to_end` and the
CPU baseline run after it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_FRAMES = 64
ALGO_IN = 2 * 1135 * 313 * 2            # composite bytes per frame
ALGO_OUT_PS = 3 * 922 * 576 * 4         # yuv444ps planes per frame
NOMINAL_GFLOP_PER_FRAME = 1.396         # 4329 tiles x 261.6 kFLOP + demod
NOMINAL_GFLOP_PER_FRAME_2D = 0.372      # 4366 field tiles x 24.9 kFLOP + demod
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def make_workload(n_frames: int, seed: int = 2):
    """Composite fields and the true modulated chroma (uint16 around 32768)."""
    from libchromadec_amd import synth
    cache = os.path.join("/tmp", f"chd_bench_c2t_{n_frames}_{seed}.npz")
    if os.path.exists(cache):
        z = np.load(cache)
        return z["fields"], z["truth"], z["u"], z["v"]
    scene = synth.PatchScene(seed=seed)
    fields, u, v, chroma = synth.encode_sequence([scene.frame(f) for f in range(n_frames)],
                                                 synth.EncodeOptions(snr_db=40, rng=np.random.default_rng(seed)),
                                                 return_truth=True)
    truth = np.clip(np.rint(chroma + 32768.0), 0, 65535).astype(np.uint16)
    u, v = u.astype(np.float32), v.astype(np.float32)
    try:
        np.savez(cache, fields=fields, truth=truth, u=u, v=v)
    except OSError:
        pass
    return fields, truth, u, v


def _c3_job(args):
    from libchromadec_amd import synth
    i, n, seed = args
    kind = i % 3
    if kind == 0:
        pics = [synth.bars_gradient(f) for f in range(n)]
    else:
        scene = synth.PatchScene(seed=seed * 1000 + i, zone_plate=(kind == 2))
        pics = [scene.frame(f) for f in range(n)]
    return synth.encode_sequence(pics, synth.EncodeOptions(snr_db=40, rng=np.random.default_rng(seed * 1000 + i)))


def make_batch_workload(seed: int = 3):
    """16 sequences, lengths U{8..40} redrawn until the total is 384 +- 8."""
    from concurrent.futures import ProcessPoolExecutor
    cache = os.path.join("/tmp", f"chd_bench_c3_{seed}.npz")
    if os.path.exists(cache):
        z = np.load(cache)
        return [z[f"s{i}"] for i in range(16)]
    rng = np.random.default_rng(seed)
    while True:
        lens = rng.integers(8, 41, 16)
        if abs(int(lens.sum()) - 384) <= 8:
            break
    with ProcessPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        seqs = list(ex.map(_c3_job, [(i, int(n), seed) for i, n in enumerate(lens)]))
    try:
        np.savez(cache, **{f"s{i}": q for i, q in enumerate(seqs)})
    except OSError:
        pass
    return seqs


def host_cpu():
    """The host the CPU baseline runs on: logical CPUs, the CPUs this process
    may use (its affinity mask, which the box sets to its share), the OpenMP
    thread limit and the CPU model name (/proc/cpuinfo, as lscpu reports it)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return {"nproc": os.cpu_count(), "affinity_cpus": avail, "omp_num_threads": omp or None, "model": model}


def _rates(n_frames, seconds):
    r = [n_frames / t for t in seconds]
    return float(np.mean(r)), float(np.median(r))


def cpu_baseline(fields: np.ndarray, n_sample: int, gpu_frames=None, c1_fields=None):
    """The FP64 oracle (a restatement of the reference algorithm) timed on the
    host, in the pattern of examples/nn_benchmark.c: warm-up kept apart, then
    mean and median frames/s.  Threads: every CPU this process may use (the
    reference's thread_count = 0 default is hardware_concurrency), capped by
    OMP_NUM_THREADS where the host sets it.  When the GPU's frames of the same
    indices are given, also reports their error against the oracle's (the
    oracle as the checker, not the thing measured)."""
    from oracle import oracle as orc
    host = host_cpu()
    threads = host["affinity_cpus"]
    if host["omp_num_threads"]:
        threads = min(threads, host["omp_num_threads"])
    p = orc.pal_params()
    thr3 = orc.uniform_thresholds(8)
    frames = list(range(n_sample))
    orc.decode_sequence(p, 8, fields, frames[:min(2, n_sample)], thr3, 0, threads=threads)   # warm-up
    walls, want = [], None
    for _ in range(3):
        t0 = time.perf_counter()
        want = orc.decode_sequence(p, 8, fields, frames, thr3, 0, threads=threads)
        walls.append(time.perf_counter() - t0)
    mean, median = _rates(n_sample, walls)
    # one core: the reference's thread_count = 1 chd_decode_frame loop
    orc.decode_sequence(p, 8, fields, [0], thr3, 0, threads=1)                              # warm-up
    per = []
    for f in range(1, 5):
        t1 = time.perf_counter()
        orc.decode_sequence(p, 8, fields, [f], thr3, 0, threads=1)
        per.append(time.perf_counter() - t1)
    mean1, median1 = _rates(1, per)
    res = {"value": median, "unit": "frames/s", "cores": threads, "kind": "port",
           "sample": f"{n_sample} frames of the same 64-frame sequence, FP64 oracle "
                     f"(reference-equivalent TransformPal3D+PalColour+OutputWriter, radix-2 FFT; FFTW is absent "
                     f"here), one frame per OpenMP thread on {threads} threads; median of 3 timed runs "
                     f"after a 2-frame warm-up, {sum(walls):.1f} s wall",
           "mean": mean, "median": median, "host": host,
           "one_core": {"value": median1, "mean": mean1, "median": median1, "unit": "frames/s", "cores": 1,
                        "sample": "frames 1-4 one at a time after a warm-up frame, one thread"}}
    if c1_fields is not None:
        # C1: Transform-2D on one bars + gradient frame
        thr2 = orc.uniform_thresholds(7)
        orc.decode_sequence(p, 7, c1_fields, [0], thr2, 0, threads=1)                       # warm-up
        per2 = []
        for _ in range(8):
            t2 = time.perf_counter()
            orc.decode_sequence(p, 7, c1_fields, [0], thr2, 0, threads=1)
            per2.append(time.perf_counter() - t2)
        m2, md2 = _rates(1, per2)
        reps = [0] * threads
        t3 = time.perf_counter()
        orc.decode_sequence(p, 7, c1_fields, reps, thr2, 0, threads=threads)
        dt3 = time.perf_counter() - t3
        res["transform_2d_c1"] = {"one_core": {"mean": m2, "median": md2, "unit": "frames/s",
                                               "sample": "8 decodes of the one C1 frame after a warm-up"},
                                  "threads": {"value": threads / dt3, "unit": "frames/s", "cores": threads,
                                              "sample": f"the C1 frame decoded {threads} times, one per thread"}}
    parity = None
    if gpu_frames is not None:
        d = gpu_frames.astype(np.float64) - want.astype(np.float64)
        parity = {"frames": n_sample, "y_max_abs": float(np.abs(d[:, 0]).max()),
                  "uv_max_abs": float(np.abs(d[:, 1:]).max()),
                  "uv_rms": float(np.sqrt(np.mean(d[:, 1:] ** 2))), "tolerance": 1e-4}
    return res, parity


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=16, help="frames for the CPU baseline (0 = skip)")
    ap.add_argument("--no-parity", action="store_true",
                    help="skip comparing the GPU frames with the oracle's in the CPU-baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="skip the learned-mask and batched legs")
    ap.add_argument("--extras", default="learned,batched,2d,abi",
                    help="comma-separated legs to run beside the headline: learned, batched, abi")
    ap.add_argument("--stress", type=int, default=0,
                    help="instead of the benchmark, run the dropout stress workload over this many frames "
                         "(libchromadec_amd/stress.py) and print its JSON line")
    ap.add_argument("--tile-phases", action="store_true",
                    help="instead of the benchmark, decode the 64-frame workload with the library built with "
                         "per-phase clock stamps in the 3D tile kernel (lib/phases, make -C libchromadec_amd "
                         "phases) and print the cycles per tile of each phase")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one plain elementwise copy of this many MiB (a known-size kernel that "
                         "scales hardware FETCH_SIZE/WRITE_SIZE counters to bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: only --gpus 1 is measured; the decode would shard by contiguous frame runs "
              "(plus a 2-frame/4-frame halo) across GPUs with no exchange between them", file=sys.stderr)
        return 2

    import torch
    from libchromadec_amd import abi

    if args.tile_phases:
        abi.LIB_PATH = os.path.join(ROOT, "libchromadec_amd", "lib", "phases", "libchromadec.so")
    if args.stress > 0:
        from libchromadec_amd import stress
        print(json.dumps({"stress": stress.run(args.stress, log=lambda m: print(m, file=sys.stderr, flush=True))}))
        return 0

    fields, truth, tu, tv = make_workload(N_FRAMES)
    dev = torch.device("cuda:0")
    plan = abi.TpalPlan(abi.pal_tpal_params(kind=abi.CHD_DEC_TRANSFORM_3D, pixel_format=abi.CHD_PIXEL_YUV444PS,
                                            max_batch_frames=N_FRAMES))
    src = torch.from_numpy(fields.view(np.int16)).to(dev)
    packed = torch.empty((fields.shape[0], plan.packed_field_elems), dtype=torch.int16, device=dev)
    plan.pack(src.data_ptr(), fields.shape[0], packed.data_ptr())
    del src
    out = torch.empty((N_FRAMES, plan.frame_bytes // 4), dtype=torch.float32, device=dev)
    seqs = [(packed.data_ptr(), N_FRAMES, 0, N_FRAMES)]
    stream = torch.cuda.current_stream().cuda_stream
    if args.tile_phases:
        print(json.dumps(tile_phases(plan, seqs, out, stream, args.steps)))
        return 0
    if args.copy_mib > 0:
        n = args.copy_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        torch.neg(a, out=b)
        del a, b
    torch.cuda.synchronize()

    # The other GPU legs run first.  From idle the GPU's clocks take about 25
    # decodes (30 ms of load) to reach their sustained level, so timed right
    # after a few warm-up steps the headline would measure that transient.
    extras = set() if args.no_extras else {e for e in args.extras.split(",") if e}
    # (the learned leg's PSNR evaluation runs on the host, so it goes first)
    legs = {}
    if "learned" in extras:
        legs["learned"] = learned_leg(plan, packed, fields, truth, tu, tv, out, seqs, stream, max(3, args.steps // 4))
    if "batched" in extras:
        legs["batched"] = batched_leg(args.steps)
    if "2d" in extras:
        legs["transform_2d"] = transform2d_leg(fields, args.steps)

    for _ in range(args.warmup):
        plan.decode(seqs, out.data_ptr(), stream)
    torch.cuda.synchronize()
    # kernel durations over the timed steps: HIP events around each launch on
    # the decode stream, recorded without waiting and summed afterwards
    plan.set_timing(2)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.decode(seqs, out.data_ptr(), stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    tile_ms, demod_ms, n_tile, _ = plan.last_timing()
    plan.set_timing(0)
    ms_per_step = 1e3 * elapsed / args.steps
    fps = N_FRAMES * args.steps / elapsed
    tile_avg = tile_ms / n_tile
    demod_avg = demod_ms / n_tile
    frames_per_launch = N_FRAMES
    algo_bytes = (ALGO_IN + ALGO_OUT_PS) * frames_per_launch
    # the roofline's unit is one decode: the tile launch and the demodulator
    # launch of a step (the algorithmic bytes, composite in and planes out,
    # belong to the pair, not to the tile kernel, which writes no plane)
    step_kernel_ms = tile_avg + demod_avg
    achieved = algo_bytes / (step_kernel_ms * 1e-3) / 1e9
    achieved_tflops = NOMINAL_GFLOP_PER_FRAME * frames_per_launch / (step_kernel_ms * 1e-3) / 1e3

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "tpal3d_tiles_c2_traffic.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get("bytes_per_decode")
        except (OSError, ValueError):
            traffic = None

    res = {
        "metric": "PAL frames/s, Transform-3D plain vs learned mask; chroma PSNR vs truth",
        "value": round(fps, 1),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "64-frame synthetic PAL sequence (moving patches over bars, 40 dB SNR), "
                               "1135x313 u16 fields, CHD_DEC_TRANSFORM_3D threshold 0.4, yuv444ps, "
                               "fields resident in HBM",
                   "frames_per_step": N_FRAMES, "tile": "8x32x16", "threshold": 0.4,
                   "value_is": "plain-threshold decode; the learned-mask decode and both PSNRs are under 'learned'"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "tpal3d_tiles + pal_demod_quad4 (one decode: the step's two launches)",
                     "avg_launch_ms": round(step_kernel_ms, 4),
                     "algorithmic_bytes_per_launch": algo_bytes,
                     "tile_kernel_avg_launch_ms": round(tile_avg, 4),
                     "demod_kernel_avg_launch_ms": round(demod_avg, 4),
                     "frac_tile_kernel_alone": round(algo_bytes / (tile_avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "valu": {"nominal_gflop_per_frame": NOMINAL_GFLOP_PER_FRAME,
                              "achieved_tflops_nominal": round(achieved_tflops, 2),
                              "peak_tflops": FP32_PEAK_TFLOPS,
                              "frac": round(achieved_tflops / FP32_PEAK_TFLOPS, 4)}},
    }
    if "learned" in legs:
        res["learned"] = legs["learned"]
        res["learned"]["per_frame_time_vs_plain"] = round(fps / res["learned"]["value"], 2)
    for k in ("batched", "transform_2d"):
        if k in legs:
            res[k] = legs[k]
    if "abi" in extras:
        res["abi_end_to_end"] = abi_leg(fields)
    if args.cpu_sample > 0:
        gpu_frames = None
        if not args.no_parity:
            plan.decode(seqs, out.data_ptr(), stream)
            torch.cuda.synchronize()
            gpu_frames = out[:args.cpu_sample].view(-1, 3, 576, 922).cpu().numpy()
        from libchromadec_amd import synth
        c1 = synth.encode_sequence([synth.bars_gradient(0)])
        res["cpu_baseline"], res["parity_vs_fp64_oracle"] = cpu_baseline(fields, args.cpu_sample, gpu_frames, c1)
    print(json.dumps(res))
    return 0


TILE_PHASES = ["forward x (unpack, window, x bins 2..6, LDS rows)", "vertical 16-point FFTs", "temporal 8-point DFTs",
               "carrier-mirror mask steps", "keep words and flag ballot", "FP64 re-decision",
               "masked inverse z", "inverse y", "inverse x and overlap-add", "tile setup"]


def tile_phases(plan, seqs, out, stream, steps):
    """Cycles per tile of each phase of the plain 3D tile kernel, from the
    clock-stamped build: wave clocks, so a phase's figure includes the time
    the wave waits (on memory, LDS or the other waves of its SIMD)."""
    import ctypes
    import torch
    from libchromadec_amd import abi
    read = ctypes.CDLL(abi.LIB_PATH).chd_tile_phases_read
    read.argtypes = [ctypes.POINTER(ctypes.c_ulonglong), ctypes.c_int]
    buf = (ctypes.c_ulonglong * 32)()
    for _ in range(3):
        plan.decode(seqs, out.data_ptr(), stream)
    torch.cuda.synchronize()
    if read(buf, 1) != 0:
        raise RuntimeError("chd_tile_phases_read failed")
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.decode(seqs, out.data_ptr(), stream)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    if read(buf, 0) != 0:
        raise RuntimeError("chd_tile_phases_read failed")
    v = list(buf)
    tiles = max(1, v[18])
    total = sum(v[:10])
    return {"tile_phases": {
        "workload": "64-frame synthetic PAL sequence, CHD_DEC_TRANSFORM_3D threshold 0.4 (the benchmark's)",
        "ms_per_step_stamped": round(ms, 4), "tiles_per_step": v[18] // steps,
        "tiles_with_redecision_per_step": v[16] // steps, "flagged_words_per_step": v[17] // steps,
        "cycles_per_tile": round(total / tiles, 1),
        "phases": [{"phase": TILE_PHASES[i], "cycles_per_tile": round(v[i] / tiles, 1),
                    "share": round(v[i] / total, 4)} for i in range(10)]}}


# pinned device-to-host copy rate measured on MI355X (one 6.4 MB copy at a
# time, hipMemcpyAsync)
LINK_GBPS = 52.5


def abi_leg(fields):
    """End to end through the public C ABI, from C: libchromadec_amd/lib/
    chd_abi_bench (abi_bench.c) opens the 64 frames as a .tbc capture and
    times chd_decode_frames_async over 6 passes (384 frames, a C callback
    frees each frame) seven times after a warm-up, and a chd_decode_frame
    loop; host planes out (PCIe included), with 2, 4 and 8 decode threads.
    Run as a child process."""
    import shutil
    import subprocess
    import tempfile
    from libchromadec_amd import synth
    exe = os.path.join(ROOT, "libchromadec_amd", "lib", "chd_abi_bench")
    if not os.path.exists(exe):
        return {"error": f"{exe} is not built"}
    d = tempfile.mkdtemp(prefix="chd_bench_")
    path = os.path.join(d, "c2.tbc")
    synth.write_tbc(path, fields)
    by = {}
    try:
        for t in (2, 4, 8):
            r = subprocess.run([exe, path, "8", "yuv444ps", str(t), "6", "7"], capture_output=True, text=True,
                               timeout=300)
            if r.returncode != 0:
                by[t] = {"error": r.stderr.strip()[-300:]}
                continue
            by[t] = json.loads(r.stdout.strip().splitlines()[-1])
    finally:
        shutil.rmtree(d, ignore_errors=True)
    ok = {t: v for t, v in by.items() if "async_median" in v}
    if not ok:
        return {"error": by}
    best = max(ok, key=lambda t: ok[t]["async_median"])
    return {"async_frames_per_s": ok[best]["async_median"], "threads": best,
            "sync_frames_per_s": ok[best]["sync"],
            "by_threads": {str(t): {k: v[k] for k in ("async_median", "async_mean", "async_min", "async_max")}
                           for t, v in ok.items()},
            "link_ceiling_frames_per_s": round(LINK_GBPS * 1e9 / ALGO_OUT_PS),
            "note": "C consumer of the ABI (libchromadec_amd/abi_bench.c): chd_decode_frames_async over 384 frames "
                    "of a .tbc file, median of 7 runs, 6.4 MB of yuv444ps host planes per frame (PCIe-inclusive; "
                    "the ceiling is the measured pinned device-to-host rate over the frame size); "
                    "sync is the chd_decode_frame loop"}


def transform2d_leg(fields, steps):
    """CHD_DEC_TRANSFORM_2D on the same 64 frames: per-field 16x32 tiles
    (transform_pal_2d.cpp:107-287) and the same demodulator, fields resident
    in HBM, timed like the headline leg."""
    import torch
    from libchromadec_amd import abi
    dev = torch.device("cuda:0")
    plan = abi.TpalPlan(abi.pal_tpal_params(kind=abi.CHD_DEC_TRANSFORM_2D, pixel_format=abi.CHD_PIXEL_YUV444PS,
                                            max_batch_frames=N_FRAMES))
    src = torch.from_numpy(fields.view(np.int16)).to(dev)
    packed = torch.empty((fields.shape[0], plan.packed_field_elems), dtype=torch.int16, device=dev)
    plan.pack(src.data_ptr(), fields.shape[0], packed.data_ptr())
    out = torch.empty((N_FRAMES, plan.frame_bytes // 4), dtype=torch.float32, device=dev)
    seqs = [(packed.data_ptr(), N_FRAMES, 0, N_FRAMES)]
    stream = torch.cuda.current_stream().cuda_stream
    for _ in range(10):
        plan.decode(seqs, out.data_ptr(), stream)
    torch.cuda.synchronize()
    plan.set_timing(2)
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.decode(seqs, out.data_ptr(), stream)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    tile_ms, demod_ms, n, _ = plan.last_timing()
    plan.set_timing(0)
    tile_avg, demod_avg = tile_ms / n, demod_ms / n
    algo = (ALGO_IN + ALGO_OUT_PS) * N_FRAMES
    tflops = NOMINAL_GFLOP_PER_FRAME_2D * N_FRAMES / ((tile_avg + demod_avg) * 1e-3) / 1e3
    return {"value": round(N_FRAMES * steps / dt, 1), "unit": "frames/s", "ms_per_step": round(1e3 * dt / steps, 4),
            "tile_kernel": "tpal2d_tiles", "tile_avg_launch_ms": round(tile_avg, 4),
            "demod_avg_launch_ms": round(demod_avg, 4),
            "hbm_frac_tile_kernel": round(algo / (tile_avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "valu": {"nominal_gflop_per_frame": NOMINAL_GFLOP_PER_FRAME_2D,
                     "achieved_tflops_nominal": round(tflops, 2),
                     "frac": round(tflops / FP32_PEAK_TFLOPS, 4)}}


def learned_leg(plain_plan, packed, fields, truth, tu, tv, out, seqs, stream, steps):
    import torch
    from libchromadec_amd import abi
    from libchromadec_amd import mask_training as mt
    weights = os.path.join(ROOT, "libchromadec_amd", "data", "tpal3d_mask.bin")
    net = abi.MaskNet(max_tiles=1)
    net.load(weights)
    plan = abi.TpalPlan(abi.pal_tpal_params(kind=abi.CHD_DEC_TRANSFORM_3D, pixel_format=abi.CHD_PIXEL_YUV444PS,
                                            max_batch_frames=N_FRAMES))
    plan.set_mask_net(net)
    for _ in range(5):
        plan.decode(seqs, out.data_ptr(), stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.decode(seqs, out.data_ptr(), stream)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    learned = out.cpu().numpy().reshape(N_FRAMES, 3, 576, 922)
    plain_plan.decode(seqs, out.data_ptr(), stream)
    torch.cuda.synchronize()
    plain = out.cpu().numpy().reshape(N_FRAMES, 3, 576, 922)
    ideal = mt.ideal_frames(plain_plan, fields, truth)
    ecb, ecr = abi.color_scales()
    return {"value": round(N_FRAMES * steps / dt, 1), "unit": "frames/s", "ms_per_step": round(1e3 * dt / steps, 3),
            "weights": "libchromadec_amd/data/tpal3d_mask.bin",
            "psnr_vs_ideal_separation_db": {"plain": round(mt.separation_psnr(plain, ideal), 3),
                                            "learned": round(mt.separation_psnr(learned, ideal), 3)},
            "psnr_vs_encoder_uv_db": {"plain": round(mt.chroma_psnr(plain, tu, tv, range(N_FRAMES), ecb, ecr), 3),
                                      "learned": round(mt.chroma_psnr(learned, tu, tv, range(N_FRAMES), ecb, ecr), 3)}}


def batched_leg(steps):
    import torch
    from libchromadec_amd import abi
    seqs = make_batch_workload()
    plan = abi.TpalPlan(abi.pal_tpal_params(kind=abi.CHD_DEC_TRANSFORM_3D, pixel_format=abi.CHD_PIXEL_YUV444PS,
                                            max_batch_frames=64))
    dev = torch.device("cuda:0")
    keep, runs = [], []
    for f in seqs:
        src = torch.from_numpy(f.view(np.int16)).to(dev)
        pk = torch.empty((f.shape[0], plan.packed_field_elems), dtype=torch.int16, device=dev)
        plan.pack(src.data_ptr(), f.shape[0], pk.data_ptr())
        keep.append(pk)
        runs.append((pk.data_ptr(), f.shape[0] // 2, 0, f.shape[0] // 2))
    total = sum(r[3] for r in runs)
    out = torch.empty((total, plan.frame_bytes // 4), dtype=torch.float32, device=dev)
    batch = abi.TpalBatch(plan, runs, out.data_ptr(), autotune=True)
    strip, groups, tile_ms = batch.config()
    stream = torch.cuda.current_stream().cuda_stream

    def replay_ms(n):
        batch.run(stream)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(n):
            batch.run(stream)
        torch.cuda.synchronize()
        return 1e3 * (time.perf_counter() - t) / n

    # both grids at the chosen strip length, then the autotuner's pick
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    shapes = {}
    for name, g in (("one_per_item", 0), ("persistent_12_per_cu", 12 * cus)):
        batch.set_shape(strip, g)
        shapes[name] = round(replay_ms(max(3, steps // 2)), 3)
    batch.set_shape(strip, groups)
    dt = replay_ms(steps) * steps / 1e3
    return {"value": round(total * steps / dt, 1), "unit": "frames/s", "frames_per_launch": total,
            "sequences": len(seqs), "lengths": [r[3] for r in runs], "ms_per_graph": round(1e3 * dt / steps, 3),
            "autotuned": {"strip_tiles": strip, "workgroups": groups,
                          "grid": "one per work item" if groups == 0 else "persistent",
                          "tile_launch_ms": round(tile_ms, 3)},
            "ms_per_graph_by_grid": shapes}


if __name__ == "__main__":
    sys.exit(main())
