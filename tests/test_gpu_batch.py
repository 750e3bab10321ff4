"""GPU: fixed multi-sequence batches (chd_tpal_batch_*): one tile launch over
every slab of several sequences of uneven length, autotuned strip length and
grid, captured in a HIP graph.  Every replay must give the frames
chd_tpal_decode gives for the same runs (which the parity tests tie to the
FP64 oracle)."""
import numpy as np
import pytest

from libchromadec_amd import abi, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sequences():
    """Three short sequences of uneven length and mixed content."""
    out = []
    for seed, n, kw in ((41, 3, dict()), (42, 5, dict(zone_plate=True)), (43, 2, dict(max_speed=16.0))):
        scene = synth.PatchScene(seed=seed, **kw)
        out.append(synth.encode_sequence([scene.frame(i) for i in range(n)],
                                         synth.EncodeOptions(snr_db=35, rng=np.random.default_rng(seed))))
    return out


def _upload(plan, seqs_fields):
    import torch
    dev = torch.device("cuda:0")
    keep, runs = [], []
    for f in seqs_fields:
        src = torch.from_numpy(f.view(np.int16)).to(dev)
        pk = torch.empty((f.shape[0], plan.packed_field_elems), dtype=torch.int16, device=dev)
        plan.pack(src.data_ptr(), f.shape[0], pk.data_ptr())
        keep.append(pk)
        n = f.shape[0] // 2
        runs.append((pk.data_ptr(), n, 0, n))
    return keep, runs


@pytest.mark.parametrize("kind", [abi.CHD_DEC_TRANSFORM_3D, abi.CHD_DEC_TRANSFORM_2D])
def test_batch_graph_replays_match_decode(sequences, kind):
    import torch
    plan = abi.TpalPlan(abi.pal_tpal_params(kind=kind, max_batch_frames=4))
    keep, runs = _upload(plan, sequences)
    total = sum(r[3] for r in runs)
    ref = torch.zeros((total, plan.frame_bytes // 4), dtype=torch.float32, device="cuda:0")
    plan.decode(runs, ref.data_ptr())
    torch.cuda.synchronize()
    out = torch.zeros_like(ref)
    batch = abi.TpalBatch(plan, runs, out.data_ptr(), autotune=(kind == abi.CHD_DEC_TRANSFORM_3D))
    strip, groups, ms = batch.config()
    assert strip in (4, 8, 12, 16, 24) and groups >= 0
    for rep in range(3):
        out.zero_()                                    # eager work between replays
        batch.run(torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (rep, (out - ref).abs().max().item(), strip, groups)
    batch.close()


def test_every_launch_shape_gives_the_same_frames(sequences):
    """Strip lengths 4/8/16 and persistent grids change only how tiles are
    assigned to workgroups, never a result bit."""
    import torch
    plan = abi.TpalPlan(abi.pal_tpal_params(kind=abi.CHD_DEC_TRANSFORM_3D, max_batch_frames=16))
    keep, runs = _upload(plan, sequences)
    total = sum(r[3] for r in runs)
    ref = torch.zeros((total, plan.frame_bytes // 4), dtype=torch.float32, device="cuda:0")
    plan.decode(runs, ref.data_ptr())
    out = torch.zeros_like(ref)
    b = abi.TpalBatch(plan, runs, out.data_ptr(), autotune=False)
    assert b.config()[0] in (4, 8, 12, 16, 24) and b.config()[1] == 0   # launch-size pick, no tuning
    for strip in (4, 8, 12, 16, 24):
        for groups in (0, 8, 256, 1024):
            b.set_shape(strip, groups)
            out.zero_()
            b.run()
            torch.cuda.synchronize()
            assert torch.equal(out, ref), (strip, groups)
    with pytest.raises(abi.ChdError):
        b.set_shape(5, 0)


def test_learned_batch_matches_decode_every_shape(sequences):
    """The learned mask in a captured batch (magnitude features, paired
    inference, the masked inverse with its staged whole-line stores) gives
    the frames chd_tpal_decode gives, for strip lengths 4/8/24 and for one
    workgroup per work item as well as persistent grids."""
    import os
    import torch
    net = abi.MaskNet(max_tiles=1)
    net.load(os.path.join(os.path.dirname(abi.__file__), "data", "tpal3d_mask.bin"))
    plan = abi.TpalPlan(abi.pal_tpal_params(kind=abi.CHD_DEC_TRANSFORM_3D, max_batch_frames=16))
    plan.set_mask_net(net)
    keep, runs = _upload(plan, sequences)
    total = sum(r[3] for r in runs)
    ref = torch.zeros((total, plan.frame_bytes // 4), dtype=torch.float32, device="cuda:0")
    plan.decode(runs, ref.data_ptr())
    torch.cuda.synchronize()
    out = torch.zeros_like(ref)
    b = abi.TpalBatch(plan, runs, out.data_ptr(), autotune=False)
    for strip in (4, 8, 24):
        for groups in (0, 8, 256):
            b.set_shape(strip, groups)
            out.zero_()
            b.run()
            torch.cuda.synchronize()
            assert torch.equal(out, ref), (strip, groups, (out - ref).abs().max().item())
    b.close()
    plan.close()
    net.close()


@pytest.mark.parametrize("kind", [abi.CHD_DEC_TRANSFORM_3D, abi.CHD_DEC_PAL_2D])
def test_luma_nr_plan_and_batch(orc, sequences, kind):
    """Luma coring (luma_nr_level, PalColour::doYNR) through chd_tpal_decode in
    chunks of 4 frames and through a captured batch with its own raw planes:
    the plan's frames match the oracle's cored decode, and every replay gives
    them bit for bit."""
    import torch
    plan = abi.TpalPlan(abi.pal_tpal_params(kind=kind, max_batch_frames=4, luma_nr_level=1.5))
    keep, runs = _upload(plan, sequences)
    total = sum(r[3] for r in runs)
    ref = torch.zeros((total, plan.frame_bytes // 4), dtype=torch.float32, device="cuda:0")
    plan.decode(runs, ref.data_ptr())
    torch.cuda.synchronize()
    got = ref.cpu().numpy().reshape(total, 3, 576, 922)
    p = orc.pal_params(luma_nr_level=1.5)
    base = 0
    for f, r in zip(sequences, runs):
        n = r[3]
        for i in sorted({0, n - 1}):
            want = orc.decode_frame(p, kind, f, i, orc.uniform_thresholds(kind), 0)
            d = np.abs(got[base + i].astype(np.float64) - want)
            assert d.max() <= 1e-4, (base + i, d.max())
        base += n
    out = torch.zeros_like(ref)
    batch = abi.TpalBatch(plan, runs, out.data_ptr(), autotune=False)
    for rep in range(2):
        out.zero_()
        batch.run(torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), rep
    batch.close()
