"""GPU: chd_tpal_decode_static, the allocation-free form of chd_tpal_decode.

It decodes in chunks out of a static per-GPU scratch array, so its frames
must equal chd_tpal_decode's bit for bit.  Two calls in a row, the second on
a smaller input, check that the array keeps no state that changes a later
call's result.  Runs that span several chunks (3D holds 4 PAL frames per
chunk) and two sequences per call cover the chunk boundaries.  The check
runs through the C ABI with torch only holding device memory."""
import numpy as np
import pytest

from libchromadec_amd import abi, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def captures():
    out = []
    for seed, n, kw in ((61, 11, dict()), (62, 5, dict(zone_plate=True))):
        scene = synth.PatchScene(seed=seed, **kw)
        out.append(synth.encode_sequence([scene.frame(i) for i in range(n)],
                                         synth.EncodeOptions(snr_db=35, rng=np.random.default_rng(seed))))
    return out


def _upload(plan, fields_list):
    import torch
    keep, runs = [], []
    for f in fields_list:
        src = torch.from_numpy(f.view(np.int16)).to("cuda:0")
        pk = torch.empty((f.shape[0], plan.packed_field_elems), dtype=torch.int16, device="cuda:0")
        plan.pack(src.data_ptr(), f.shape[0], pk.data_ptr())
        keep.append(pk)
        n = f.shape[0] // 2
        runs.append((pk.data_ptr(), n, 0, n))
    return keep, runs


def _both(plan, runs):
    import torch
    total = sum(r[3] for r in runs)
    # int32 words: the comparison is bitwise whatever the output format
    ref = torch.full((total, plan.frame_bytes // 4), 0x5A5A5A5A, dtype=torch.int32, device="cuda:0")
    got = torch.full_like(ref, -0x3C3C3C3D)
    plan.decode(runs, ref.data_ptr())
    plan.decode_static(runs, got.data_ptr())
    torch.cuda.synchronize()
    return ref, got


CASES = [
    dict(kind=abi.CHD_DEC_TRANSFORM_3D),
    dict(kind=abi.CHD_DEC_TRANSFORM_3D, pixel_format=abi.CHD_PIXEL_YUV444P16, luma_nr_level=2.0),
    dict(kind=abi.CHD_DEC_TRANSFORM_2D),
    dict(kind=abi.CHD_DEC_PAL_2D, pixel_format=abi.CHD_PIXEL_RGBS),
]


@pytest.mark.parametrize("case", CASES, ids=["3d", "3d-p16-coring", "2d", "pal2d-rgbs"])
def test_static_decode_matches_decode_on_two_calls(captures, case):
    import torch
    plan = abi.TpalPlan(abi.pal_tpal_params(max_batch_frames=64, **case))
    keep, runs = _upload(plan, captures)
    # first call: 11 + 5 frames, several chunks and two sequences
    ref, got = _both(plan, runs)
    assert torch.equal(ref, got), (case, int((ref != got).sum()))
    # second call, smaller: a 3-frame run starting inside the first sequence
    small = [(runs[0][0], runs[0][1], 4, 3)]
    ref2, got2 = _both(plan, small)
    assert torch.equal(ref2, got2), (case, int((ref2 != got2).sum()))
    assert torch.equal(ref2, ref[4:7])     # a run decodes as the same frames of the whole sequence
    plan.close()


def test_static_decode_rejects_a_learned_plan(captures):
    plan = abi.TpalPlan(abi.pal_tpal_params(max_batch_frames=8))
    net = abi.MaskNet(max_tiles=1)
    plan.set_mask_net(net)
    keep, runs = _upload(plan, captures[:1])
    with pytest.raises(abi.ChdError) as ei:
        plan.decode_static(runs, keep[0].data_ptr())
    assert ei.value.status == 17          # CHD_E_UNSUPPORTED
    plan.close()
    net.close()
