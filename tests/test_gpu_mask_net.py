"""GPU: learned carrier-mirror mask (BF16 matrix-core network, tile spectra,
soft-mask inverse) against the CPU oracle (oracle/mask_net.py,
oracle/chd_oracle.c).

Tolerances: spectra within the kernel's own FP32 error bound (8 u L1 per
tile); network outputs and gradients against the oracle evaluated with the
same BF16 roundings (differences come from FP32 summation order and the odd
BF16 rounding that falls the other way); decoded frames with the learned
mask within 1e-4 of the FP64 oracle decode given the same masks.
"""
import numpy as np
import pytest

from libchromadec_amd import abi
from oracle import mask_net as mn

pytestmark = pytest.mark.gpu

F32_TOL = 1e-4
BLACK = 0x4000


def _torch():
    import torch
    return torch


def _win(n):
    k = np.arange(n)
    return 0.5 - 0.5 * np.cos(2 * np.pi * (k + 0.5) / n)


W3D = _win(8)[:, None, None] * _win(32)[None, :, None] * _win(16)[None, None, :]


def _tile(fields, slab, ty, tx, first=44, last=619, avs=185):
    """TransformPal3D::forwardFFTTile input (transform_pal_3d.cpp:168-202),
    fields 2*slab .. 2*slab+7 (black outside the sequence)."""
    t = np.full((8, 32, 16), float(BLACK))
    tile_y, tile_x = first - 16 + 16 * ty, avs - 8 + 8 * tx
    for z in range(8):
        fi = 2 * slab + z
        if not 0 <= fi < fields.shape[0]:
            continue
        off = z & 1
        for y in range(32):
            ln = tile_y + y
            if first <= ln <= last and ln % 2 == off:
                t[z, y] = fields[fi, ln // 2, tile_x:tile_x + 16]
    return t


def _device_spectra(plan, fields, first_slab, n_slabs):
    torch = _torch()
    dev = torch.device("cuda:0")
    src = torch.from_numpy(fields.view(np.int16)).to(dev)
    packed = torch.empty((fields.shape[0], plan.packed_field_elems), dtype=torch.int16, device=dev)
    plan.pack(src.data_ptr(), fields.shape[0], packed.data_ptr())
    tps = plan.tiles_per_slab
    spec = torch.zeros((n_slabs * tps, abi.TILE_BINS, 2), dtype=torch.float32, device=dev)
    plan.slab_spectra(packed.data_ptr(), fields.shape[0], first_slab, n_slabs, spec.data_ptr())
    torch.cuda.synchronize()
    return packed, spec


def _as_complex(spec_t):
    a = spec_t.cpu().numpy().astype(np.float64)
    return (a[..., 0] + 1j * a[..., 1]).reshape(-1, 5, 32, 8)


@pytest.fixture(scope="module")
def plan3d():
    return abi.TpalPlan(abi.pal_tpal_params(kind=abi.CHD_DEC_TRANSFORM_3D, max_batch_frames=8))


def test_slab_spectra_match_fft(plan3d, patch_capture):
    _, fields = patch_capture
    tps = plan3d.tiles_per_slab
    assert tps == 37 * 117
    _, spec = _device_spectra(plan3d, fields, -1, 3)   # slabs -1, 0, 1
    got = _as_complex(spec)
    rng = np.random.default_rng(0)
    for _ in range(12):
        j, ty, tx = int(rng.integers(0, 3)), int(rng.integers(0, 37)), int(rng.integers(0, 117))
        t = _tile(fields, j - 1, ty, tx)
        ref = np.fft.rfftn(t * W3D)[:, :, 2:7]                       # [z, y, x]
        g = got[j * tps + ty * 117 + tx]                              # [x-2, y, z]
        l1 = np.abs((t - BLACK) * W3D).sum()
        err = np.abs(np.transpose(g, (2, 1, 0)) - ref).max()
        assert err <= 8 * 2.0 ** -24 * l1 + 1e-3, (j, ty, tx, err, l1)


def _params(seed):
    th = mn.init_params(seed)
    th[-1] = 0.25
    return th


def test_forward_matches_oracle(plan3d, patch_capture):
    torch = _torch()
    _, fields = patch_capture
    _, spec = _device_spectra(plan3d, fields, 1, 1)
    n = 512
    spec = spec[1500:1500 + n].contiguous()
    net = abi.MaskNet(max_tiles=200)                      # three groups: 200, 200, 112
    th = _params(11)
    net.set_params(th)
    assert np.array_equal(net.params(), th)
    mask = torch.zeros((n, abi.TILE_VISITS), dtype=torch.float32, device="cuda:0")
    net.forward(spec.data_ptr(), n, mask.data_ptr())
    torch.cuda.synchronize()
    got = mask.cpu().numpy()
    want, _ = mn.forward(th, mn.features(_as_complex(spec)), bf16=True)
    d = np.abs(got - want)
    assert d.max() < 2e-2 and d.mean() < 1e-3, (d.max(), d.mean())
    # deterministic: the same call gives the same bits
    mask2 = torch.zeros_like(mask)
    net.forward(spec.data_ptr(), n, mask2.data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(mask, mask2)


def test_gradient_matches_oracle(plan3d):
    torch = _torch()
    from libchromadec_amd import synth
    scene = synth.PatchScene(seed=21)
    fields, _, _, chroma = synth.encode_sequence([scene.frame(i) for i in range(3)],
                                                 synth.EncodeOptions(snr_db=40, rng=np.random.default_rng(21)),
                                                 return_truth=True)
    truth = np.clip(np.rint(chroma + 32768.0), 0, 65535).astype(np.uint16)
    _, spec = _device_spectra(plan3d, fields, 0, 1)
    tplan = abi.TpalPlan(abi.pal_tpal_params(kind=abi.CHD_DEC_TRANSFORM_3D, max_batch_frames=8, black=32768))
    _, tspec = _device_spectra(tplan, truth, 0, 1)
    sel = np.arange(1200, 1200 + 96 * 17, 17)[:96]
    idx = torch.from_numpy(sel).to("cuda:0")
    s, ts = spec[idx].contiguous(), tspec[idx].contiguous()
    n = len(sel)
    net = abi.MaskNet(max_tiles=n)
    th = _params(12)
    net.set_params(th)
    grad = torch.zeros(abi.MASK_NET_PARAMS, dtype=torch.float32, device="cuda:0")
    F, Ct = _as_complex(s), _as_complex(ts)
    w, num = mn.targets(F, Ct)
    scale = float(1.0 / w.sum())
    loss = net.gradient(s.data_ptr(), ts.data_ptr(), n, scale, grad.data_ptr())
    g = grad.cpu().numpy().astype(np.float64)
    feat = mn.features(F)
    want = mn.backward(th, feat, w, num, scale, bf16=True)
    m, _ = mn.forward(th, feat, bf16=True)
    assert abs(loss - mn.loss(m, w, num, scale)) <= 1e-3 * abs(mn.loss(m, w, num, scale)) + 1e-6
    offs = np.cumsum([0, 27 * 4 * 32, 32, 27 * 32 * 32, 32, 27 * 32, 1])
    for i, name in enumerate(["W1", "b1", "W2", "b2", "W3", "b3"]):
        a, b = g[offs[i]:offs[i + 1]], want[offs[i]:offs[i + 1]]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        assert rel < 3e-2, (name, rel)
    cos = g @ want / (np.linalg.norm(g) * np.linalg.norm(want))
    assert cos > 0.999, cos
    # a few Adam steps on this batch lower the loss
    l0 = loss
    for _ in range(30):
        loss = net.gradient(s.data_ptr(), ts.data_ptr(), n, scale, grad.data_ptr())
        net.adam_step(grad.data_ptr(), 3e-3)
    assert loss < l0


def _decode_device(plan, packed, n_frames, first, count):
    torch = _torch()
    out = torch.zeros((count, plan.frame_bytes // 4), dtype=torch.float32, device="cuda:0")
    plan.decode([(packed.data_ptr(), n_frames, first, count)], out.data_ptr())
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(count, 3, 576, 922)


def test_learned_decode_matches_oracle_with_same_masks(orc, patch_capture):
    torch = _torch()
    _, fields = patch_capture
    plan = abi.TpalPlan(abi.pal_tpal_params(kind=abi.CHD_DEC_TRANSFORM_3D, max_batch_frames=4))
    net = abi.MaskNet(max_tiles=3000)
    th = _params(13)
    net.set_params(th)
    plan.set_mask_net(net)
    nf = fields.shape[0] // 2
    packed, spec = _device_spectra(plan, fields, -2, nf + 2)        # slabs -2 .. nf-1
    tps = plan.tiles_per_slab
    masks = torch.zeros((spec.shape[0], abi.TILE_VISITS), dtype=torch.float32, device="cuda:0")
    net.forward(spec.data_ptr(), spec.shape[0], masks.data_ptr())
    got = _decode_device(plan, packed, nf, 0, nf)                   # learned path, 2 chunks
    m = masks.cpu().numpy().reshape(nf + 2, 37, 117, 768)
    assert 0.05 < m.mean() < 0.95
    p = orc.pal_params()
    for f in (0, 3):
        want = orc.decode_frame_masked(p, fields, f, np.stack([m[f], m[f + 2]]))
        d = np.abs(got[f].astype(np.float64) - want)
        assert d.max() <= F32_TOL, (f, d.max())


def test_learned_decode_independent_of_chunking(patch_capture):
    """With alternating field parity the network runs paired: two tiles' halves
    per work unit.  A tile's mask depends on that tile alone, so decodes in
    chunks of 3 frames (5 slabs: an odd tile count, the last unit half
    empty) and in one chunk of 6 frames (8 slabs) are bit-identical."""
    torch = _torch()
    _, fields = patch_capture
    net = abi.MaskNet(max_tiles=4096)
    net.set_params(_params(17))
    nf = fields.shape[0] // 2
    outs = []
    for chunk in (3, 6):
        plan = abi.TpalPlan(abi.pal_tpal_params(kind=abi.CHD_DEC_TRANSFORM_3D, max_batch_frames=chunk))
        assert (chunk + 2) * plan.tiles_per_slab % 2 == (1 if chunk == 3 else 0)
        plan.set_mask_net(net)
        packed, _ = _device_spectra(plan, fields, 0, 1)
        outs.append(_decode_device(plan, packed, nf, 0, nf))
        plan.close()
    net.close()
    assert np.array_equal(outs[0], outs[1])


def test_all_ones_mask_equals_threshold_zero_bitwise(patch_capture):
    _, fields = patch_capture
    nf = fields.shape[0] // 2
    th = np.zeros(abi.MASK_NET_PARAMS, np.float32)
    th[-1] = 40.0                                      # sigmoid(40) rounds to 1.0f
    net = abi.MaskNet(max_tiles=4096)
    net.set_params(th)
    learned = abi.TpalPlan(abi.pal_tpal_params(kind=abi.CHD_DEC_TRANSFORM_3D, max_batch_frames=8))
    learned.set_mask_net(net)
    plain = abi.TpalPlan(abi.pal_tpal_params(kind=abi.CHD_DEC_TRANSFORM_3D, max_batch_frames=8, threshold=0.0))
    packed, _ = _device_spectra(plain, fields, 0, 1)
    a = _decode_device(learned, packed, nf, 0, nf)
    b = _decode_device(plain, packed, nf, 0, nf)
    assert np.array_equal(a, b)
    # and the plan falls back to the threshold when the network is removed
    learned.set_mask_net(None)
    c = _decode_device(learned, packed, nf, 0, nf)
    d = _decode_device(abi.TpalPlan(abi.pal_tpal_params(kind=abi.CHD_DEC_TRANSFORM_3D, max_batch_frames=8)),
                       packed, nf, 0, nf)
    assert np.array_equal(c, d)


def test_all_ones_mask_non_alternating_windows_bitwise(patch_capture, tmp_path):
    """A second field marked first (SourceField::getOffset,
    source_field.h:77-79) breaks the line-parity alternation of every window
    holding it, so those frames take the learned path's full form (stored
    spectra, one tile per inference unit, soft mask applied to the stored
    bins) while the later frames take the paired form.  With a network whose
    mask is 1 everywhere, both equal the threshold-0 decode bit for bit
    through chd_decode_frame."""
    import json
    from libchromadec_amd import synth
    _, fields = patch_capture
    path = str(tmp_path / "order.tbc")
    side = synth.write_tbc(path, fields)
    with open(side) as fh:
        meta = json.load(fh)
    meta["fields"][3]["isFirstField"] = True   # second field of frame 1
    with open(side, "w") as fh:
        json.dump(meta, fh)
    th = np.zeros(abi.MASK_NET_PARAMS, np.float32)
    th[-1] = 40.0                                      # sigmoid(40) rounds to 1.0f
    net = abi.MaskNet(max_tiles=1)
    net.set_params(th)
    weights = str(tmp_path / "ones.bin")
    net.save(weights)
    net.close()
    v = abi.Video(path)
    learned = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_3D, output_format="yuv444ps", transform_mask="learned",
                          transform_mask_weights=weights).commit()
    plain = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_3D, output_format="yuv444ps", transform_threshold=0.0).commit()
    nf = fields.shape[0] // 2
    for f in (0, 1, 3, nf - 1):                         # windows 0-3 hold field 3; the last does not
        assert np.array_equal(learned.decode(f), plain.decode(f)), f
    learned.close()
    plain.close()


def test_decode_separated_matches_oracle(orc):
    """Demodulation of given (true) chroma: the ideal-separation reference
    the learned mask is measured against."""
    torch = _torch()
    from libchromadec_amd import synth
    scene = synth.PatchScene(seed=31)
    fields, _, _, chroma = synth.encode_sequence([scene.frame(i) for i in range(2)],
                                                 synth.EncodeOptions(snr_db=40, rng=np.random.default_rng(31)),
                                                 return_truth=True)
    plan = abi.TpalPlan(abi.pal_tpal_params(kind=abi.CHD_DEC_TRANSFORM_3D, max_batch_frames=1))
    dev = torch.device("cuda:0")
    src = torch.from_numpy(fields.view(np.int16)).to(dev)
    packed = torch.empty((fields.shape[0], plan.packed_field_elems), dtype=torch.int16, device=dev)
    plan.pack(src.data_ptr(), fields.shape[0], packed.data_ptr())
    ch = torch.from_numpy(chroma.astype(np.float32)).to(dev)
    out = torch.zeros((2, plan.frame_bytes // 4), dtype=torch.float32, device=dev)
    plan.decode_separated(packed.data_ptr(), ch.data_ptr(), 2, out.data_ptr())
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(2, 3, 576, 922)
    p = orc.pal_params()
    for f in range(2):
        want = orc.decode_frame_from_chroma(p, fields[2 * f], fields[2 * f + 1],
                                            chroma[2 * f].astype(np.float32), chroma[2 * f + 1].astype(np.float32))
        d = np.abs(got[f].astype(np.float64) - want)
        assert d.max() <= F32_TOL, (f, d.max())


def _packaged_weights():
    import os
    return os.path.join(os.path.dirname(abi.LIB_PATH), "..", "data", "tpal3d_mask.bin")


def test_abi_learned_option(patch_capture, tmp_path):
    """transform_mask = "learned" through chd_decode_frame: the packaged
    weights by default, the same frame the device API gives with them, and a
    loud error for an explicit weights file that does not load."""
    torch = _torch()
    path, fields = patch_capture
    v = abi.Video(path)
    learned = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_3D, output_format="yuv444ps", transform_mask="learned").commit()
    plain = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_3D, output_format="yuv444ps").commit()
    a, b = learned.decode(2), plain.decode(2)
    assert not np.array_equal(a, b)
    net = abi.MaskNet(max_tiles=4096)
    net.load(_packaged_weights())
    plan = abi.TpalPlan(abi.pal_tpal_params(kind=abi.CHD_DEC_TRANSFORM_3D, max_batch_frames=4))
    plan.set_mask_net(net)
    packed, _ = _device_spectra(plan, fields, 0, 1)
    nf = fields.shape[0] // 2
    dev_frames = _decode_device(plan, packed, nf, 0, nf)
    assert np.array_equal(dev_frames[2], a)
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"not a weights file")
    with pytest.raises(abi.ChdError) as e:
        abi.Decoder(v, abi.CHD_DEC_TRANSFORM_3D, transform_mask="learned",
                    transform_mask_weights=str(bad)).commit()
    assert e.value.name == "CHD_E_NN_MODEL_LOAD"
