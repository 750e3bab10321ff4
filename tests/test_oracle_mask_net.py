"""Pins of the learned-mask oracle (oracle/mask_net.py and the soft-mask
decode of oracle/chd_oracle.c) against independent computations:

* the network's forward and backward against torch.nn.functional.conv3d
  with autograd (float64), and against finite differences;
* the visit and partner index maps against a brute-force walk of the
  reference's applyFilter loop (transform_pal_3d.cpp:277-330);
* the soft mask with the hard rule's own decisions as 0/1 values against
  the hard rule (orc_apply_filter3d), bit for bit;
* the soft-mask frame decode with an all-ones mask against the threshold-0
  decode, and its linearity in the mask.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import mask_net as mn

torch = pytest.importorskip("torch")


def _torch_net(theta):
    W1, b1, W2, b2, W3, b3 = [torch.tensor(np.asarray(v, np.float64)) for v in mn.split_params(theta)]
    return [W1, b1, W2, b2, W3, b3]


def _torch_forward(params, feat):
    """Independent forward: conv3d on [T, C, Z, Y, X] with circular padding in
    z and y and zero padding in x."""
    W1, b1, W2, b2, W3, b3 = params

    def conv(x, W, b):
        cin, cout = W.shape[1], W.shape[2]
        k = W.reshape(3, 3, 3, cin, cout).permute(4, 3, 0, 1, 2)  # [cout, cin, dz, dy, dx]
        x = torch.nn.functional.pad(x, (0, 0, 1, 1, 1, 1), mode="circular")
        x = torch.nn.functional.pad(x, (1, 1))
        return torch.nn.functional.conv3d(x, k, b)

    t = feat.shape[0]
    x = feat.reshape(t, mn.Z, mn.Y, mn.X, 4).permute(0, 4, 1, 2, 3)
    x = torch.relu(conv(x, W1, b1))
    x = torch.relu(conv(x, W2, b2))
    x = conv(x, W3, b3)
    return torch.sigmoid(x[:, 0].reshape(t, mn.P))


def _random_case(seed, t=3):
    rng = np.random.default_rng(seed)
    spec = (rng.normal(size=(t, 5, 32, 8)) + 1j * rng.normal(size=(t, 5, 32, 8))) * np.exp(rng.uniform(6, 14, (t, 1, 1, 1)))
    truth = spec * rng.uniform(0, 1, spec.shape) + 0.1 * rng.normal(size=spec.shape) * np.abs(spec)
    theta = mn.init_params(seed).astype(np.float64)
    theta[mn.N_PARAMS - 1] = 0.3
    return spec, truth, theta


def test_param_count_matches_header():
    import os
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "chromadec", "tpal_mask.h")).read()
    assert f"#define CHD_MASK_NET_PARAMS {mn.N_PARAMS}" in hdr
    assert mn.N_PARAMS == 32033


def test_pair_view_and_bin_map_follow_the_reference_loop():
    rng = np.random.default_rng(1)
    spec = rng.normal(size=(2, 5, 32, 8)) + 1j * rng.normal(size=(2, 5, 32, 8))
    a, b = mn.pair_view(spec)
    m = rng.uniform(size=(2, mn.P))
    bins = mn.visit_to_bins(m)
    owner = {}
    for z in range(8):
        for y in range(32):
            for x in (2, 3, 4):  # transform_pal_3d.cpp:277-330 visit loop
                zr, yr, xr = (2 + 8 - z) % 8, (8 + 32 - y) % 32, 8 - x
                v = (z * 32 + y) * 3 + (x - 2)
                assert np.array_equal(a[:, v], spec[:, x - 2, y, z])
                assert np.array_equal(b[:, v], spec[:, xr - 2, yr, zr])
                owner.setdefault((x, y, z), v)
                if xr != x:
                    owner[(xr, yr, zr)] = v
    assert len(owner) == 5 * 32 * 8
    for (x, y, z), v in owner.items():
        assert np.array_equal(bins[:, x - 2, y, z], m[:, v])


def test_forward_matches_torch_conv3d():
    spec, _, theta = _random_case(2)
    feat = mn.features(spec)
    m, _ = mn.forward(theta, feat)
    ref = _torch_forward(_torch_net(theta), torch.tensor(feat)).numpy()
    assert np.abs(m - ref).max() < 1e-12


def test_backward_matches_torch_autograd():
    spec, truth, theta = _random_case(3)
    feat = mn.features(spec)
    w, num = mn.targets(spec, truth)
    scale = 1.0 / w.sum()
    g = mn.backward(theta, feat, w, num, scale)
    params = [p.requires_grad_() for p in _torch_net(theta)]
    mt = _torch_forward(params, torch.tensor(feat))
    loss = scale * torch.sum(torch.tensor(w) * mt * mt - 2.0 * torch.tensor(num) * mt)
    loss.backward()
    ref = np.concatenate([p.grad.numpy().ravel() for p in params])
    assert abs(float(loss) - mn.loss(mt.detach().numpy(), w, num, scale)) < 1e-12 * max(1.0, abs(float(loss)))
    assert np.abs(g - ref).max() <= 1e-10 * np.abs(ref).max()


def test_backward_matches_finite_differences():
    spec, truth, theta = _random_case(4, t=2)
    feat = mn.features(spec)
    w, num = mn.targets(spec, truth)
    scale = 1.0 / w.sum()
    g = mn.backward(theta, feat, w, num, scale)
    rng = np.random.default_rng(0)
    offs = np.cumsum([0, 27 * 4 * 32, 32, 27 * 32 * 32, 32, 27 * 32, 1])
    picks = [int(rng.integers(offs[i], offs[i + 1])) for i in range(6) for _ in range(3)]
    for k in picks:
        h = 1e-6 * max(1.0, abs(theta[k]))
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (mn.loss(mn.forward(tp, feat)[0], w, num, scale) - mn.loss(mn.forward(tm, feat)[0], w, num, scale)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-4 * abs(g[k]), (k, fd, g[k])


def test_targets_make_the_truth_the_optimum():
    spec, _, _ = _random_case(5)
    w, num = mn.targets(spec, spec)          # truth = composite: keep everything
    assert np.allclose(num, w)
    w0, num0 = mn.targets(spec, 0 * spec)    # no chroma: discard everything
    assert np.all(num0 == 0) and np.allclose(w0, w)


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5 - 2 ** -9, 3.0e-3], np.float32)
    r = mn.bf16_round(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -2.5
    assert abs(r[4] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_soft_mask_with_hard_decisions_equals_hard_rule(orc):
    rng = np.random.default_rng(6)
    L = orc.lib()
    L.orc_apply_soft_mask3d.argtypes = [C.POINTER(C.c_double)] * 4 + [C.POINTER(C.c_float)]
    for trial in range(4):
        h = (rng.normal(size=(8, 32, 9)) + 1j * rng.normal(size=(8, 32, 9))) * 100.0
        t = rng.uniform(0.2, 0.9)
        thr = np.full(768, t * t)
        hard = orc.apply_filter(h, thr, 8)
        # the reference's decision per visit (transform_pal_3d.cpp:313-327)
        m = np.zeros(768, np.float32)
        for z in range(8):
            for y in range(32):
                for x in (2, 3, 4):
                    zr, yr = (10 - z) % 8, (40 - y) % 32
                    a, b = h[z, y, x], h[zr, yr, 8 - x]
                    ma, mb = abs(a) ** 2, abs(b) ** 2
                    selfm = x == 4 and y == yr and z == zr
                    m[(z * 32 + y) * 3 + x - 2] = 1.0 if selfm or not (ma < mb * t * t or mb < ma * t * t) else 0.0
        re, im = np.ascontiguousarray(h.real), np.ascontiguousarray(h.imag)
        ore, oim = np.zeros_like(re), np.zeros_like(im)
        dp = lambda v: v.ctypes.data_as(C.POINTER(C.c_double))
        L.orc_apply_soft_mask3d(dp(re), dp(im), dp(ore), dp(oim), m.ctypes.data_as(C.POINTER(C.c_float)))
        assert np.array_equal(ore + 1j * oim, hard)


def test_masked_decode_all_ones_is_threshold_zero_and_linear(orc, patch_capture):
    _, fields = patch_capture
    geo = dict(first_line=200, last_line=263, avs=400, ave=528)
    p = orc.pal_params(**geo)
    # tile lattice over this crop (transform_pal_3d.cpp:151-153)
    n_ty = (263 - (200 - 16)) // 16 + 1
    n_tx = (528 - (400 - 8) + 7) // 8
    ones = np.ones((2, n_ty, n_tx, 768), np.float32)
    got = orc.decode_frame_masked(p, fields, 2, ones)
    want = orc.decode_frame(p, 8, fields, 2, orc.uniform_thresholds(8, 0.0))
    assert np.array_equal(got, want)
    rng = np.random.default_rng(7)
    m1 = rng.uniform(0, 0.5, ones.shape).astype(np.float32)
    m2 = rng.uniform(0, 0.5, ones.shape).astype(np.float32)
    zero = orc.decode_frame_masked(p, fields, 2, 0 * ones)
    lhs = orc.decode_frame_masked(p, fields, 2, m1 + m2)
    rhs = orc.decode_frame_masked(p, fields, 2, m1) + orc.decode_frame_masked(p, fields, 2, m2) - zero
    assert np.abs(lhs[1:] - rhs[1:]).max() < 1e-5   # U/V linear in the mask
    assert np.all(zero[1:] == 0.0)


def test_demod_from_chroma_reproduces_the_full_decode(orc, patch_capture):
    """decode_frame_from_chroma is the second half of decode_window: fed the
    transform's own chroma it gives the same frame."""
    _, fields = patch_capture
    p = orc.pal_params(first_line=200, last_line=263, avs=400, ave=528)
    fw, offs, s, e = orc.frame_window(fields, 2, 8)
    full, chroma = orc.decode_window(p, 8, fw, offs, s, e, orc.uniform_thresholds(8), want_chroma=True)
    got = orc.decode_frame_from_chroma(p, fields[4], fields[5], chroma[0], chroma[1])
    assert np.array_equal(got, full)


def test_tile_spectra_match_rfftn(orc):
    """orc_tile_spectra3d against numpy's rfftn of a brute-force windowed
    tile built from transform_pal_3d.cpp:168-202 (lines outside the active
    lines or of the other parity hold the fill value)."""
    p = orc.pal_params(first_line=200, last_line=231, avs=400, ave=416)
    rng = np.random.default_rng(1)
    fields = [rng.integers(10000, 60000, size=(313, 1135)).astype(np.float64) for _ in range(8)]
    offs = [z & 1 for z in range(8)]
    spec = orc.tile_spectra3d(p, fields, offs, 16384.0)
    assert spec.shape == (9, 5, 32, 8)
    for t, (tile_y, tile_x) in enumerate([(184 + 16 * a, 392 + 8 * b) for a in range(3) for b in range(3)]):
        tile = np.zeros((8, 32, 16))
        for z in range(8):
            for y in range(32):
                fy = tile_y + y
                v = (fields[z][fy // 2, tile_x:tile_x + 16]
                     if 200 <= fy <= 231 and fy % 2 == offs[z] else np.full(16, 16384.0))
                tile[z, y] = v * orc.window(z, 8) * orc.window(y, 32) * np.array([orc.window(x, 16) for x in range(16)])
        want = np.transpose(np.fft.rfftn(tile)[:, :, 2:7], (2, 1, 0))
        assert np.abs(spec[t] - want).max() <= 1e-9 * np.abs(want).max(), t


def test_wiener_optimal_mask_ceiling(orc):
    """The least-squares optimal soft mask (per visit, from the encoder's
    true modulated chroma: m = sum Re(conj(C) F) / sum |F|^2 over the visit's
    bins, clipped to [0, 1]) is an upper bound for what any learned mask of
    this form can reach.  Against the ideal separation (the demodulator fed
    the true chroma) it gains several dB over the 0.4 threshold; against the
    encoder's U/V it gains well under 1 dB, because the demodulator's own
    1.1 MHz, 7-line chroma filter sets that error for every mask: the +1 dB
    against encoder U/V is out of reach of any chroma/luma separation mask."""
    from libchromadec_amd import abi, synth
    from libchromadec_amd import mask_training as mt
    scene = synth.PatchScene(seed=5, n_patches=8)
    fields, u, v, chroma = synth.encode_sequence([scene.frame(i) for i in range(7)],
                                                 synth.EncodeOptions(snr_db=40, rng=np.random.default_rng(5)),
                                                 return_truth=True)
    p = orc.pal_params()
    f = 2                                   # its window is frames 0..6
    offs = [z & 1 for z in range(8)]
    n_ty, n_tx = 37, 117
    masks = []
    for s0 in (0, 4):                       # the window's two Z tiles: fields 2f-4 and 2f
        F = orc.tile_spectra3d(p, fields[s0:s0 + 8], offs, float(p.black16b))
        Ct = orc.tile_spectra3d(p, chroma[s0:s0 + 8], offs, 0.0)
        w, num = mn.targets(F, Ct)
        m = np.where(w > 0, np.clip(num / np.where(w > 0, w, 1.0), 0.0, 1.0), 0.0)
        masks.append(m.reshape(n_ty, n_tx, 768))
    best = orc.decode_frame_masked(p, fields, f, np.stack(masks), 0)[None]
    plain = orc.decode_frame(p, 8, fields, f, orc.uniform_thresholds(8), 0)[None]
    ideal = orc.decode_frame_from_chroma(p, fields[2 * f], fields[2 * f + 1], chroma[2 * f], chroma[2 * f + 1], 0)[None]
    ecb, ecr = abi.color_scales()
    sep = {k: mt.separation_psnr(x, ideal) for k, x in (("plain", plain), ("best", best))}
    uv = {k: mt.chroma_psnr(x, u, v, [f], ecb, ecr) for k, x in (("plain", plain), ("best", best))}
    print("optimal mask: separation PSNR", sep, "encoder U/V PSNR", uv)
    assert sep["best"] - sep["plain"] >= 3.0, (sep, uv)
    assert 0.0 <= uv["best"] - uv["plain"] < 0.5, (sep, uv)
