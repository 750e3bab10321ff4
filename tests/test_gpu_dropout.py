"""GPU: chd_decode_frame with dropout concealment (chd_decoder_set_dropout)
against the FP64 oracle decode of the same window after the oracle's
concealment (every frame of the look-around window is corrected, as
c_api_decoder.cpp:297-315 does)."""
import numpy as np
import pytest

from libchromadec_amd import abi, synth
from oracle import dropout as od

pytestmark = pytest.mark.gpu

PAL_VP = dict(field_width=1135, field_height=313, colour_burst_end=138, active_video_end=1107,
              first_active_frame_line=44, last_active_frame_line=619, system="PAL")


@pytest.fixture(scope="module")
def dropout_capture(tmp_path_factory):
    rng = np.random.default_rng(6)
    scene = synth.PatchScene(seed=6, max_speed=16.0)
    fields = synth.encode_sequence([scene.frame(i) for i in range(5)],
                                   synth.EncodeOptions(snr_db=30, phase_jitter_deg=1.0, rng=rng))
    drops = synth.add_dropouts(fields, rng)
    path = str(tmp_path_factory.mktemp("drop") / "stress.tbc")
    synth.write_tbc(path, fields, drops)
    return path, fields, drops


@pytest.mark.parametrize("kind", [abi.CHD_DEC_TRANSFORM_3D, abi.CHD_DEC_TRANSFORM_2D])
def test_corrected_decode_matches_oracle(orc, dropout_capture, kind):
    path, fields, drops = dropout_capture
    corrected = fields.copy()
    stats = []
    for f in range(fields.shape[0] // 2):
        a, b, st = od.correct_frame(PAL_VP, fields[2 * f], fields[2 * f + 1], drops[2 * f], drops[2 * f + 1])
        corrected[2 * f], corrected[2 * f + 1] = a, b
        stats.append(st)
    v = abi.Video(path)
    dec = abi.Decoder(v, kind, output_format="yuv444ps").set_dropout(True).commit()
    plain = abi.Decoder(v, kind, output_format="yuv444ps").commit()
    p = orc.pal_params()
    for f in (1, 3):
        got = dec.decode(f)
        assert dec.last_dropout_stats() == stats[f]
        want = orc.decode_frame(p, kind, corrected, f, orc.uniform_thresholds(kind), 0)
        assert np.abs(got.astype(np.float64) - want).max() <= 1e-4
        assert not np.array_equal(got, plain.decode(f))
    # switching correction off again gives the uncorrected decode
    dec.set_dropout(False)
    assert np.array_equal(dec.decode(3), plain.decode(3))
    assert dec.last_dropout_stats() == (0, 0, 0)


def test_overlapping_pairs_conceal_each_pair(orc, tmp_path):
    """Field 3 marked second and field 4 first: frame 2 pairs fields 4 and 5
    (both marked first, decoded as is) and frame 3 pairs fields 5 and 6
    (1-based), so field 5 is concealed twice, once with each partner, and
    one 3D window holds both copies.  Each frame of the window is concealed
    as its own pair, as the reference does (c_api_decoder.cpp:297-315)."""
    import json
    rng = np.random.default_rng(61)
    scene = synth.PatchScene(seed=61)
    fields = synth.encode_sequence([scene.frame(i) for i in range(6)],
                                   synth.EncodeOptions(snr_db=35, rng=rng))
    drops = [[] for _ in range(fields.shape[0])]
    drops[4] = [(300, 640, 101), (500, 900, 140)]     # field 5: startx, endx, field line (1-based)
    drops[3] = [(250, 420, 120)]
    drops[5] = [(400, 700, 90)]
    for i in (3, 4, 5):
        for sx, ex, ln in drops[i]:
            fields[i, ln - 1, sx:ex] = 0x0100
    path = str(tmp_path / "overlap.tbc")
    side = synth.write_tbc(path, fields, drops)
    with open(side) as fh:
        meta = json.load(fh)
    meta["fields"][2]["isFirstField"] = False
    meta["fields"][3]["isFirstField"] = True
    with open(side, "w") as fh:
        json.dump(meta, fh)
    first = [bool(r["isFirstField"]) for r in meta["fields"]]
    # the library's pairing (Metadata::fieldNumber): advance to a first field
    pairs, k = [], 0
    while 2 * k + 1 < fields.shape[0]:
        a = 2 * k
        while a < fields.shape[0] - 1 and not first[a]:
            a += 1
        if a + 1 >= fields.shape[0]:
            break
        pairs.append((a, a + 1))
        k += 1
    assert pairs[1] == (3, 4) and pairs[2] == (4, 5)    # 0-based: field 5 in two pairs
    conc = []
    for a, b in pairs:
        ca, cb, _ = od.correct_frame(PAL_VP, fields[a], fields[b], drops[a], drops[b])
        conc.append((ca, cb, 0 if first[a] else 1, 0 if first[b] else 1))
    assert not np.array_equal(conc[1][1], conc[2][0])   # the two concealments of field 5 differ
    v = abi.Video(path)
    dec = abi.Decoder(v, abi.CHD_DEC_TRANSFORM_3D, output_format="yuv444ps").set_dropout(True).commit()
    p = orc.pal_params()
    black = np.full(fields.shape[1:], p.black16b, np.uint16)
    for f in (1, 2, 3):
        win, offs = [], []
        for fr in range(f - 2, f + 5):
            if 0 <= fr < len(pairs):
                win += [conc[fr][0], conc[fr][1]]
                offs += [conc[fr][2], conc[fr][3]]
            else:
                win += [black, black]
                offs += [0, 1]
        want = orc.decode_window(p, 8, np.stack(win), offs, 4, 6, orc.uniform_thresholds(8), 0)
        got = dec.decode(f)
        assert np.abs(got.astype(np.float64) - want).max() <= 1e-4, f
    dec.close()
    v.close()


@pytest.mark.parametrize("kind", [abi.CHD_DEC_TRANSFORM_3D, abi.CHD_DEC_TRANSFORM_2D])
def test_async_batches_with_concealment_match_sync(dropout_capture, kind):
    """Async workers decode runs of consecutive frames as one submission
    (concealed fields staged and packed together): each frame equals its
    synchronous decode with the same concealment."""
    path, _, _ = dropout_capture
    v = abi.Video(path)
    dec = abi.Decoder(v, kind, output_format="yuv444ps", thread_count=2).set_dropout(True).commit()
    want = {f: dec.decode(f) for f in range(5)}
    got = dec.decode_async([0, 1, 2, 3, 4, 4, 3, 2, 1, 0])
    assert sorted(got) == list(range(5))
    for f in range(5):
        assert np.array_equal(got[f], want[f]), f
    dec.close()
    v.close()
