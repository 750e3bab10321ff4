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
