import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def synth():
    from libchromadec_amd import synth as s
    return s


@pytest.fixture(scope="session")
def capture_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("captures"))


@pytest.fixture(scope="session")
def bars_capture(synth, capture_dir):
    """3 frames of colour bars + moving gradient (the CPU 2D workload's content)."""
    fields = synth.encode_sequence([synth.bars_gradient(f) for f in range(3)])
    path = os.path.join(capture_dir, "bars.tbc")
    synth.write_tbc(path, fields)
    return path, fields


@pytest.fixture(scope="session")
def patch_capture(synth, capture_dir):
    """6 frames of moving patches over bars with 40 dB noise (3D workload content)."""
    scene = synth.PatchScene(seed=2)
    fields = synth.encode_sequence([scene.frame(f) for f in range(6)],
                                   synth.EncodeOptions(snr_db=40, rng=np.random.default_rng(2)))
    path = os.path.join(capture_dir, "patches.tbc")
    synth.write_tbc(path, fields)
    return path, fields


_PARITY_ROWS = []


@pytest.fixture(scope="session")
def parity_rows():
    """Worst GPU - oracle errors per benchmark configuration, collected by
    tests/test_gpu_configs.py: (config, frames, output format, stats)."""
    return _PARITY_ROWS


def pytest_sessionfinish(session, exitstatus):
    """With CHD_PARITY_TABLE=<path>, write the collected rows as a markdown
    table (float planes: full-scale units; yuv444p16: LSB)."""
    path = os.environ.get("CHD_PARITY_TABLE")
    if not path or not _PARITY_ROWS:
        return
    lines = ["| Config | Frames checked | Output | E'Y max | U/V max | U/V RMS |", "|---|---|---|---|---|---|"]
    for name, frames, fmt, s in _PARITY_ROWS:
        if fmt == "p16":
            vals = f"{s['y_max']:.0f} LSB | {s['uv_max']:.0f} LSB | {s['uv_rms']:.3f} LSB"
        else:
            vals = f"{s['y_max']:.2e} | {s['uv_max']:.2e} | {s['uv_rms']:.2e}"
        lines.append(f"| {name} | {frames} | {fmt} | {vals} |")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
