"""The command-line driver end to end on the reference's own input file
(needs an MI355X): float32 [n][T] in, mode_matrix.bin [N_imf][n][T] float32 out,
checked bit-for-bit against the FP64 oracle rounded to float32."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "memd-gpu_amd", "memd_cli")
GOLD = os.path.join(ROOT, "tests", "golden", "eeg_sample_32x30504.npz")


def _sample_file(tmp_path):
    import memd_amd
    import synth
    y = synth.load_eeg_sample(GOLD)
    p = tmp_path / "EEGsample_32CH_30504.bin"
    memd_amd.write_bin(p, y)
    return p, y


def test_cli_main_layout(tmp_path, oracle):
    p, y = _sample_file(tmp_path)
    out = tmp_path / "mode_matrix.bin"
    r = subprocess.run([CLI, "--input", str(p), "--output", str(out), "--imfs", "3", "--sifts", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("Channel: 32, Length: 30504, DirVec: 64, Time: ")
    got = np.fromfile(out, np.float32).reshape(3, 32, 30504)
    imfs_o, _ = oracle.memd(y.astype(np.float64), oracle.directions(32, 64), 3, 2)
    assert np.array_equal(got, imfs_o.astype(np.float32))


def test_cli_bench_lengths(tmp_path, oracle):
    # runBenchmark: the first L samples of every channel, one line per length
    p, y = _sample_file(tmp_path)
    out = tmp_path / "mode_matrix.bin"
    r = subprocess.run([CLI, "--input", str(p), "--output", str(out), "--bench", "--lengths", "2049,4097,99999",
                        "--channels", "32", "--directions", "64", "--imfs", "2", "--sifts", "1"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert [ln.split(",")[1].strip() for ln in lines] == ["Length: 2049", "Length: 4097"]   # 99999 > file
    got = np.fromfile(out, np.float32).reshape(2, 32, 4097)
    imfs_o, _ = oracle.memd(y[:, :4097].astype(np.float64), oracle.directions(32, 64), 2, 1)
    assert np.array_equal(got, imfs_o.astype(np.float32))
