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


def test_cli_bench_channel_sweep(tmp_path, oracle):
    # runBenchmark's file table (:1377-1383): one file per channel count, K = 2n
    import memd_amd
    import synth
    paths, ys = [], []
    for n in (4, 8):
        y = synth.eeg_epoch(n, 5000, 80 + n).astype(np.float32)
        p = tmp_path / f"InputData_{n}CH_5000.bin"
        memd_amd.write_bin(p, y)
        paths.append(str(p))
        ys.append(y)
    out = tmp_path / "mode_matrix.bin"
    r = subprocess.run([CLI, "--output", str(out), "--bench", "--files", ",".join(paths), "--channel-list", "4,8",
                        "--lengths", "2049,4097", "--imfs", "2", "--sifts", "1"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    heads = [ln.rsplit(",", 1)[0] for ln in r.stdout.strip().splitlines()]
    assert heads == ["Channel: 4, Length: 2049, DirVec: 8", "Channel: 4, Length: 4097, DirVec: 8",
                     "Channel: 8, Length: 2049, DirVec: 16", "Channel: 8, Length: 4097, DirVec: 16"]
    # the last run's output: 8 channels, 4097 samples, K = 16
    got = np.fromfile(out, np.float32).reshape(2, 8, 4097)
    imfs_o, _ = oracle.memd(ys[1][:, :4097].astype(np.float64), oracle.directions(8, 16), 2, 1)
    assert np.array_equal(got, imfs_o.astype(np.float32))


def test_cli_end_options(tmp_path, oracle):
    # --spline / --boundary reach the run: memd.m's setting against the oracle
    p, y = _sample_file(tmp_path)
    out = tmp_path / "mode_matrix.bin"
    r = subprocess.run([CLI, "--input", str(p), "--output", str(out), "--length", "3000", "--imfs", "3",
                        "--sifts", "2", "--spline", "not-a-knot", "--boundary", "mirror"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, np.float32).reshape(3, 32, 3000)
    imfs_o, _ = oracle.memd(y[:, :3000].astype(np.float64), oracle.directions(32, 64), 3, 2,
                            spline=oracle.SPLINE_NOT_A_KNOT, boundary=oracle.BOUNDARY_MIRROR, threads=16)
    assert np.array_equal(got, imfs_o.astype(np.float32))
