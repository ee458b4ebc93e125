"""C-ABI checks that need no GPU: the library loads, exports every declared
symbol, and its host-only entry points behave."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "memd_amd.h")


def _declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:memd_status|const char\*|void|size_t)\s+(memd_\w+)\s*\(", src, re.M)))


def test_header_declares_expected_api():
    import memd_amd
    assert sorted(memd_amd.EXPORTS) == _declared()


def test_library_exports_every_declared_symbol():
    import memd_amd
    lib = memd_amd.load_library()
    for name in _declared():
        assert hasattr(lib, name), name


def test_defaults_and_version():
    import memd_amd
    o = memd_amd.default_opts()
    assert (o.mode, o.window, o.use_graph) == (0, 0, 1)
    assert (o.sd, o.sd2, o.tol) == (0.075, 0.75, 0.075)
    assert b"gfx950" in memd_amd.load_library().memd_version()


def test_workspace_bytes_scale():
    import memd_amd
    lib = memd_amd.load_library()
    small = lib.memd_workspace_bytes(32, 4096, 64)
    big = lib.memd_workspace_bytes(32, 30504, 64)
    assert 0 < small < big
    assert lib.memd_workspace_bytes(0, 4096, 64) == 0


def test_null_context_is_rejected():
    import memd_amd
    lib = memd_amd.load_library()
    assert lib.memd_run_f64(None, 1, 10, 1, 2, 1, None, None, None, None, None, None) == 1
    assert lib.memd_last_error(None) == b"null context"


def test_context_without_gpu_reports_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import memd_amd
    h = ctypes.c_void_p()
    rc = memd_amd.load_library().memd_ctx_create(0, ctypes.byref(h))
    assert rc == 2 and not h.value


def test_python_mirror_refuses_missing_library(tmp_path):
    import memd_amd
    saved = memd_amd._lib
    try:
        memd_amd._lib = None
        with pytest.raises(ImportError):
            memd_amd.load_library(str(tmp_path / "nope.so"))
    finally:
        memd_amd._lib = saved


def test_bin_files_round_trip(tmp_path):
    # channel-major float32 [n][T], no header (the reference demo's file format)
    import numpy as np
    import memd_amd
    rng = np.random.default_rng(3)
    y = rng.standard_normal((5, 123)).astype(np.float32)
    p = tmp_path / "in.bin"
    memd_amd.write_bin(p, y)
    assert p.stat().st_size == y.nbytes
    assert np.array_equal(np.fromfile(p, np.float32).reshape(5, 123), y)
    assert memd_amd.bin_length(p, 5) == 123
    assert np.array_equal(memd_amd.read_bin(p, 5), y)
    # the first T samples of every channel of a longer file (runBenchmark's slicing)
    assert np.array_equal(memd_amd.read_bin(p, 5, 50), y[:, :50])
    # trailing bytes: the length floors like the reference (bytes / 4 / n)
    with open(p, "ab") as f:
        f.write(b"\0" * 7)
    assert memd_amd.bin_length(p, 5) == (y.nbytes + 7) // 4 // 5 == 123


def test_bin_files_errors(tmp_path):
    import numpy as np
    import memd_amd
    with pytest.raises(memd_amd.MemdError) as ei:
        memd_amd.bin_length(tmp_path / "missing.bin", 4)
    assert ei.value.status == 8 and "missing.bin" in str(ei.value)
    p = tmp_path / "short.bin"
    memd_amd.write_bin(p, np.zeros(10, np.float32))
    lib = memd_amd.load_library()
    buf = np.empty(40, np.float32)
    # asking for more samples per channel than the file holds
    rc = lib.memd_read_bin_f32(os.fsencode(p), 2, 20, 20, buf.ctypes.data_as(ctypes.c_void_p))
    assert rc == 8 and b"too short" in lib.memd_last_error(None)
    rc = lib.memd_read_bin_f32(os.fsencode(p), 2, 5, 6, buf.ctypes.data_as(ctypes.c_void_p))
    assert rc == 1
    with pytest.raises(memd_amd.MemdError) as ei:
        memd_amd.write_bin(tmp_path / "no" / "dir" / "x.bin", np.zeros(3, np.float32))
    assert ei.value.status == 8


def test_reference_eeg_sample_file_layout(tmp_path):
    # the reference's own input (32 ch x 30504 float32, kept as a fixture) written in
    # its on-disk layout has the reference file's size, and its length rule gives T back
    import numpy as np
    import memd_amd
    import synth
    y = synth.load_eeg_sample(os.path.join(ROOT, "tests", "golden", "eeg_sample_32x30504.npz"))
    assert y.shape == (32, 30504) and y.dtype == np.float32
    p = tmp_path / "EEGsample_32CH_30504.bin"
    memd_amd.write_bin(p, y)
    assert p.stat().st_size == 3904512
    assert memd_amd.bin_length(p, 32) == 30504
    assert np.array_equal(memd_amd.read_bin(p, 32), y)
    assert np.array_equal(memd_amd.read_bin(p, 32, 2049), y[:, :2049])


def test_cli_error_paths(tmp_path):
    # the demo driver reports failures as status codes (the reference exits with -1)
    import subprocess
    cli = os.path.join(ROOT, "memd-gpu_amd", "memd_cli")
    r = subprocess.run([cli, "--input", str(tmp_path / "none.bin")], capture_output=True, text=True)
    assert r.returncode == 8 and "cannot open" in r.stderr
    r = subprocess.run([cli, "--no-such-flag", "1"], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
