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


class _NoLibrary:
    """Stands in for the loaded library: any call means a buffer reached the C ABI."""

    def __getattr__(self, name):
        raise AssertionError(f"{name} was called with mis-shaped buffers")


def _offline_ctx():
    import memd_amd
    ctx = object.__new__(memd_amd.MemdContext)
    ctx.lib, ctx.h, ctx.device = _NoLibrary(), None, 0
    return ctx


@pytest.mark.parametrize("case", ["V_channels", "V_rank", "out_shape", "out_slots", "x_len", "y_rank", "n_imf"])
def test_run_refuses_mis_shaped_buffers(case):
    # the C ABI receives pointers and sizes only, so the mirror checks the shapes
    # before any library call (a mis-shaped buffer would be read or written past its end)
    import torch
    import memd_amd
    n, T, K = 4, 100, 8
    y, V = torch.zeros(n, T, dtype=torch.float64), torch.zeros(K, n, dtype=torch.float64)
    out, x, n_imf = torch.zeros(3, n, T, dtype=torch.float64), torch.arange(T, dtype=torch.int32), 3
    if case == "V_channels":
        V = torch.zeros(K, n + 1, dtype=torch.float64)
    elif case == "V_rank":
        V = torch.zeros(K * n, dtype=torch.float64)
    elif case == "out_shape":
        out = torch.zeros(3, n, T - 1, dtype=torch.float64)
    elif case == "out_slots":
        n_imf = 4
    elif case == "x_len":
        x = torch.arange(T - 1, dtype=torch.int32)
    elif case == "y_rank":
        y = torch.zeros(n * T, dtype=torch.float64)
    elif case == "n_imf":
        n_imf, out = 0, None
    with pytest.raises(ValueError):
        _offline_ctx().run(y, V, n_imf, 2, x=x, out=out)


def test_batch_and_server_refuse_mis_shaped_buffers():
    import torch
    import memd_amd
    y, V = torch.zeros(4, 100, dtype=torch.float64), torch.zeros(8, 5, dtype=torch.float64)
    with pytest.raises(ValueError):
        _offline_ctx().run_batch([{"y": y, "V": V, "n_imf": 3, "S": 2}])
    srv = object.__new__(memd_amd.MemdServer)
    srv.ctx, srv.lib, srv.h, srv._req = _offline_ctx(), _NoLibrary(), None, {}
    with pytest.raises(ValueError):
        srv.submit(y, V, 3, 2)
    with pytest.raises(ValueError):
        srv.submit(y, torch.zeros(8, 4, dtype=torch.float64), 3, 2, out=torch.zeros(2, 4, 100, dtype=torch.float64))


@pytest.mark.parametrize("case", ["numel", "out", "primes", "index"])
def test_mex_layout_refuses_mis_sized_buffers(case):
    import torch
    n, K, N_imf, S, T = 3, 6, 4, 2, 50
    data, out = torch.zeros(n * T, dtype=torch.float32), torch.zeros(N_imf * n * T, dtype=torch.float32)
    primes, index = torch.tensor([2, 3, 5], dtype=torch.int32), torch.arange(T, dtype=torch.int32)
    if case == "numel":
        data = torch.zeros(n * T + 1, dtype=torch.float32)
    elif case == "out":
        out = torch.zeros(N_imf * n * T - 1, dtype=torch.float32)
    elif case == "primes":
        primes = torch.tensor([2, 3], dtype=torch.int32)
    elif case == "index":
        index = torch.arange(T + 1, dtype=torch.int32)
    with pytest.raises(ValueError):
        _offline_ctx().mex_compat(n, K, N_imf, S, data, index, primes, out)


def test_workspace_bytes_follow_mode_and_slots():
    import memd_amd
    lib = memd_amd.load_library()
    fix = memd_amd.default_opts()
    stop = memd_amd.default_opts(mode=memd_amd.MODE_STOP)
    base = lib.memd_workspace_bytes(32, 4096, 64)
    assert lib.memd_workspace_bytes_opts(32, 4096, 64, 9, None) == base
    assert lib.memd_workspace_bytes_opts(32, 4096, 64, 64, ctypes.byref(fix)) == base
    assert lib.memd_workspace_bytes_opts(32, 4096, 64, 65, ctypes.byref(fix)) > base
    # stop mode: the mean [n][T] and the per-block amplitude sums [K][ceil(n/16)][T]
    assert lib.memd_workspace_bytes_opts(32, 4096, 64, 9, ctypes.byref(stop)) >= base + 64 * 2 * 4096 * 8
    # mirrored ends: the record array, 2K * n * (T/2 + 2) intervals of 32 bytes
    mirror = memd_amd.default_opts(boundary=memd_amd.BOUNDARY_MIRROR)
    assert lib.memd_workspace_bytes_opts(32, 4096, 64, 9, ctypes.byref(mirror)) >= base + 128 * 32 * 2050 * 32
    assert lib.memd_workspace_bytes_opts(32, 4096, 64, 0, None) == 0
