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
