"""Python mirror of the MEMD engine's host interface (ctypes over libmemd_amd.so).

The names and argument meaning follow the reference's host API:

* ``GenerateDirectionVector(SignalDim, NumDirVector, d_DirectionVector)``
  (memReducedEpochs_dev.cu:1029-1072)
* ``memd(SignalDim, SignalLength, NumDirVector, num_IMFs, max_iter, d_x, d_y,
  d_DirectionVectors, d_IMFs)`` returning elapsed seconds
  (memReducedEpochs_dev.cu:1074-1075, 1367)
* ``memdMatlabCuda(signalDim, numDirVector, numIMFs, maxiter, data, samplesIndex,
  primesSet, outputData)`` (memdMatlabCuda.cu:996-1039, assistMatlabCuda.m:17)

Device buffers are torch tensors on a ROCm device (torch is only the allocator
here).  Every call goes through the C ABI declared in include/memd_amd.h; there
is no CPU fallback: if the shared library is missing, loading raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmemd_amd.so")

MEMD_OK = 0
STATUS_NAMES = {0: "MEMD_OK", 1: "MEMD_ERR_ARG", 2: "MEMD_ERR_HIP", 3: "MEMD_ERR_NOMEM",
                4: "MEMD_ERR_COORDS", 5: "MEMD_ERR_UNSUPPORTED", 6: "MEMD_ERR_ADMISSION",
                7: "MEMD_ERR_NOT_FOUND", 8: "MEMD_ERR_IO"}
MODE_FIX, MODE_STOP = 0, 1
WINDOW_REFWARP32, WINDOW_FULL = 0, 1
SPLINE_NATURAL, SPLINE_NOT_A_KNOT = 0, 1
BOUNDARY_CLAMP, BOUNDARY_MIRROR = 0, 1

_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_i32 = ctypes.c_int32


class MemdOpts(ctypes.Structure):
    _fields_ = [("mode", _i32), ("window", _i32), ("sd", ctypes.c_double), ("sd2", ctypes.c_double),
                ("tol", ctypes.c_double), ("use_graph", _i32), ("d_knot_counts", _vp),
                ("max_passes", _i32), ("h_sift_counts", ctypes.POINTER(_i32)),
                ("h_n_imf_out", ctypes.POINTER(_i32)), ("spline", _i32), ("boundary", _i32)]


class MemdServerInfo(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("probes", ctypes.c_int64), ("rebuilds", ctypes.c_int64),
                ("retired", ctypes.c_int64), ("withdrawn", ctypes.c_int64), ("deadlines", ctypes.c_int64),
                ("rebuild_s", ctypes.c_double)]


class MemdEpoch(ctypes.Structure):
    _fields_ = [("n", _sz), ("T", _sz), ("K", _sz), ("N_imf", _sz), ("S", _sz),
                ("d_y", _vp), ("d_V", _vp), ("d_IMFs", _vp),
                ("h_sift_counts", ctypes.POINTER(_i32)), ("h_n_imf_out", ctypes.POINTER(_i32))]


class MemdKernelProfile(ctypes.Structure):
    _fields_ = [("ms_per_launch", ctypes.c_double * 10), ("launches", ctypes.c_int64 * 10),
                ("steps", ctypes.c_int64), ("knots", ctypes.c_int64), ("knot_channels", ctypes.c_int64),
                ("eval_points", ctypes.c_int64), ("samples", ctypes.c_int64)]


class MemdTuning(ctypes.Structure):
    _fields_ = [("n", _i32), ("block_groups", _i32), ("eval_tile_words", _i32), ("reserved", _i32),
                ("block_ms", ctypes.c_double), ("eval_ms", ctypes.c_double)]

    def as_dict(self):
        return {"n": self.n, "block_groups": self.block_groups, "eval_tile_words": self.eval_tile_words,
                "block_ms": self.block_ms, "eval_ms": self.eval_ms}


#: kernel order of MemdKernelProfile.ms_per_launch
KERNEL_NAMES = ("k_flags", "k_compact", "k_factor", "k_block", "k_reduced", "k_coef", "k_eval", "k_advance",
                "k_stop_apply", "k_init")


class MemdError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


#: every symbol include/memd_amd.h declares (checked by the CPU test-suite)
EXPORTS = ("memd_ctx_create", "memd_ctx_destroy", "memd_ctx_set_stream", "memd_last_error",
           "memd_opts_default", "memd_version", "memd_directions_f64", "memd_directions_f32", "memd_run_f64",
           "memd_run_f32", "memd_run_noalloc_f64",
           "memd_mex_compat_f32", "memd_mex_compat_f64", "memd_run_batch_f64", "memd_profile_batch_f64",
           "memd_workspace_bytes", "memd_workspace_bytes_opts", "memd_server_create", "memd_server_destroy", "memd_server_submit",
           "memd_server_step", "memd_server_completed", "memd_server_poll", "memd_server_release",
           "memd_server_pool_used", "memd_bin_length", "memd_read_bin_f32", "memd_write_bin_f32",
           "memd_autotune", "memd_ctx_set_tuning", "memd_ctx_get_tuning", "memd_server_cancel",
           "memd_server_set_deadline", "memd_server_get_info", "memd_ctx_set_direction_cache",
           "memd_ctx_direction_cache_bytes", "memd_debug_fail_allocations")
#: exports that do not return a memd_status
_NOT_STATUS = ("memd_last_error", "memd_version", "memd_opts_default", "memd_workspace_bytes",
               "memd_workspace_bytes_opts", "memd_server_pool_used", "memd_ctx_direction_cache_bytes")
REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_CANCELLED, REQ_TIMED_OUT = 0, 1, 2, 3, 4
#: MEMD_NOALLOC_SCRATCH_BYTES: the per-device workspace of memd_run_noalloc_f64
NOALLOC_SCRATCH_BYTES = 128 * 1024 * 1024

_lib = None


def load_library(path=LIB_PATH):
    """Load libmemd_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `make -C memd-gpu_amd` "
                          "(or __graft_entry__.build()); there is no fallback path")
    # torch ships its own HIP runtime (file libamdhip64.so, soname libamdhip64.so.7)
    # and loads it by file name.  Importing torch first makes the library's
    # libamdhip64.so.7 dependency resolve to that same runtime; loading the
    # library first would put a second HIP runtime into the process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(path)
    L.memd_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(_vp)]
    L.memd_ctx_destroy.argtypes = [_vp]
    L.memd_ctx_set_stream.argtypes = [_vp, _vp]
    L.memd_last_error.argtypes = [_vp]
    L.memd_last_error.restype = ctypes.c_char_p
    L.memd_version.restype = ctypes.c_char_p
    L.memd_opts_default.argtypes = [ctypes.POINTER(MemdOpts)]
    L.memd_opts_default.restype = None
    L.memd_workspace_bytes.argtypes = [_sz, _sz, _sz]
    L.memd_workspace_bytes.restype = _sz
    L.memd_workspace_bytes_opts.argtypes = [_sz, _sz, _sz, _sz, ctypes.POINTER(MemdOpts)]
    L.memd_workspace_bytes_opts.restype = _sz
    L.memd_directions_f64.argtypes = [_vp, _sz, _sz, ctypes.POINTER(_i32), _vp]
    L.memd_directions_f32.argtypes = L.memd_directions_f64.argtypes
    L.memd_run_f64.argtypes = [_vp, _sz, _sz, _sz, _sz, _sz, _vp, _vp, _vp, _vp,
                               ctypes.POINTER(MemdOpts), ctypes.POINTER(ctypes.c_double)]
    L.memd_run_f32.argtypes = L.memd_run_f64.argtypes
    L.memd_run_noalloc_f64.argtypes = L.memd_run_f64.argtypes
    L.memd_mex_compat_f32.argtypes = [_vp, _i32, _i32, _i32, _i32, _sz, _vp, _vp, _vp, _vp]
    L.memd_mex_compat_f64.argtypes = L.memd_mex_compat_f32.argtypes
    L.memd_run_batch_f64.argtypes = [_vp, ctypes.POINTER(MemdEpoch), _sz, ctypes.POINTER(MemdOpts),
                                     ctypes.POINTER(ctypes.c_double)]
    L.memd_profile_batch_f64.argtypes = [_vp, ctypes.POINTER(MemdEpoch), _sz, ctypes.POINTER(MemdOpts),
                                         ctypes.POINTER(MemdKernelProfile)]
    L.memd_server_create.argtypes = [_vp, _sz, _i32, ctypes.POINTER(MemdOpts), ctypes.POINTER(_vp)]
    L.memd_server_destroy.argtypes = [_vp]
    L.memd_server_submit.argtypes = [_vp, ctypes.POINTER(MemdEpoch), ctypes.POINTER(ctypes.c_int64)]
    L.memd_server_step.argtypes = [_vp, _i32, ctypes.POINTER(_i32), ctypes.POINTER(_i32)]
    L.memd_server_completed.argtypes = [_vp, ctypes.POINTER(ctypes.c_int64), _i32, ctypes.POINTER(_i32)]
    L.memd_server_poll.argtypes = [_vp, ctypes.c_int64, ctypes.POINTER(_i32)]
    L.memd_server_release.argtypes = [_vp, ctypes.c_int64]
    L.memd_server_pool_used.argtypes = [_vp]
    L.memd_server_pool_used.restype = _sz
    L.memd_server_cancel.argtypes = [_vp, ctypes.c_int64]
    L.memd_server_set_deadline.argtypes = [_vp, ctypes.c_int64, ctypes.c_double]
    L.memd_server_get_info.argtypes = [_vp, ctypes.POINTER(MemdServerInfo)]
    L.memd_ctx_set_direction_cache.argtypes = [_vp, _sz]
    L.memd_ctx_direction_cache_bytes.argtypes = [_vp]
    L.memd_ctx_direction_cache_bytes.restype = _sz
    L.memd_debug_fail_allocations.argtypes = [_vp, _i32]
    L.memd_bin_length.argtypes = [ctypes.c_char_p, _sz, ctypes.POINTER(_sz)]
    L.memd_read_bin_f32.argtypes = [ctypes.c_char_p, _sz, _sz, _sz, _vp]
    L.memd_write_bin_f32.argtypes = [ctypes.c_char_p, _vp, _sz]
    L.memd_autotune.argtypes = [_vp, _sz, _sz, _sz, ctypes.POINTER(MemdTuning)]
    L.memd_ctx_set_tuning.argtypes = [_vp, ctypes.POINTER(MemdTuning)]
    L.memd_ctx_get_tuning.argtypes = [_vp, _i32, ctypes.POINTER(MemdTuning)]
    for name in EXPORTS:
        if name not in _NOT_STATUS:
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def default_opts(mode=MODE_FIX, window=WINDOW_REFWARP32, sd=0.075, sd2=0.75, tol=0.075, use_graph=True,
                 spline=SPLINE_NATURAL, boundary=BOUNDARY_CLAMP):
    o = MemdOpts()
    load_library().memd_opts_default(ctypes.byref(o))
    o.mode, o.window, o.sd, o.sd2, o.tol, o.use_graph = mode, window, sd, sd2, tol, int(bool(use_graph))
    o.spline = spline
    o.boundary = boundary
    return o


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _shape(t):
    return tuple(int(v) for v in t.shape)


def check_epoch_shapes(y, V, n_imf, S, out=None, x=None):
    """Shapes of one epoch's buffers, checked before the C ABI sees them (it only
    receives pointers and sizes, so a mis-shaped buffer would be read or written
    past its end on the device): y [n][T], V [K][n], out [n_imf][n][T], x [T].
    Returns (n, T, K)."""
    if len(y.shape) != 2:
        raise ValueError(f"d_y must be [n][T], got shape {_shape(y)}")
    n, T = _shape(y)
    if len(V.shape) != 2 or int(V.shape[1]) != n:
        raise ValueError(f"d_DirectionVectors must be [K][{n}], got shape {_shape(V)}")
    K = int(V.shape[0])
    if int(n_imf) < 1 or int(S) < 1:
        raise ValueError(f"n_imf and S must be >= 1 (got {n_imf}, {S})")
    if out is not None and _shape(out) != (int(n_imf), n, T):
        raise ValueError(f"d_IMFs must be [{n_imf}][{n}][{T}], got shape {_shape(out)}")
    if x is not None and (len(x.shape) != 1 or int(x.shape[0]) != T):
        raise ValueError(f"d_x must be [{T}], got shape {_shape(x)}")
    return n, T, K


def _require_device(t, dtype, name):
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a device tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


class MemdContext:
    """One engine context (stream binding, cached workspace, direction cache)."""

    def __init__(self, device=0):
        self.lib = load_library()
        h = ctypes.c_void_p()
        rc = self.lib.memd_ctx_create(device, ctypes.byref(h))
        if rc != MEMD_OK:
            raise MemdError(rc, "memd_ctx_create failed: " + self.lib.memd_last_error(None).decode())
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            self.lib.memd_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != MEMD_OK:
            raise MemdError(rc, self.lib.memd_last_error(self.h).decode())

    def set_stream(self, stream):
        """Bind a torch.cuda.Stream (or None for the legacy default stream)."""
        self._check(self.lib.memd_ctx_set_stream(self.h, ctypes.c_void_p(stream.cuda_stream) if stream else None))

    # ---- direction cache and the allocation-failure hook (no reference counterpart)
    def set_direction_cache(self, max_bytes):
        self._check(self.lib.memd_ctx_set_direction_cache(self.h, int(max_bytes)))

    def direction_cache_bytes(self):
        return int(self.lib.memd_ctx_direction_cache_bytes(self.h))

    def fail_allocations(self, count):
        """Make the next `count` device allocations of this context fail (MEMD_ERR_NOMEM)."""
        self._check(self.lib.memd_debug_fail_allocations(self.h, int(count)))

    # ---- tile shapes (no reference counterpart; never change results)
    def autotune(self, n, T, K):
        """Time the candidate tile shapes for (n, T, K) epochs on this device and
        install the fastest for n; returns the choice as a dict."""
        t = MemdTuning()
        self._check(self.lib.memd_autotune(self.h, n, T, K, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, n, block_groups, eval_tile_words):
        t = MemdTuning(n=n, block_groups=block_groups, eval_tile_words=eval_tile_words)
        self._check(self.lib.memd_ctx_set_tuning(self.h, ctypes.byref(t)))

    def get_tuning(self, n):
        t = MemdTuning()
        self._check(self.lib.memd_ctx_get_tuning(self.h, n, ctypes.byref(t)))
        return t.as_dict()

    # ---- GenerateDirectionVector (memReducedEpochs_dev.cu:1029-1072)
    def directions(self, n, K, primes=None, out=None):
        import torch
        if out is None:
            out = torch.empty((K, n), dtype=torch.float64, device=f"cuda:{self.device}")
        f32 = out.dtype == torch.float32
        _require_device(out, torch.float32 if f32 else torch.float64, "d_DirectionVector")
        p = None
        if primes is not None:
            arr = (ctypes.c_int32 * n)(*[int(v) for v in primes])
            p = ctypes.cast(arr, ctypes.POINTER(ctypes.c_int32))
        fn = self.lib.memd_directions_f32 if f32 else self.lib.memd_directions_f64
        self._check(fn(self.h, n, K, p, _ptr(out)))
        return out

    # ---- memd (memReducedEpochs_dev.cu:1074-1368)
    def run(self, y, V, n_imf, S, x=None, out=None, mode=MODE_FIX, window=WINDOW_REFWARP32,
            sd=0.075, sd2=0.75, tol=0.075, use_graph=True, trace_passes=0, spline=SPLINE_NATURAL,
            noalloc=False, boundary=BOUNDARY_CLAMP):
        """Decompose one epoch y [n][T] (float64 or float32 device tensor).

        Returns (imfs [n_imf][n][T], info) with info = elapsed_s, sift_counts,
        n_imf_out and, when trace_passes > 0, knot_counts [passes][K][2].
        noalloc=True calls memd_run_noalloc_f64 (float64, out given)."""
        import torch
        n, T, K = check_epoch_shapes(y, V, n_imf, S, out, x)
        f32 = y.dtype == torch.float32
        if noalloc and (f32 or out is None or trace_passes):
            raise ValueError("the allocation-free entry point takes float64 buffers, a preallocated out "
                             "and no knot-count trace buffer")
        dt = torch.float32 if f32 else torch.float64
        _require_device(y, dt, "d_y")
        _require_device(V, dt, "d_DirectionVectors")
        if out is None:
            out = torch.empty((n_imf, n, T), dtype=dt, device=y.device)
        _require_device(out, dt, "d_IMFs")
        if x is not None:
            _require_device(x, torch.int32, "d_x")
        o = default_opts(mode, window, sd, sd2, tol, use_graph, spline, boundary)
        sc = (ctypes.c_int32 * n_imf)()
        nout = ctypes.c_int32(0)
        o.h_sift_counts = ctypes.cast(sc, ctypes.POINTER(ctypes.c_int32))
        o.h_n_imf_out = ctypes.pointer(nout)
        kc = None
        if trace_passes:
            kc = torch.zeros((trace_passes, K, 2), dtype=torch.int32, device=y.device)
            o.d_knot_counts = kc.data_ptr()
            o.max_passes = trace_passes
        el = ctypes.c_double(0.0)
        fn = self.lib.memd_run_f32 if f32 else self.lib.memd_run_f64
        if noalloc:
            fn = self.lib.memd_run_noalloc_f64
        self._check(fn(self.h, n, T, K, n_imf, S, _ptr(x), _ptr(y), _ptr(V), _ptr(out), ctypes.byref(o),
                       ctypes.byref(el)))
        info = {"elapsed_s": el.value, "sift_counts": np.array(sc[:], np.int32), "n_imf_out": nout.value}
        if kc is not None:
            info["knot_counts"] = kc
        return out, info

    # ---- mexFunction argument layout (memdMatlabCuda.cu:996-1039)
    def mex_compat(self, signalDim, numDirVector, numIMFs, maxiter, data, samplesIndex, primesSet, outputData):
        import torch
        n, N_imf = int(signalDim), int(numIMFs)
        if n < 1 or data.numel() % n != 0:
            raise ValueError(f"data holds {data.numel()} values, not a multiple of signalDim = {n}")
        T = data.numel() // n   # memdMatlabCuda.cu:1038-1039
        if outputData.numel() != N_imf * n * T:
            raise ValueError(f"outputData must hold numIMFs * signalDim * T = {N_imf * n * T} values, "
                             f"got {outputData.numel()}")
        if primesSet.numel() != n:
            raise ValueError(f"primesSet must hold signalDim = {n} values, got {primesSet.numel()}")
        if samplesIndex is not None and samplesIndex.numel() != T:
            raise ValueError(f"samplesIndex must hold T = {T} values, got {samplesIndex.numel()}")
        f32 = data.dtype == torch.float32
        dt = torch.float32 if f32 else torch.float64
        _require_device(data, dt, "data")
        _require_device(outputData, dt, "outputData")
        _require_device(primesSet, torch.int32, "primesSet")
        if samplesIndex is not None:
            _require_device(samplesIndex, torch.int32, "samplesIndex")
        fn = self.lib.memd_mex_compat_f32 if f32 else self.lib.memd_mex_compat_f64
        self._check(fn(self.h, int(signalDim), int(numDirVector), int(numIMFs), int(maxiter), data.numel(),
                       _ptr(data), _ptr(samplesIndex), _ptr(primesSet), _ptr(outputData)))
        return outputData

    # ---- batched epochs (no reference counterpart)
    def run_batch(self, epochs, mode=MODE_FIX, window=WINDOW_REFWARP32, use_graph=True, profile=False,
                  spline=SPLINE_NATURAL, boundary=BOUNDARY_CLAMP, sd=0.075, sd2=0.75, tol=0.075):
        """epochs: list of dicts with y, V, n_imf, S (and optional out).
        Returns (list of imfs tensors, info); with profile=True info["profile"]
        holds per-kernel average launch times (ms) and the work counters."""
        import torch
        arr = (MemdEpoch * len(epochs))()
        outs, scs, nouts = [], [], []
        for i, ep in enumerate(epochs):
            y, V = ep["y"], ep["V"]
            n, T, K = check_epoch_shapes(y, V, ep["n_imf"], ep["S"], ep.get("out"))
            _require_device(y, torch.float64, "d_y")
            _require_device(V, torch.float64, "d_V")
            out = ep.get("out")
            if out is None:
                out = torch.empty((ep["n_imf"], n, T), dtype=torch.float64, device=y.device)
            sc = (ctypes.c_int32 * ep["n_imf"])()
            no = ctypes.c_int32(0)
            arr[i] = MemdEpoch(n, T, K, ep["n_imf"], ep["S"], y.data_ptr(), V.data_ptr(),
                               out.data_ptr(), ctypes.cast(sc, ctypes.POINTER(ctypes.c_int32)),
                               ctypes.pointer(no))
            outs.append(out)
            scs.append(sc)
            nouts.append(no)
        o = default_opts(mode, window, sd, sd2, tol, use_graph=use_graph, spline=spline, boundary=boundary)
        el = ctypes.c_double(0.0)
        if profile:
            pr = MemdKernelProfile()
            self._check(self.lib.memd_profile_batch_f64(self.h, arr, len(epochs), ctypes.byref(o), ctypes.byref(pr)))
        else:
            self._check(self.lib.memd_run_batch_f64(self.h, arr, len(epochs), ctypes.byref(o), ctypes.byref(el)))
        info = {"elapsed_s": el.value, "sift_counts": [np.array(s[:], np.int32) for s in scs],
                "n_imf_out": [v.value for v in nouts]}
        if profile:
            info["profile"] = {
                "ms_per_launch": dict(zip(KERNEL_NAMES, pr.ms_per_launch[:])),
                "launches": dict(zip(KERNEL_NAMES, pr.launches[:])),
                "steps": pr.steps, "knots": pr.knots, "knot_channels": pr.knot_channels,
                "eval_points": pr.eval_points, "samples": pr.samples}
        return outs, info


class MemdServer:
    """Served continuous batching over one context (no reference counterpart).

    submit() queues an epoch and returns its id; step() admits what fits the
    HBM pool, advances the active set and retires finished epochs; completed()
    returns the ids finished since the last call; result() gives the IMFs and
    sift counts of a finished request."""

    def __init__(self, ctx, pool_bytes, max_active=0, mode=MODE_FIX, window=WINDOW_REFWARP32,
                 sd=0.075, sd2=0.75, tol=0.075, spline=SPLINE_NATURAL, boundary=BOUNDARY_CLAMP):
        self.ctx = ctx
        self.lib = ctx.lib
        o = default_opts(mode, window, sd, sd2, tol, spline=spline, boundary=boundary)
        h = ctypes.c_void_p()
        ctx._check(self.lib.memd_server_create(ctx.h, int(pool_bytes), int(max_active), ctypes.byref(o),
                                               ctypes.byref(h)))
        self.h = h
        self._req = {}

    def close(self):
        if getattr(self, "h", None):
            self.lib.memd_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, y, V, n_imf, S, out=None):
        import torch
        n, T, K = check_epoch_shapes(y, V, n_imf, S, out)
        _require_device(y, torch.float64, "d_y")
        _require_device(V, torch.float64, "d_V")
        if out is None:
            out = torch.empty((n_imf, n, T), dtype=torch.float64, device=y.device)
        _require_device(out, torch.float64, "d_IMFs")
        sc = (ctypes.c_int32 * n_imf)()
        no = ctypes.c_int32(0)
        ep = MemdEpoch(n, T, K, n_imf, S, y.data_ptr(), V.data_ptr(), out.data_ptr(),
                       ctypes.cast(sc, ctypes.POINTER(ctypes.c_int32)), ctypes.pointer(no))
        rid = ctypes.c_int64(0)
        self.ctx._check(self.lib.memd_server_submit(self.h, ctypes.byref(ep), ctypes.byref(rid)))
        # keep the buffers alive until the request is released
        self._req[rid.value] = (y, V, out, sc, no)
        return rid.value

    def step(self, steps=1):
        a, q = ctypes.c_int32(0), ctypes.c_int32(0)
        self.ctx._check(self.lib.memd_server_step(self.h, int(steps), ctypes.byref(a), ctypes.byref(q)))
        return a.value, q.value

    def completed(self, cap=4096):
        ids = (ctypes.c_int64 * cap)()
        n = ctypes.c_int32(0)
        self.ctx._check(self.lib.memd_server_completed(self.h, ids, cap, ctypes.byref(n)))
        return list(ids[:n.value])

    def poll(self, rid):
        st = ctypes.c_int32(0)
        self.ctx._check(self.lib.memd_server_poll(self.h, int(rid), ctypes.byref(st)))
        return st.value

    def result(self, rid):
        if self.poll(rid) != REQ_DONE:
            raise MemdError(1, f"request {rid} is not finished")
        _, _, out, sc, no = self._req[rid]
        return out, {"sift_counts": np.array(sc[:], np.int32), "n_imf_out": no.value}

    def release(self, rid):
        self.ctx._check(self.lib.memd_server_release(self.h, int(rid)))
        self._req.pop(rid, None)

    def pool_used(self):
        return int(self.lib.memd_server_pool_used(self.h))

    def cancel(self, rid):
        self.ctx._check(self.lib.memd_server_cancel(self.h, int(rid)))

    def set_deadline(self, rid, seconds):
        self.ctx._check(self.lib.memd_server_set_deadline(self.h, int(rid), float(seconds)))

    def info(self):
        i = MemdServerInfo()
        self.ctx._check(self.lib.memd_server_get_info(self.h, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in MemdServerInfo._fields_}


# ---- raw sample files (memReducedEpochs_dev.cu:896-932, used at :1484-1501) ----

def _io_check(rc):
    if rc != MEMD_OK:
        raise MemdError(rc, load_library().memd_last_error(None).decode())


def bin_length(path, n):
    """Samples per channel of a float32 [n][T] file: bytes / 4 / n, floored (:1486)."""
    T = _sz(0)
    _io_check(load_library().memd_bin_length(os.fsencode(path), int(n), ctypes.byref(T)))
    return T.value


def read_bin(path, n, T=None):
    """float32 array [n][T]: the first T samples of every channel (default: all)."""
    T_file = bin_length(path, n)
    T = T_file if T is None else int(T)
    out = np.empty((n, T), np.float32)
    _io_check(load_library().memd_read_bin_f32(os.fsencode(path), int(n), T_file, T,
                                              out.ctypes.data_as(ctypes.c_void_p)))
    return out


def write_bin(path, arr):
    """Write a float32 array as raw bytes in C order (writeBin)."""
    a = np.ascontiguousarray(arr, dtype=np.float32)
    _io_check(load_library().memd_write_bin_f32(os.fsencode(path), a.ctypes.data_as(ctypes.c_void_p), a.size))


# ---- module-level functions with the reference's names -----------------------

_default_ctx = None


def _ctx():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = MemdContext(0)
    return _default_ctx


def GenerateDirectionVector(SignalDim, NumDirVector, d_DirectionVector):
    """memReducedEpochs_dev.cu:1029-1072: fill d_DirectionVector [K][n] in place."""
    _ctx().directions(SignalDim, NumDirVector, out=d_DirectionVector)


def memd(SignalDim, SignalLength, NumDirVector, num_IMFs, max_iter, d_x, d_y, d_DirectionVectors, d_IMFs):
    """memReducedEpochs_dev.cu:1074-1075: decompose d_y into d_IMFs; returns seconds."""
    if tuple(d_y.shape) != (SignalDim, SignalLength):
        raise ValueError("d_y must be [SignalDim][SignalLength]")
    if tuple(d_DirectionVectors.shape) != (NumDirVector, SignalDim):
        raise ValueError("d_DirectionVectors must be [NumDirVector][SignalDim]")
    _, info = _ctx().run(d_y, d_DirectionVectors, num_IMFs, max_iter, x=d_x, out=d_IMFs)
    return info["elapsed_s"]


def memdMatlabCuda(signalDim, numDirVector, numIMFs, maxiter, data, samplesIndex, primesSet, outputData):
    """memdMatlabCuda.cu:996-1039: returns the IMF array (memory [N_imf][n][T])."""
    return _ctx().mex_compat(signalDim, numDirVector, numIMFs, maxiter, data, samplesIndex, primesSet, outputData)
