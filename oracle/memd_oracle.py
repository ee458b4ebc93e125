"""ctypes view of the FP64 oracle (``libmemd_oracle.so``) -- test infrastructure only.

The C++ restatement lives in ``memd_oracle.cpp``; every function there cites the
reference line it follows (memReducedEpochs_dev.cu / memdMatlabCuda.cu).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmemd_oracle.so")

MODE_FIX, MODE_STOP = 0, 1
WINDOW_REFWARP32, WINDOW_FULL = 0, 1
SPLINE_NATURAL, SPLINE_NOT_A_KNOT = 0, 1
BOUNDARY_CLAMP, BOUNDARY_MIRROR = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class Opts(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("window", ctypes.c_int),
                ("sd", ctypes.c_double), ("sd2", ctypes.c_double), ("tol", ctypes.c_double),
                ("threads", ctypes.c_int), ("spline", ctypes.c_int), ("boundary", ctypes.c_int)]


class Trace(ctypes.Structure):
    _fields_ = [("knot_counts", _ip), ("max_sifts", ctypes.c_int32), ("sift_counts", _ip),
                ("n_imf_out", _ip), ("n_passes", _ip)]


def build():
    """Compile the oracle with its Makefile (g++, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_directions.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _dp]
        L.oracle_radical_inverse.argtypes = [ctypes.c_int, ctypes.c_uint]
        L.oracle_radical_inverse.restype = ctypes.c_double
        L.oracle_prime_table.argtypes = [ctypes.c_int, ctypes.c_int, _ip]
        L.oracle_project.argtypes = [ctypes.c_int] * 3 + [_dp, _dp, _dp]
        L.oracle_extrema.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _ip]
        L.oracle_clamp.argtypes = [ctypes.c_int, _ip, _dp, ctypes.c_int]
        L.oracle_spline_solve.argtypes = [ctypes.c_int, _ip, ctypes.c_int, _dp, _dp]
        L.oracle_spline_solve_bc.argtypes = [ctypes.c_int, _ip, ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.oracle_spline_coefs.argtypes = [ctypes.c_int, _ip, _dp, _dp, _dp]
        L.oracle_envelope.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _dp, _dp]
        L.oracle_tridiag_blocked.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.oracle_envelope_mean.argtypes = [ctypes.c_int] * 3 + [_dp, _dp, ctypes.POINTER(Opts), _dp, _dp, _ip]
        L.oracle_memd.argtypes = [ctypes.c_int] * 5 + [_dp, _dp, _dp, ctypes.POINTER(Opts), ctypes.POINTER(Trace)]
        L.oracle_window_limit.argtypes = [ctypes.c_int, ctypes.c_int]
        _pp = ctypes.POINTER(ctypes.c_void_p)
        L.oracle_memd_batch.argtypes = [ctypes.c_int] + [_ip] * 5 + [_pp] * 3 + [ctypes.POINTER(Opts), ctypes.c_int]
        L.oracle_mirror_knots.argtypes = [ctypes.c_int, _dp, _ip, ctypes.c_int, _ip, ctypes.c_int] + [_ip] * 6
        L.oracle_amp_block_sum.argtypes = [_dp, ctypes.c_int]
        L.oracle_amp_block_sum.restype = ctypes.c_double
        L.oracle_stop_decision.argtypes = [ctypes.c_int] * 3 + [_dp, _dp, _ip] + [ctypes.c_double] * 3
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def make_opts(mode=MODE_FIX, window=WINDOW_REFWARP32, sd=0.075, sd2=0.75, tol=0.075, threads=1,
              spline=SPLINE_NATURAL, boundary=BOUNDARY_CLAMP):
    return Opts(mode, window, sd, sd2, tol, threads, spline, boundary)


def solver_block():
    return lib().oracle_solver_block()


def window_limit(T, window=WINDOW_REFWARP32):
    return lib().oracle_window_limit(T, window)


def radical_inverse(base, i):
    return lib().oracle_radical_inverse(base, i)


def prime_table(n, K):
    P = np.zeros(n, np.int32)
    if lib().oracle_prime_table(n, K, _i(P)) != 0:
        raise ValueError("bad n/K")
    return P


def directions(n, K, primes=None):
    V = np.zeros((K, n), np.float64)
    p = None
    if primes is not None:
        primes = np.ascontiguousarray(primes, np.int32)
        p = _i(primes)
    if lib().oracle_directions(n, K, p, _d(V)) != 0:
        raise ValueError("bad direction arguments")
    return V


def project(cur, V):
    cur = np.ascontiguousarray(cur, np.float64)
    V = np.ascontiguousarray(V, np.float64)
    n, T = cur.shape
    K = V.shape[0]
    P = np.zeros((K, T), np.float64)
    lib().oracle_project(n, T, K, _d(cur), _d(V), _d(P))
    return P


def extrema(Pk, env, window=WINDOW_REFWARP32):
    Pk = np.ascontiguousarray(Pk, np.float64)
    X = np.zeros(len(Pk), np.int32)
    E = lib().oracle_extrema(len(Pk), _d(Pk), env, window, _i(X))
    return X[:E].copy()


def mirror_knots(Pk, Xmax, Xmin):
    """Mirrored knot lists of one direction from its clamp-mode lists (ends included).

    Returns ((P_max, S_max), (P_min, S_min), (lsym, rsym)): knot positions, the
    samples whose values they carry, and the two mirror points."""
    Pk = np.ascontiguousarray(Pk, np.float64)
    Xmax = np.ascontiguousarray(Xmax, np.int32)
    Xmin = np.ascontiguousarray(Xmin, np.int32)
    out = [np.zeros(len(X) + 2, np.int32) for X in (Xmax, Xmax, Xmin, Xmin)]
    E = np.zeros(2, np.int32)
    sym = np.zeros(2, np.int32)
    rc = lib().oracle_mirror_knots(len(Pk), _d(Pk), _i(Xmax), len(Xmax), _i(Xmin), len(Xmin),
                                   *[_i(a) for a in out], _i(E), _i(sym))
    if rc != 0:
        raise ValueError("mirroring needs at least 4 knots per list")
    return ((out[0][:E[0]].copy(), out[1][:E[0]].copy()), (out[2][:E[1]].copy(), out[3][:E[1]].copy()),
            (int(sym[0]), int(sym[1])))


def clamp(X, Y, env):
    X = np.ascontiguousarray(X, np.int32)
    Y = np.array(Y, np.float64)
    lib().oracle_clamp(len(X), _i(X), _d(Y), env)
    return Y


def spline_solve(X, Y, spline=SPLINE_NATURAL):
    """Y: [nrhs][E] clamped knot values -> M: [nrhs][E] (natural: M[:,0] = M[:,-1] = 0)."""
    X = np.ascontiguousarray(X, np.int32)
    Y = np.ascontiguousarray(np.atleast_2d(Y), np.float64)
    M = np.zeros_like(Y)
    if spline == SPLINE_NATURAL:
        lib().oracle_spline_solve(len(X), _i(X), Y.shape[0], _d(Y), _d(M))
    else:
        lib().oracle_spline_solve_bc(len(X), _i(X), Y.shape[0], _d(Y), _d(M), spline)
    return M


def spline_coefs(X, Y, M):
    X = np.ascontiguousarray(X, np.int32)
    Y = np.ascontiguousarray(Y, np.float64)
    M = np.ascontiguousarray(M, np.float64)
    coef = np.zeros((len(X) - 1, 4), np.float64)
    lib().oracle_spline_coefs(len(X), _i(X), _d(Y), _d(M), _d(coef))
    return coef


def envelope(T, X, coef):
    X = np.ascontiguousarray(X, np.int32)
    coef = np.ascontiguousarray(coef, np.float64)
    env = np.zeros(T, np.float64)
    lib().oracle_envelope(T, len(X), _i(X), _d(coef), _d(env))
    return env


def tridiag_blocked(lo, di, up, r):
    """Rows 1..N of length-(N+2) arrays; returns x with x[0] = x[N+1] = 0."""
    lo, di, up, r = (np.ascontiguousarray(a, np.float64) for a in (lo, di, up, r))
    N = len(di) - 2
    x = np.zeros(N + 2, np.float64)
    lib().oracle_tridiag_blocked(N, _d(lo), _d(di), _d(up), _d(r), _d(x))
    return x


def envelope_mean(cur, V, window=WINDOW_REFWARP32, want_amp=False, threads=1, spline=SPLINE_NATURAL,
                  boundary=BOUNDARY_CLAMP):
    cur = np.ascontiguousarray(cur, np.float64)
    V = np.ascontiguousarray(V, np.float64)
    n, T = cur.shape
    K = V.shape[0]
    mean = np.zeros((n, T), np.float64)
    amp = np.zeros(T, np.float64) if want_amp else None
    counts = np.zeros((K, 2), np.int32)
    o = make_opts(window=window, threads=threads, spline=spline, boundary=boundary)
    lib().oracle_envelope_mean(n, T, K, _d(cur), _d(V), ctypes.byref(o), _d(mean),
                               _d(amp) if want_amp else None, _i(counts))
    return mean, amp, counts


def stop_decision(mean, amp, counts, sd=0.075, sd2=0.75, tol=0.075):
    """Stop mode's decision on one pass: True = stop sifting this IMF."""
    mean = np.ascontiguousarray(mean, np.float64)
    amp = np.ascontiguousarray(amp, np.float64)
    counts = np.ascontiguousarray(counts, np.int32)
    n, T = mean.shape
    K = counts.shape[0]
    return bool(lib().oracle_stop_decision(n, T, K, _d(mean), _d(amp), _i(counts), sd, sd2, tol))


def memd(y, V, n_imf, S, mode=MODE_FIX, window=WINDOW_REFWARP32, sd=0.075, sd2=0.75, tol=0.075,
         threads=1, trace=True, spline=SPLINE_NATURAL, boundary=BOUNDARY_CLAMP):
    """Whole epoch. Returns (imfs [N_imf][n][T], info dict)."""
    y = np.ascontiguousarray(y, np.float64)
    V = np.ascontiguousarray(V, np.float64)
    n, T = y.shape
    K = V.shape[0]
    imfs = np.zeros((n_imf, n, T), np.float64)
    max_passes = (n_imf - 1) * (S + 1) + 1
    kc = np.zeros((max_passes, K, 2), np.int32) if trace else None
    sc = np.zeros(n_imf, np.int32)
    nout = np.zeros(1, np.int32)
    npass = np.zeros(1, np.int32)
    tr = Trace(_i(kc) if trace else None, max_passes, _i(sc), _i(nout), _i(npass))
    o = make_opts(mode, window, sd, sd2, tol, threads, spline, boundary)
    rc = lib().oracle_memd(n, T, K, n_imf, S, _d(y), _d(V), _d(imfs), ctypes.byref(o), ctypes.byref(tr))
    if rc != 0:
        raise ValueError("oracle_memd rejected its arguments")
    info = {"sift_counts": sc.copy(), "n_imf_out": int(nout[0]), "n_passes": int(npass[0])}
    if trace:
        info["knot_counts"] = kc[: int(npass[0])].copy()
    return imfs, info


def memd_batch(epochs, n_imf, S, threads=1, mode=MODE_FIX, window=WINDOW_REFWARP32, sd=0.075, sd2=0.75,
               tol=0.075, spline=SPLINE_NATURAL, boundary=BOUNDARY_CLAMP):
    """epochs: list of (y [n][T], V [K][n]).  Epochs run in parallel on `threads`
    threads, each serially; returns the list of imfs [n_imf][n][T]."""
    ys = [np.ascontiguousarray(y, np.float64) for y, _ in epochs]
    Vs = [np.ascontiguousarray(V, np.float64) for _, V in epochs]
    outs = [np.zeros((n_imf, y.shape[0], y.shape[1]), np.float64) for y in ys]
    c = len(ys)
    ia = lambda v: np.ascontiguousarray(v, np.int32)
    n, T, K = ia([y.shape[0] for y in ys]), ia([y.shape[1] for y in ys]), ia([V.shape[0] for V in Vs])
    nimf, ss = ia([n_imf] * c), ia([S] * c)
    pa = lambda arrs: (ctypes.c_void_p * c)(*[a.ctypes.data for a in arrs])
    o = make_opts(mode, window, sd, sd2, tol, 1, spline, boundary)
    rc = lib().oracle_memd_batch(c, _i(n), _i(T), _i(K), _i(nimf), _i(ss), pa(ys), pa(Vs), pa(outs),
                                 ctypes.byref(o), threads)
    if rc != 0:
        raise ValueError("oracle_memd_batch rejected its arguments")
    return outs


def amp_block_sum(d):
    """sum_c d[c]^2 in the engine's order (groups of 4, blocks of 16)."""
    d = np.ascontiguousarray(d, np.float64)
    return lib().oracle_amp_block_sum(_d(d), len(d))
