"""Writes tests/golden/golden_small.npz: independent numpy/scipy answers that pin the oracle.

Run once in a container with numpy + scipy (it does not read the reference):

    python tests/golden/make_golden.py

Each block restates a piece of the reference by a DIFFERENT route from the
oracle, so that a plausible slip in the oracle (dropped term, wrong sign or
index, transposed operand) disagrees with at least one vector:

* directions: closed form v = (b_0..b_{n-2}, |b_{n-1}|)/|b| with b = 2h - 1 and
  the Hammersley h from exact integer digit expansion (fractions), instead of the
  atan2/sin/cos chain of Sequence2Vector (memReducedEpochs_dev.cu:948-990).
* extrema window: a lane-by-lane simulation of find_extrema_shfl_max/min
  (memReducedEpochs_dev.cu:77-141) on the launch grid of :1191-1193, instead of
  the closed-form limit 240*ceil(T/242).
* end-knot clamp + spline envelope: np.polyfit line through knots 1,2 (and
  m-1,m-2) and scipy.interpolate.CubicSpline(bc_type='natural') evaluated at the
  integer samples, instead of the blocked Thomas solve + Horner.
* tridiagonal: dense np.linalg.solve of the natural-spline system.
* one envelope-mean pass of a small epoch: the whole chain above with V @ y.
"""
import os
from fractions import Fraction

import numpy as np
from scipy.interpolate import CubicSpline

HERE = os.path.dirname(os.path.abspath(__file__))


def primes_upto(count):
    out, c = [], 2
    while len(out) < count:
        if all(c % p for p in out if p * p <= c):
            out.append(c)
        c += 1
    return out


def radinv_exact(base, i):
    f, denom = Fraction(0), 1
    while i:
        denom *= base
        f += Fraction(i % base, denom)
        i //= base
    return float(f)


def directions_closed_form(n, K, P):
    V = np.zeros((K, n))
    for k in range(K):
        h = np.array([(k + 1.5) / (-P[0])] + [radinv_exact(P[j], k + 1) for j in range(1, n)])
        b = 2.0 * h - 1.0
        if n == 1:
            V[k] = [1.0]
            continue
        v = b.copy()
        v[-1] = abs(b[-1])
        V[k] = v / np.linalg.norm(b)
    return V


def simulate_extrema_kernel(P, env):
    """Lane-level model of find_extrema_shfl_max/min: 256-thread blocks, 32-lane
    warps, sample e = g - 2*(g // 32), shuffle neighbours clamp at lanes 0/31."""
    T = len(P)
    blocks = T // 242 + (0 if T % 242 == 0 else 1)
    flags = np.zeros(T, bool)
    for g0 in range(0, blocks * 256, 32):          # one warp at a time
        lanes = np.arange(32)
        g = g0 + lanes
        e = g - 2 * (g // 32)
        active = e < T
        vals = np.where(active, P[np.minimum(e, T - 1)], np.nan)
        up = np.where(lanes > 0, np.roll(vals, 1), vals)
        down = np.where(lanes < 31, np.roll(vals, -1), vals)
        with np.errstate(invalid="ignore"):
            hit = (vals > up) & (vals > down) if env == 0 else (vals < up) & (vals < down)
        for ln in range(32):
            if active[ln] and hit[ln]:
                flags[e[ln]] = True
            if active[ln] and g[ln] == 0:
                flags[0] = True
                flags[T - 1] = True
    return np.flatnonzero(flags).astype(np.int32)


def clamp_polyfit(X, Y, env):
    Y = Y.astype(np.float64).copy()
    E = len(X)
    if E < 6:
        return Y
    left = np.polyval(np.polyfit(X[1:3].astype(float), Y[1:3], 1), float(X[0]))
    m = E - 1
    right = np.polyval(np.polyfit(X[m - 2:m].astype(float), Y[m - 2:m], 1), float(X[m]))
    if (left > Y[0]) if env == 0 else (left < Y[0]):
        Y[0] = left
    if (right > Y[m]) if env == 0 else (right < Y[m]):
        Y[m] = right
    return Y


def spline_env(T, X, Y):
    return CubicSpline(X.astype(float), Y, bc_type="natural")(np.arange(T, dtype=float))


def envelope_mean_reference(y, V):
    n, T = y.shape
    K = V.shape[0]
    P = V @ y
    acc = np.zeros((n, T))
    nv = 0
    counts = np.zeros((K, 2), np.int32)
    for k in range(K):
        Xs = [simulate_extrema_kernel(P[k], 0), simulate_extrema_kernel(P[k], 1)]
        counts[k] = [len(Xs[0]), len(Xs[1])]
        if not (len(Xs[0]) > 3 and len(Xs[1]) > 3):
            continue
        nv += 1
        for c in range(n):
            envs = []
            for e in range(2):
                Yc = clamp_polyfit(Xs[e], y[c, Xs[e]], e)
                envs.append(spline_env(T, Xs[e], Yc))
            acc[c] += (envs[0] + envs[1]) / 2.0
    return acc / nv, counts


def main():
    rng = np.random.default_rng(20240601)
    out = {}
    # directions
    for n, K in [(1, 4), (2, 8), (3, 5), (8, 16), (32, 64)]:
        P = [-K] + primes_upto(n - 1)
        out[f"dirs_ref_{n}_{K}"] = directions_closed_form(n, K, P)
    for n, K in [(3, 7), (32, 64)]:
        P = primes_upto(n)   # MEX path: assistMatlabCuda.m passes primes(2000)(1:n)
        out[f"dirs_mex_{n}_{K}"] = directions_closed_form(n, K, P)
        out[f"dirs_mex_{n}_{K}_primes"] = np.array(P, np.int32)
    # extrema window
    for T in [5, 100, 242, 483, 484, 1000, 4096, 8193]:
        P = np.round(rng.standard_normal(T), 1)          # rounding plants exact ties
        out[f"ext_P_{T}"] = P
        out[f"ext_max_{T}"] = simulate_extrema_kernel(P, 0)
        out[f"ext_min_{T}"] = simulate_extrema_kernel(P, 1)
    # clamp + spline envelope
    for case, (T, m) in enumerate([(40, 4), (60, 5), (200, 6), (500, 40), (3000, 400)]):
        X = np.sort(rng.choice(np.arange(1, T - 1), m - 2, replace=False))
        X = np.r_[0, X, T - 1].astype(np.int32)
        Y = rng.standard_normal(m) * 3.0
        out[f"spl_X_{case}"] = X
        out[f"spl_Y_{case}"] = Y
        out[f"spl_T_{case}"] = np.array(T)
        for e in range(2):
            Yc = clamp_polyfit(X, Y, e)
            out[f"spl_Yc_{case}_{e}"] = Yc
            out[f"spl_env_{case}_{e}"] = spline_env(T, X, Yc)
    # tridiagonal (natural-spline shaped systems)
    for N in [1, 2, 31, 32, 33, 63, 64, 65, 200, 1000]:
        h = rng.integers(1, 9, N + 1).astype(float)
        lo = np.r_[0.0, h[:N], 0.0]
        up = np.r_[0.0, h[1:N + 1], 0.0]
        di = np.r_[0.0, 2.0 * (h[:N] + h[1:N + 1]), 0.0]
        r = np.r_[0.0, rng.standard_normal(N), 0.0]
        A = np.diag(di[1:N + 1]) + np.diag(lo[2:N + 1], -1) + np.diag(up[1:N], 1)
        x = np.linalg.solve(A, r[1:N + 1])
        out[f"tri_{N}"] = np.stack([lo, di, up, r])
        out[f"tri_x_{N}"] = x
    # one envelope-mean pass of a small epoch
    n, T, K = 3, 300, 6
    t = np.arange(T)
    y = np.stack([np.sin(2 * np.pi * t / 37.0 + c) + 0.3 * rng.standard_normal(T) for c in range(n)])
    V = directions_closed_form(n, K, [-K] + primes_upto(n - 1))
    mean, counts = envelope_mean_reference(y, V)
    out["pass_y"] = y
    out["pass_V"] = V
    out["pass_mean"] = mean
    out["pass_counts"] = counts
    np.savez_compressed(os.path.join(HERE, "golden_small.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
