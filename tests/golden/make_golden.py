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
* one envelope-mean pass of a small epoch: the whole chain above with V @ y,
  including memd.m's amplitude sum_valid sqrt(sum_c (e+ - e-)^2) / 2 / nvalid.
* whole decompositions: the fixed-count loop of memReducedEpochs_dev.cu:1174-1179,
  1328-1338 and the stop-mode loop of the public memd.m (stop(): sx = |mean|_2,
  divided by amp only when every amp(t) != 0, sift while (mean(sx > sd) > tol |
  any(sx > sd2)) & any(nem > 2); stop_emd(): no IMF is started once every
  direction has fewer than 3 interior extrema), written as plain Python loops over
  the stage restatements above.  The generator checks that no decision is within
  rounding of a threshold (sx near sd or sd2, a fraction near tol), so the
  oracle's different rounding cannot flip one.
* mirrored ends (memd.m's boundary_conditions, nbsym = 2; not in the reference):
  a transcription with MATLAB's 1-based indices, inclusive slices and
  min(end, .) / max(end - ., 1) bounds, instead of the oracle's 0-based
  reversed-range helper; the envelopes are scipy CubicSplines (natural or
  not-a-knot) through the mirrored knots, evaluated at the integer samples.
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


def boundary_conditions(indmin, indmax, x, nbsym=2):
    """memd.m's boundary_conditions for one direction, with 1-based sample indices
    as in MATLAB (t(i) = i).  indmin / indmax: 0-based interior extrema; x: the
    projection.  Returns 0-based (positions, sources) for the min and max lists
    and the mirror points (lsym, rsym), and the branch taken at each end."""
    lx = len(x)
    X = lambda i: x[i - 1]
    imin = [int(i) + 1 for i in indmin]
    imax = [int(i) + 1 for i in indmax]
    fl = lambda v: v[::-1]

    def sl(v, a, b):   # MATLAB v(a:b), inclusive, empty when b < a
        return list(v[a - 1:b]) if b >= a else []

    nmax, nmin = len(imax), len(imin)
    if imax[0] < imin[0]:
        if X(1) > X(imin[0]):
            lmax = fl(sl(imax, 2, min(nmax, nbsym + 1)))
            lmin = fl(sl(imin, 1, min(nmin, nbsym)))
            lsym, lb = imax[0], 0
        else:
            lmax = fl(sl(imax, 1, min(nmax, nbsym)))
            lmin = fl(sl(imin, 1, min(nmin, nbsym - 1))) + [1]
            lsym, lb = 1, 1
    else:
        if X(1) < X(imax[0]):
            lmax = fl(sl(imax, 1, min(nmax, nbsym)))
            lmin = fl(sl(imin, 2, min(nmin, nbsym + 1)))
            lsym, lb = imin[0], 2
        else:
            lmax = fl(sl(imax, 1, min(nmax, nbsym - 1))) + [1]
            lmin = fl(sl(imin, 1, min(nmin, nbsym)))
            lsym, lb = 1, 3
    if imax[-1] < imin[-1]:
        if X(lx) < X(imax[-1]):
            rmax = fl(sl(imax, max(nmax - nbsym + 1, 1), nmax))
            rmin = fl(sl(imin, max(nmin - nbsym, 1), nmin - 1))
            rsym, rb = imin[-1], 0
        else:
            rmax = [lx] + fl(sl(imax, max(nmax - nbsym + 2, 1), nmax))
            rmin = fl(sl(imin, max(nmin - nbsym + 1, 1), nmin))
            rsym, rb = lx, 1
    else:
        if X(lx) > X(imin[-1]):
            rmax = fl(sl(imax, max(nmax - nbsym, 1), nmax - 1))
            rmin = fl(sl(imin, max(nmin - nbsym + 1, 1), nmin))
            rsym, rb = imax[-1], 2
        else:
            rmax = fl(sl(imax, max(nmax - nbsym + 1, 1), nmax))
            rmin = [lx] + fl(sl(imin, max(nmin - nbsym + 2, 1), nmin))
            rsym, rb = lx, 3
    tl = lambda L: [2 * lsym - i for i in L]
    tr = lambda L: [2 * rsym - i for i in L]
    if tl(lmin)[0] > 1 or tl(lmax)[0] > 1:   # symmetrized parts do not reach t(1)
        if lsym == imax[0]:
            lmax = fl(sl(imax, 1, min(nmax, nbsym)))
        else:
            lmin = fl(sl(imin, 1, min(nmin, nbsym)))
        assert lsym != 1
        lsym, lb = 1, lb + 4
    if tr(rmin)[-1] < lx or tr(rmax)[-1] < lx:
        if rsym == imax[-1]:
            rmax = fl(sl(imax, max(nmax - nbsym + 1, 1), nmax))
        else:
            rmin = fl(sl(imin, max(nmin - nbsym + 1, 1), nmin))
        assert rsym != lx
        rsym, rb = lx, rb + 4
    tmin = np.array(tl(lmin) + imin + tr(rmin)) - 1
    tmax = np.array(tl(lmax) + imax + tr(rmax)) - 1
    smin = np.array(lmin + imin + rmin) - 1
    smax = np.array(lmax + imax + rmax) - 1
    return (tmin, smin), (tmax, smax), (lsym - 1, rsym - 1), (lb, rb)


def spline_env_bc(T, X, Y, bc):
    return CubicSpline(X.astype(float), Y, bc_type=bc)(np.arange(T, dtype=float))


def envelope_stats_mirror(y, V, bc="natural"):
    """One pass with mirrored ends: the reference's extrema (ends dropped), memd.m's
    boundary_conditions, CubicSpline envelopes through the mirrored knots, no clamp."""
    n, T = y.shape
    K = V.shape[0]
    P = V @ y
    acc = np.zeros((n, T))
    amp = np.zeros(T)
    nv = 0
    counts = np.zeros((K, 2), np.int32)
    for k in range(K):
        Xs = [simulate_extrema_kernel(P[k], 0), simulate_extrema_kernel(P[k], 1)]
        counts[k] = [len(Xs[0]), len(Xs[1])]
        if not (len(Xs[0]) > 3 and len(Xs[1]) > 3):
            continue
        nv += 1
        (tmin, smin), (tmax, smax), _, _ = boundary_conditions(Xs[1][1:-1], Xs[0][1:-1], P[k])
        emax = spline_env_bc(T, tmax, y[:, smax].T, bc).T
        emin = spline_env_bc(T, tmin, y[:, smin].T, bc).T
        acc += (emax + emin) / 2.0
        amp += np.sqrt(np.sum((emax - emin) ** 2, axis=0)) / 2.0
    with np.errstate(invalid="ignore", divide="ignore"):
        return acc / nv, amp / nv, counts, nv


def envelope_stats_reference(y, V):
    """One pass: mean over valid directions, memd.m amplitude, knot counts, nvalid."""
    n, T = y.shape
    K = V.shape[0]
    P = V @ y
    acc = np.zeros((n, T))
    amp = np.zeros(T)
    nv = 0
    counts = np.zeros((K, 2), np.int32)
    for k in range(K):
        Xs = [simulate_extrema_kernel(P[k], 0), simulate_extrema_kernel(P[k], 1)]
        counts[k] = [len(Xs[0]), len(Xs[1])]
        if not (len(Xs[0]) > 3 and len(Xs[1]) > 3):
            continue
        nv += 1
        emax = np.stack([spline_env(T, Xs[0], clamp_polyfit(Xs[0], y[c, Xs[0]], 0)) for c in range(n)])
        emin = np.stack([spline_env(T, Xs[1], clamp_polyfit(Xs[1], y[c, Xs[1]], 1)) for c in range(n)])
        acc += (emax + emin) / 2.0
        amp += np.sqrt(np.sum((emax - emin) ** 2, axis=0)) / 2.0
    with np.errstate(invalid="ignore", divide="ignore"):
        return acc / nv, amp / nv, counts, nv


def envelope_mean_reference(y, V):
    mean, _, counts, _ = envelope_stats_reference(y, V)
    return mean, counts


def memd_fix_reference(y, V, n_imf, S, stats=None):
    """memReducedEpochs_dev.cu:1174-1179, 1328-1338: S sifts per IMF, N_imf - 1 IMFs + residual."""
    stats = stats or envelope_stats_reference
    running = y.copy()
    out = []
    for _ in range(n_imf - 1):
        cur = running.copy()
        for _ in range(S):
            cur = cur - stats(cur, V)[0]
        out.append(cur)
        running = running - cur
    out.append(running)
    return np.stack(out)


class Fragile(Exception):
    pass


def stop_reference(mean, amp, counts, nv, sd, sd2, tol, margin=1e-9):
    """memd.m stop() on one pass; raises Fragile when an sx(t) is within `margin`
    (relative) of sd or sd2, where the oracle's rounding (about 1e-13 relative
    from scipy's) could put it on the other side."""
    if nv == 0:
        return True   # no valid direction: memd.m's zero mean, stop
    sx = np.sqrt(np.sum(mean ** 2, axis=0))
    if np.all(amp != 0):
        sx = sx / amp
    for thr in (sd, sd2):
        if np.any(np.abs(sx - thr) <= margin * thr):
            raise Fragile(f"sx within {margin} of {thr}")
    T = mean.shape[1]
    frac = np.count_nonzero(sx > sd) / T
    nem = (counts[:, 0] - 2) + (counts[:, 1] - 2)
    return not ((frac > tol or np.any(sx > sd2)) and np.any(nem > 2))


def memd_stop_reference(y, V, n_imf, S, sd, sd2, tol, stats=None):
    """memd.m's loop: stop_emd before each IMF, stop() before each subtraction, at
    most S subtractions per IMF, at most n_imf - 1 IMFs; returns (imfs, sifts, used)."""
    stats = stats or envelope_stats_reference
    running = y.copy()
    out, sifts = [], []
    for _ in range(n_imf - 1):
        cur = running.copy()
        mean, amp, counts, nv = stats(cur, V)
        if np.all((counts[:, 0] - 2) + (counts[:, 1] - 2) < 3):
            break   # stop_emd
        nsift = 0
        while not stop_reference(mean, amp, counts, nv, sd, sd2, tol):
            cur = cur - mean
            nsift += 1
            if nsift == S:
                break
            mean, amp, counts, nv = stats(cur, V)
        out.append(cur)
        sifts.append(nsift)
        running = running - cur
    out.append(running)
    sc = np.zeros(n_imf, np.int32)
    sc[:len(sifts)] = sifts
    return np.stack(out), sc, len(out)


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
    mean, amp, counts, _ = envelope_stats_reference(y, V)
    out["pass_y"] = y
    out["pass_V"] = V
    out["pass_mean"] = mean
    out["pass_amp"] = amp
    out["pass_counts"] = counts
    # the fixed-count loop of a small epoch
    y = np.stack([np.sin(2 * np.pi * t / 23.0 + c) + 0.5 * np.sin(2 * np.pi * t / 71.0 + 2 * c)
                  + 0.2 * rng.standard_normal(T) for c in range(n)])
    out["fix_y"] = y
    out["fix_V"] = V
    out["fix_imfs"] = memd_fix_reference(y, V, 4, 3)
    # stop-mode decompositions (memd.m's loop)
    import sys
    sys.path.insert(0, os.path.join(HERE, "..", "..", "memd-gpu_amd"))
    import synth
    cases = [
        ("c0", lambda: synth.sines_epoch(8, 1000, 2), 16, 12, 50, (0.075, 0.75, 0.075)),   # the 8 x 1000 host case
        ("uni", lambda: synth.sines_epoch(1, 700, 21), 4, 10, 30, (0.075, 0.75, 0.075)),   # univariate
        ("end", lambda: synth.sines_epoch(3, 64, 23), 6, 8, 10, (0.075, 0.75, 0.075)),     # ends by stop_emd
        ("thr", lambda: synth.eeg_epoch(4, 1500, 22), 8, 8, 20, (0.05, 0.5, 0.05)),        # other thresholds
    ]
    for name, gen, K, n_imf, S, (sd, sd2, tol) in cases:
        y = gen()
        n = y.shape[0]
        V = directions_closed_form(n, K, [-K] + primes_upto(n - 1))
        imfs, sc, used = memd_stop_reference(y, V, n_imf, S, sd, sd2, tol)
        out[f"stop_{name}_y"] = y
        out[f"stop_{name}_V"] = V
        out[f"stop_{name}_cfg"] = np.array([n_imf, S, sd, sd2, tol])
        out[f"stop_{name}_imfs"] = imfs
        out[f"stop_{name}_sifts"] = sc
        out[f"stop_{name}_used"] = np.array(used)
        print(name, "sifts", sc.tolist(), "slots used", used)
    # mirrored ends: knot lists of short projections, every branch of both ends
    rng2 = np.random.default_rng(20261204)
    T = 64
    rows, seen = [], set()
    for trial in range(8000):
        tt = np.arange(T)
        Pk = np.sin(2 * np.pi * tt / rng2.uniform(6, 30) + rng2.uniform(0, 6.3)) + rng2.uniform(0, 0.8) * rng2.standard_normal(T)
        if trial % 2:   # monotone ramps at the ends: reflections that fall short of the end sample
            a, b = rng2.integers(0, 24, 2)
            Pk[:a] = Pk[a] + rng2.uniform(-1.5, 1.5) * (a - tt[:a]) / max(a, 1)
            Pk[T - b:] = Pk[T - b - 1] + rng2.uniform(-1.5, 1.5) * (tt[T - b:] - (T - b - 1)) / max(b, 1)
        Xa, Xi = simulate_extrema_kernel(Pk, 0), simulate_extrema_kernel(Pk, 1)
        if len(Xa) < 4 or len(Xi) < 4:
            continue
        res = boundary_conditions(Xi[1:-1], Xa[1:-1], Pk)
        key = res[3]
        if key in seen and trial % 50:
            continue
        seen.add(key)
        rows.append((Pk, res))
    lbs = {r[1][3][0] for r in rows}
    rbs = {r[1][3][1] for r in rows}
    assert lbs >= {0, 1, 2, 3, 4, 6} and rbs >= {0, 1, 2, 3, 4, 6}, (lbs, rbs)
    C, W = len(rows), T // 2 + 4
    mk = {k: np.zeros((C, W), np.int32) for k in ("pmax", "smax", "pmin", "smin")}
    mk_e = np.zeros((C, 2), np.int32)
    mk_sym = np.zeros((C, 2), np.int32)
    mk_br = np.zeros((C, 2), np.int32)
    for i, (Pk, ((tmin, smin), (tmax, smax), sym, br)) in enumerate(rows):
        mk["pmax"][i, :len(tmax)], mk["smax"][i, :len(smax)] = tmax, smax
        mk["pmin"][i, :len(tmin)], mk["smin"][i, :len(smin)] = tmin, smin
        mk_e[i] = [len(tmax), len(tmin)]
        mk_sym[i] = sym
        mk_br[i] = br
    out["mir_P"] = np.stack([r[0] for r in rows])
    for k, v in mk.items():
        out[f"mir_{k}"] = v
    out["mir_E"] = mk_e
    out["mir_sym"] = mk_sym
    out["mir_branch"] = mk_br
    print("mirror knot cases", C, "left branches", sorted(lbs), "right branches", sorted(rbs))
    # mirrored ends: one pass (both end conditions), a fixed-count loop and memd.m's own
    # configuration (stop criterion, mirrored ends, not-a-knot splines)
    n, T, K = 3, 300, 8
    t = np.arange(T)
    y = np.stack([np.sin(2 * np.pi * t / 29.0 + 1.3 * c) + 0.4 * np.sin(2 * np.pi * t / 83.0 + c)
                  + 0.3 * rng2.standard_normal(T) for c in range(n)])
    V = directions_closed_form(n, K, [-K] + primes_upto(n - 1))
    out["mpass_y"] = y
    out["mpass_V"] = V
    for bc, tag in (("natural", "nat"), ("not-a-knot", "nak")):
        mean, amp, counts, _ = envelope_stats_mirror(y, V, bc)
        out[f"mpass_mean_{tag}"] = mean
        out[f"mpass_amp_{tag}"] = amp
        out[f"mpass_counts_{tag}"] = counts
    out["mfix_imfs"] = memd_fix_reference(y, V, 4, 3, stats=lambda c, v: envelope_stats_mirror(c, v, "natural"))
    y = synth.sines_epoch(4, 900, 31)
    V = directions_closed_form(4, 12, [-12] + primes_upto(3))
    imfs, sc, used = memd_stop_reference(y, V, 6, 20, 0.075, 0.75, 0.075,
                                         stats=lambda c, v: envelope_stats_mirror(c, v, "not-a-knot"))
    out["mstop_y"] = y
    out["mstop_V"] = V
    out["mstop_imfs"] = imfs
    out["mstop_sifts"] = sc
    out["mstop_used"] = np.array(used)
    print("memd.m configuration sifts", sc.tolist(), "slots used", used)
    np.savez_compressed(os.path.join(HERE, "golden_small.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
