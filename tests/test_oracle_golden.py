"""Pins the FP64 oracle against independent answers (CPU only).

Golden vectors come from tests/golden/make_golden.py (numpy/scipy, different
algorithms from the oracle); the rest are closed forms and known values.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "golden_small.npz")


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLD, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


# ---- directions (memReducedEpochs_dev.cu:934-1072) ---------------------------

def test_radical_inverse_known_values(oracle):
    known = {(2, 1): 0.5, (2, 2): 0.25, (2, 3): 0.75, (2, 4): 0.125, (2, 5): 0.625,
             (3, 1): 1 / 3, (3, 2): 2 / 3, (3, 3): 1 / 9, (3, 4): 4 / 9, (5, 7): 2 / 5 + 1 / 25}
    for (b, i), v in known.items():
        assert oracle.radical_inverse(b, i) == pytest.approx(v, abs=1e-15)


def test_prime_table_reference_rule(oracle):
    # memReducedEpochs_dev.cu:1033-1046: P[0] = -K, then 2,3,5,7,...
    P = oracle.prime_table(8, 64)
    assert P.tolist() == [-64, 2, 3, 5, 7, 11, 13, 17]


@pytest.mark.parametrize("n,K", [(1, 4), (2, 8), (3, 5), (8, 16), (32, 64)])
def test_directions_closed_form(oracle, gold, n, K):
    V = oracle.directions(n, K)
    np.testing.assert_allclose(V, gold[f"dirs_ref_{n}_{K}"], rtol=0, atol=2e-14)
    np.testing.assert_allclose(np.linalg.norm(V, axis=1), 1.0, atol=1e-14)


@pytest.mark.parametrize("n,K", [(3, 7), (32, 64)])
def test_directions_mex_primes(oracle, gold, n, K):
    P = gold[f"dirs_mex_{n}_{K}_primes"]
    V = oracle.directions(n, K, primes=P)
    np.testing.assert_allclose(V, gold[f"dirs_mex_{n}_{K}"], rtol=0, atol=2e-14)
    # the MEX path differs from the C++ path (P[0] = 2 instead of -K)
    assert not np.allclose(V, oracle.directions(n, K))


def test_direction_of_single_channel_is_one(oracle):
    assert np.array_equal(oracle.directions(1, 9), np.ones((9, 1)))


# ---- projection ---------------------------------------------------------------

def test_projection_matches_matmul(oracle):
    rng = np.random.default_rng(3)
    cur = rng.standard_normal((7, 333))
    V = rng.standard_normal((5, 7))
    np.testing.assert_allclose(oracle.project(cur, V), V @ cur, rtol=1e-14, atol=1e-14)


# ---- extrema window (memReducedEpochs_dev.cu:77-141, :1191-1193) ---------------

@pytest.mark.parametrize("T", [5, 100, 242, 483, 484, 1000, 4096, 8193])
def test_extrema_match_kernel_simulation(oracle, gold, T):
    P = gold[f"ext_P_{T}"]
    assert np.array_equal(oracle.extrema(P, 0), gold[f"ext_max_{T}"])
    assert np.array_equal(oracle.extrema(P, 1), gold[f"ext_min_{T}"])


@pytest.mark.parametrize("T,gap", [(1000, 0), (4096, 14), (8193, 31), (16384, 62), (30504, 22),
                                   (60000, 478), (241665, 1903)])
def test_window_gap_lengths(oracle, T, gap):
    assert (T - 2) - oracle.window_limit(T, oracle.WINDOW_REFWARP32) == gap
    assert oracle.window_limit(T, oracle.WINDOW_FULL) == T - 2


def test_full_window_tests_every_interior_sample(oracle):
    rng = np.random.default_rng(5)
    P = rng.standard_normal(4096)
    want = np.r_[0, 1 + np.flatnonzero((P[1:-1] > P[:-2]) & (P[1:-1] > P[2:])), 4095]
    assert np.array_equal(oracle.extrema(P, 0, oracle.WINDOW_FULL), want)


# ---- clamp + natural spline (memReducedEpochs_dev.cu:447-584, 704-830) --------

@pytest.mark.parametrize("case", range(5))
def test_clamp_and_envelope_match_scipy(oracle, gold, case):
    X, Y, T = gold[f"spl_X_{case}"], gold[f"spl_Y_{case}"], int(gold[f"spl_T_{case}"])
    for e in range(2):
        Yc = oracle.clamp(X, Y, e)
        np.testing.assert_allclose(Yc, gold[f"spl_Yc_{case}_{e}"], rtol=1e-11, atol=1e-12)  # polyfit lstsq rounding
        M = oracle.spline_solve(X, Yc[None, :])[0]
        env = oracle.envelope(T, X, oracle.spline_coefs(X, Yc, M))
        ref = gold[f"spl_env_{case}_{e}"]
        np.testing.assert_allclose(env, ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())


def test_clamp_only_from_six_knots(oracle):
    # scan[T-1] > 4 <=> E >= 6 (memReducedEpochs_dev.cu:454)
    X = np.array([0, 2, 5, 9, 14], np.int32)
    Y = np.array([-10.0, 1.0, 2.0, 3.0, -10.0])
    assert np.array_equal(oracle.clamp(X, Y, 0), Y)
    X6 = np.array([0, 2, 5, 9, 14, 20], np.int32)
    Y6 = np.array([-10.0, 1.0, 2.0, 5.0, 3.0, -10.0])
    got = oracle.clamp(X6, Y6, 0)
    assert got[0] == pytest.approx(1.0 - (1.0 / 3.0) * 2)      # line through knots 1,2 at X0
    assert got[-1] == pytest.approx(3.0 + (-2.0 / 5.0) * 6)    # line through knots 3,4 at X5
    # min envelope only lowers
    got_min = oracle.clamp(X6, Y6, 1)
    assert got_min[0] == -10.0 and got_min[-1] == -10.0


@pytest.mark.parametrize("N", [1, 2, 31, 32, 33, 63, 64, 65, 200, 1000])
def test_blocked_tridiagonal_matches_dense(oracle, gold, N):
    lo, di, up, r = gold[f"tri_{N}"]
    x = oracle.tridiag_blocked(lo, di, up, r)
    ref = gold[f"tri_x_{N}"]
    assert x[0] == 0.0 and x[N + 1] == 0.0
    np.testing.assert_allclose(x[1:N + 1], ref, rtol=0, atol=4e-15 * np.abs(ref).max())


def test_blocked_solver_block_boundaries_shift_nothing(oracle):
    # a planted-error check: the residual of the blocked solve must be at rounding level
    rng = np.random.default_rng(11)
    for N in [95, 96, 97, 500]:
        h = rng.integers(1, 20, N + 1).astype(float)
        lo = np.r_[0.0, h[:N], 0.0]
        up = np.r_[0.0, h[1:N + 1], 0.0]
        di = np.r_[0.0, 2.0 * (h[:N] + h[1:N + 1]), 0.0]
        r = np.r_[0.0, rng.standard_normal(N), 0.0]
        x = oracle.tridiag_blocked(lo, di, up, r)
        res = di[1:N + 1] * x[1:N + 1] + lo[1:N + 1] * x[0:N] + up[1:N + 1] * x[2:N + 2] - r[1:N + 1]
        assert np.abs(res).max() < 1e-13


# ---- one envelope-mean pass (memReducedEpochs_dev.cu:1181-1318) ---------------

def test_envelope_mean_pass_matches_independent_pipeline(oracle, gold):
    y, V = gold["pass_y"], gold["pass_V"]
    mean, _, counts = oracle.envelope_mean(y, V)
    assert np.array_equal(counts, gold["pass_counts"])
    ref = gold["pass_mean"]
    np.testing.assert_allclose(mean, ref, rtol=0, atol=1e-11 * np.abs(ref).max())


@pytest.mark.parametrize("seed", range(6))
def test_not_a_knot_matches_scipy(oracle, seed):
    # optional end condition (the reference's dead tridiagonal_setup_nak, restated
    # pivot-free): the spline must be scipy's not-a-knot interpolant, including equal
    # end gaps (where the reference's row-0 form has a zero diagonal) and knot counts
    # that span several solver blocks
    from scipy.interpolate import CubicSpline
    rng = np.random.default_rng(100 + seed)
    for E in (4, 5, 7, 33, 34, 65, 200):
        gaps = rng.integers(1, 15, size=E - 1)
        if seed % 2 == 0:
            gaps[:3] = 4
            gaps[-3:] = 9
        X = np.concatenate([[0], np.cumsum(gaps)]).astype(np.int32)
        Y = rng.standard_normal((2, E)) * 50
        M = oracle.spline_solve(X, Y, spline=oracle.SPLINE_NOT_A_KNOT)
        T = int(X[-1]) + 1
        for c in range(2):
            coef = oracle.spline_coefs(X, Y[c], M[c])
            env = oracle.envelope(T, X, coef)
            ref = CubicSpline(X.astype(float), Y[c], bc_type="not-a-knot")(np.arange(T, dtype=float))
            assert np.abs(env - ref).max() <= 1e-11 * np.abs(ref).max()
            # not-a-knot: the cubic term is shared by the first two and the last two intervals
            d = coef[:, 3]
            assert abs(d[0] - d[1]) <= 1e-9 * max(abs(d).max(), 1e-300)
            assert abs(d[-1] - d[-2]) <= 1e-9 * max(abs(d).max(), 1e-300)
            # and it interpolates
            assert np.allclose(env[X], Y[c], rtol=0, atol=1e-12 * np.abs(Y[c]).max())


def test_not_a_knot_four_knots_is_one_cubic(oracle):
    # E = 4: both end conditions leave a single cubic through the four knots
    X = np.array([0, 3, 4, 9], np.int32)
    Y = np.array([[1.0, -2.0, 0.5, 3.0]])
    M = oracle.spline_solve(X, Y, spline=oracle.SPLINE_NOT_A_KNOT)
    coef = oracle.spline_coefs(X, Y[0], M[0])
    env = oracle.envelope(10, X, coef)
    p = np.polyfit(X.astype(float), Y[0], 3)
    assert np.allclose(env, np.polyval(p, np.arange(10.0)), rtol=0, atol=1e-12)


def test_not_a_knot_memd_reconstructs(oracle):
    import synth
    y = synth.sines_epoch(4, 900, 3)
    V = oracle.directions(4, 8)
    imfs, info = oracle.memd(y, V, 4, 3, spline=oracle.SPLINE_NOT_A_KNOT)
    assert info["n_imf_out"] == 4
    np.testing.assert_allclose(imfs.sum(0), y, rtol=0, atol=1e-12 * np.abs(y).max())
    nat, _ = oracle.memd(y, V, 4, 3)
    assert not np.array_equal(nat, imfs)   # the option changes the envelopes


@pytest.mark.parametrize("mode", [0, 1])
def test_threaded_oracle_equals_serial(oracle, mode):
    # the OpenMP path (channels in parallel, one solver factorisation shared by them)
    # must give the serial path's results bit for bit: it is the checker of the
    # full-size GPU tests and the CPU baseline; 8000 samples give many solver blocks
    import synth
    y = synth.eeg_epoch(16, 8000, 3)
    V = oracle.directions(16, 8)
    a, ia = oracle.memd(y, V, 3, 2, mode=mode, threads=1)
    for th in (2, 8):
        b, ib = oracle.memd(y, V, 3, 2, mode=mode, threads=th)
        assert np.array_equal(ia["knot_counts"], ib["knot_counts"])
        assert np.array_equal(a, b)


# ---- memd.m amplitude, the IMF loops and the stop criterion ---------------------

def test_envelope_amplitude_matches_independent_pipeline(oracle, gold):
    # memd.m's amp(t) = sum over valid directions of |e+(t) - e-(t)|_2 / 2, / nvalid
    _, amp, _ = oracle.envelope_mean(gold["pass_y"], gold["pass_V"], want_amp=True)
    ref = gold["pass_amp"]
    np.testing.assert_allclose(amp, ref, rtol=0, atol=1e-11 * np.abs(ref).max())


def test_fix_loop_matches_independent_pipeline(oracle, gold):
    # memReducedEpochs_dev.cu:1174-1179, 1328-1338: 3 IMFs of 3 sifts + residual
    imfs, info = oracle.memd(gold["fix_y"], gold["fix_V"], 4, 3)
    ref = gold["fix_imfs"]
    assert info["n_imf_out"] == 4 and info["sift_counts"].tolist() == [3, 3, 3, 0]
    np.testing.assert_allclose(imfs, ref, rtol=0, atol=1e-10 * np.abs(ref).max())


def test_fix_loop_composes_single_sifts(oracle):
    # the loop against compositions of the (separately pinned) envelope-mean pass,
    # bit for bit: a sift is cur - mean(cur); the next IMF starts from running - IMF
    import synth
    y = synth.sines_epoch(4, 800, 41)
    V = oracle.directions(4, 8)
    m = lambda x: oracle.envelope_mean(x, V)[0]
    one, _ = oracle.memd(y, V, 2, 1)
    assert np.array_equal(one[0], y - m(y))
    assert np.array_equal(one[1], y - one[0])        # residual slot
    two, _ = oracle.memd(y, V, 2, 2)
    s1 = y - m(y)
    assert np.array_equal(two[0], s1 - m(s1))         # S = 2: the second sift acts on the first's result
    three, _ = oracle.memd(y, V, 3, 1)
    r1 = y - one[0]
    assert np.array_equal(three[0], one[0])
    assert np.array_equal(three[1], r1 - m(r1))       # IMF 2 sifts the residual, not the signal
    assert np.array_equal(three[2], r1 - three[1])


@pytest.mark.parametrize("name", ["c0", "uni", "end", "thr"])
def test_stop_mode_matches_memd_m_restatement(oracle, gold, name):
    # memd.m's loop (stop_emd, stop(), at most S subtractions per IMF) written as
    # plain loops over numpy/scipy stages; c0 is the 8 x 1000, K = 16 host case
    y, V = gold[f"stop_{name}_y"], gold[f"stop_{name}_V"]
    n_imf, S, sd, sd2, tol = gold[f"stop_{name}_cfg"]
    imfs, info = oracle.memd(y, V, int(n_imf), int(S), mode=oracle.MODE_STOP, sd=sd, sd2=sd2, tol=tol)
    used = int(gold[f"stop_{name}_used"])
    assert info["n_imf_out"] == used
    assert info["sift_counts"].tolist() == gold[f"stop_{name}_sifts"].tolist()
    ref = gold[f"stop_{name}_imfs"]
    for i in range(used):
        assert np.linalg.norm(imfs[i] - ref[i]) <= 1e-9 * np.linalg.norm(y), f"slot {i}"


def _counts(nem_per_dir):
    # knot counts (endpoints included) giving these interior-extrema totals
    return np.array([[2 + v // 2, 2 + v - v // 2] for v in nem_per_dir], np.int32)


def test_stop_decision_known_answers(oracle):
    T = 40
    amp = np.ones(T)
    mean = np.zeros((2, T))
    rich = _counts([5, 4])
    # zero mean: sx = 0 everywhere, stop
    assert oracle.stop_decision(mean, amp, rich)
    # one sample above sd2 keeps sifting (any(sx > sd2))
    m = mean.copy(); m[0, 7] = 0.8
    assert not oracle.stop_decision(m, amp, rich)
    # ... unless no direction has more than 2 interior extrema
    assert oracle.stop_decision(m, amp, _counts([2, 1]))
    assert not oracle.stop_decision(m, amp, _counts([2, 3]))
    # fraction above sd: 3 of 40 is exactly tol = 0.075, not above it -> stop;
    # 4 of 40 is above -> sift
    m = mean.copy(); m[1, :3] = 0.1
    assert oracle.stop_decision(m, amp, rich)
    m[1, 3] = 0.1
    assert not oracle.stop_decision(m, amp, rich)
    # sx is compared strictly: a sample exactly at sd2 does not count as above it,
    # nor do 4 of 40 samples exactly at sd (sqrt(x * x) == x, so sx is exactly 0.075)
    m = mean.copy(); m[0, 0] = 0.75
    assert oracle.stop_decision(m, amp, rich)
    m = mean.copy(); m[0, :4] = 0.075
    assert oracle.stop_decision(m, amp, rich)
    m[0, 4] = 0.0750000000000001
    assert oracle.stop_decision(m, amp, rich)          # 1 of 40 above sd
    m[0, :4] = 0.0750000000000001
    assert not oracle.stop_decision(m, amp, rich)      # 5 of 40 above sd
    # sx is the 2-norm over channels: 0.6 and 0.6 give 0.849 > sd2
    m = mean.copy(); m[0, 5] = 0.6; m[1, 5] = 0.6
    assert not oracle.stop_decision(m, amp, rich)
    # amplitude scaling: |mean| = 0.5 everywhere (above sd at every sample) ...
    m = np.zeros((1, T)); m[0, :] = 0.5
    assert not oracle.stop_decision(m, np.ones(T), _counts([4]))
    # ... divided by amp = 10 it is 0.05 < sd: stop
    assert oracle.stop_decision(m, np.full(T, 10.0), _counts([4]))
    # ... but with one amp(t) = 0 memd.m does not divide at all: 0.5 > sd, sift
    # (dividing anyway would give 0.05 elsewhere and 0/0 at the zero: stop)
    a = np.full(T, 10.0); a[9] = 0.0
    m[0, 9] = 0.0
    assert not oracle.stop_decision(m, a, _counts([4]))


def test_stop_mode_with_zero_thresholds_is_fix_mode(oracle):
    # sd = sd2 = tol = 0: every pass with any nonzero mean sample keeps sifting, so
    # each IMF takes exactly S subtractions, like fix mode, bit for bit
    import synth
    y = synth.sines_epoch(4, 900, 42)
    V = oracle.directions(4, 8)
    fix, finfo = oracle.memd(y, V, 4, 3)
    stop, sinfo = oracle.memd(y, V, 4, 3, mode=oracle.MODE_STOP, sd=0.0, sd2=0.0, tol=0.0)
    assert sinfo["n_imf_out"] == 4 and sinfo["sift_counts"].tolist() == finfo["sift_counts"].tolist()
    assert np.array_equal(stop, fix)


def test_stop_mode_with_infinite_thresholds_never_sifts(oracle):
    # sd = sd2 = inf: no sample is ever above a threshold, the first IMF is the
    # signal itself after 0 sifts, and the zero residual has no extrema left
    import synth
    y = synth.sines_epoch(4, 900, 43)
    V = oracle.directions(4, 8)
    imfs, info = oracle.memd(y, V, 6, 5, mode=oracle.MODE_STOP, sd=np.inf, sd2=np.inf)
    assert info["n_imf_out"] == 2
    assert info["sift_counts"].tolist() == [0] * 6
    assert np.array_equal(imfs[0], y)
    assert not imfs[1].any()


@pytest.mark.parametrize("periods,slots", [(1.0, 1), (1.6, 2)])
def test_stop_mode_ends_below_three_interior_extrema(oracle, periods, slots):
    # memd.m stop_emd: no IMF is started once every direction has fewer than 3
    # interior extrema.  One channel (every direction is [1]): one sine period has
    # one maximum and one minimum inside, so the decomposition ends at once and the
    # only slot is the residual; 1.6 periods have 3 and start an IMF
    T = 64
    t = np.arange(T)
    y = np.sin(2 * np.pi * periods * t / T + 0.3)[None, :]
    nint = len(oracle.extrema(y[0], 0)) - 2 + len(oracle.extrema(y[0], 1)) - 2
    assert nint == (2 if periods == 1.0 else 3)
    V = oracle.directions(1, 4)
    imfs, info = oracle.memd(y, V, 5, 4, mode=oracle.MODE_STOP)
    assert info["n_imf_out"] == slots
    if slots == 1:
        assert np.array_equal(imfs[0], y)


def test_epoch_parallel_batch_equals_single_epochs(oracle):
    # the CPU baseline's batched form: every epoch as oracle_memd gives it alone
    import synth
    eps = [(synth.eeg_epoch(n, T, 40 + i), oracle.directions(n, K))
           for i, (n, T, K) in enumerate([(4, 900, 8), (8, 1500, 16), (2, 700, 4), (16, 1200, 32), (3, 64, 6)])]
    for mode in (oracle.MODE_FIX, oracle.MODE_STOP):
        outs = oracle.memd_batch(eps, 4, 3, threads=4, mode=mode)
        for (y, V), got in zip(eps, outs):
            want, _ = oracle.memd(y, V, 4, 3, mode=mode)
            assert np.array_equal(got, want)


def test_amplitude_channel_sum_order(oracle):
    # sum_c d^2 in groups of 4 channels (sequential), 4 groups to a block of 16, blocks in
    # order: known answers where that order shows in the rounding (ulp(2^54) = 4, so a
    # single 1.0 added to 2^54 is lost while a group's 4.0 is kept)
    big = 2.0 ** 27                                                        # big^2 = 2^54
    assert oracle.amp_block_sum(np.array([3.0, 4.0])) == 25.0
    assert oracle.amp_block_sum(np.ones(37)) == 37.0
    # plain sequential order would give 2^54 (every 1.0 lost); groups keep the second group's 4
    assert oracle.amp_block_sum(np.array([big] + [1.0] * 7)) == 2.0 ** 54 + 4.0
    # block 0: 2^54 + 3 groups of 4; block 1: 4 groups of 4 x 0.25 = 1 each, summed to 4
    # before it meets block 0 (groups added one by one would lose all four)
    assert oracle.amp_block_sum(np.array([big] + [1.0] * 15 + [0.5] * 16)) == 2.0 ** 54 + 16.0
    # n <= 4: the plain sequential sum
    d = np.array([big, 1.0, 1.0, 1.0])
    assert oracle.amp_block_sum(d) == 2.0 ** 54
