"""Pins the oracle's mirrored ends (memd.m's boundary_conditions, nbsym = 2) on the CPU.

The mirrored-end option is not in the reference (its ends are the forced
samples 0 and T-1 with the clamp of memReducedEpochs_dev.cu:447-486, 545-584).
It follows the public memd.m toolbox, so the pins are:

* a transcription of boundary_conditions with MATLAB's 1-based indices and
  inclusive slices (tests/golden/make_golden.py), compared knot by knot on short
  projections that take every reachable branch at both ends;
* what the construction fixes: added knots are reflections of extrema of their
  own list (or of the end sample) about one mirror point per end, the lists stay
  strictly increasing and cover samples 0 .. T-1, the interior is untouched;
* scipy's CubicSpline (natural and not-a-knot) through the mirrored knots for a
  whole envelope-mean pass, the fixed-count loop and memd.m's own configuration
  (stop criterion, mirrored ends, not-a-knot);
* a closed form: extrema of one height give flat envelopes.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "golden_small.npz")


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLD, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _cases(oracle, gold):
    for i, Pk in enumerate(gold["mir_P"]):
        Xa, Xi = oracle.extrema(Pk, 0), oracle.extrema(Pk, 1)
        yield i, Pk, Xa, Xi, oracle.mirror_knots(Pk, Xa, Xi)


def test_mirror_knots_match_memd_m_transcription(oracle, gold):
    branches = set()
    for i, Pk, Xa, Xi, ((pa, sa), (pi, si), sym) in _cases(oracle, gold):
        Ea, Ei = gold["mir_E"][i]
        assert (len(pa), len(pi)) == (Ea, Ei), i
        assert np.array_equal(pa, gold["mir_pmax"][i, :Ea]), i
        assert np.array_equal(sa, gold["mir_smax"][i, :Ea]), i
        assert np.array_equal(pi, gold["mir_pmin"][i, :Ei]), i
        assert np.array_equal(si, gold["mir_smin"][i, :Ei]), i
        assert sym == tuple(gold["mir_sym"][i]), i
        branches.add(tuple(gold["mir_branch"][i]))
    lb = {b[0] for b in branches}
    rb = {b[1] for b in branches}
    # four cases per end plus the two reachable "does not extend enough" corrections
    assert lb == {0, 1, 2, 3, 4, 6} and rb == {0, 1, 2, 3, 4, 6}


def test_mirror_knots_are_reflections_about_the_end_points(oracle, gold):
    for i, Pk, Xa, Xi, lists in _cases(oracle, gold):
        T = len(Pk)
        lsym, rsym = lists[2]
        assert 0 <= lsym < T and 0 <= rsym < T
        for (P, S), X in zip(lists[:2], (Xa, Xi)):
            inner = X[1:-1]
            # the interior extrema, unchanged, between 1 or 2 added knots per end
            nl = int(np.searchsorted(P, inner[0]))
            nr = len(P) - nl - len(inner)
            assert 1 <= nl <= 2 and 1 <= nr <= 2, i
            assert np.array_equal(P[nl:nl + len(inner)], inner) and np.array_equal(S[nl:nl + len(inner)], inner)
            assert np.all(np.diff(P) > 0), i
            assert P[0] <= 0 and P[-1] >= T - 1, i
            # each added knot mirrors an extremum of its own list, or the end sample
            for j in range(nl):
                assert P[j] == 2 * lsym - S[j]
                assert S[j] in inner or S[j] == 0
            for j in range(len(P) - nr, len(P)):
                assert P[j] == 2 * rsym - S[j]
                assert S[j] in inner or S[j] == T - 1
            # the mirror point is an extremum nearest the end, or the end sample itself
            assert lsym in (0, Xa[1], Xi[1]) and rsym in (T - 1, Xa[-2], Xi[-2])


@pytest.mark.parametrize("tag,spline", [("nat", 0), ("nak", 1)])
def test_mirror_pass_matches_scipy(oracle, gold, tag, spline):
    y, V = gold["mpass_y"], gold["mpass_V"]
    mean, amp, counts = oracle.envelope_mean(y, V, want_amp=True, spline=spline, boundary=oracle.BOUNDARY_MIRROR)
    # knot counts stay the clamp lists' (interior extrema + 2): stop mode's nem reads them
    assert np.array_equal(counts, gold[f"mpass_counts_{tag}"])
    ref = gold[f"mpass_mean_{tag}"]
    np.testing.assert_allclose(mean, ref, rtol=0, atol=1e-11 * np.abs(ref).max())
    ref = gold[f"mpass_amp_{tag}"]
    np.testing.assert_allclose(amp, ref, rtol=0, atol=1e-11 * np.abs(ref).max())
    clamp, _, _ = oracle.envelope_mean(y, V, spline=spline)
    assert np.abs(clamp - mean).max() > 1e-3 * np.abs(mean).max()   # the option changes the ends


def test_mirror_fix_loop_matches_independent_pipeline(oracle, gold):
    imfs, info = oracle.memd(gold["mpass_y"], gold["mpass_V"], 4, 3, boundary=oracle.BOUNDARY_MIRROR)
    ref = gold["mfix_imfs"]
    assert info["sift_counts"].tolist() == [3, 3, 3, 0]
    np.testing.assert_allclose(imfs, ref, rtol=0, atol=1e-10 * np.abs(ref).max())


def test_memd_m_configuration_matches_transcription(oracle, gold):
    # memd.m as published: stop criterion (0.075, 0.75, 0.075), mirrored ends, not-a-knot
    imfs, info = oracle.memd(gold["mstop_y"], gold["mstop_V"], 6, 20, mode=oracle.MODE_STOP,
                             spline=oracle.SPLINE_NOT_A_KNOT, boundary=oracle.BOUNDARY_MIRROR)
    used = int(gold["mstop_used"])
    assert info["n_imf_out"] == used
    assert np.array_equal(info["sift_counts"], gold["mstop_sifts"])
    ref = gold["mstop_imfs"]
    np.testing.assert_allclose(imfs[:used], ref, rtol=0, atol=1e-9 * np.abs(ref).max())


@pytest.mark.parametrize("spline", [0, 1])
def test_mirror_of_level_extrema_is_flat(oracle, spline):
    # a triangle wave whose maxima are all 1 and minima all -1: every knot of the
    # upper list (mirrored ones included) is 1, so both splines are constant and
    # the mean is exactly 0 -- whatever phase the ends cut the wave at
    for phase in range(8):
        y = np.array([[1.0, 0.5, 0.0, -0.5, -1.0, -0.5, 0.0, 0.5][(t + phase) % 8] for t in range(203)])[None, :]
        V = np.ones((1, 1))
        mean, amp, counts = oracle.envelope_mean(y, V, want_amp=True, spline=spline, boundary=oracle.BOUNDARY_MIRROR)
        assert np.all(mean == 0.0), phase
        assert np.all(amp == 1.0), phase


def test_mirror_threaded_oracle_equals_serial(oracle, gold):
    import synth
    y = synth.eeg_epoch(6, 1200, 5)
    V = oracle.directions(6, 12)
    a, _ = oracle.memd(y, V, 3, 3, boundary=oracle.BOUNDARY_MIRROR)
    b, _ = oracle.memd(y, V, 3, 3, boundary=oracle.BOUNDARY_MIRROR, threads=4)
    assert np.array_equal(a, b)
