"""HIP engine vs the FP64 oracle, through the C ABI (needs an MI355X).

Bar: bit-exact.  The engine and the oracle share one arithmetic contract
(FP64, contraction off, fma where both write it), so knot counts per direction
per sift, sift counts, IMF count and every IMF sample must be identical.  The
floating-point tolerance for IMF samples (1e-9 relative L2) is asserted too, as
a weaker second line.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD_DIR = os.path.join(os.path.dirname(__file__), "golden")


def _torch():
    import torch
    return torch


def _dev(a, dtype=None):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0") if dtype is None else \
        torch.from_numpy(np.ascontiguousarray(a, dtype)).to("cuda:0")


def _rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _check_epoch(ctx, oracle, y, K, n_imf, S, window=0, primes=None, use_graph=True, mode=0):
    n, T = y.shape
    V = oracle.directions(n, K, primes)
    imfs_o, info_o = oracle.memd(y, V, n_imf, S, window=window, mode=mode)
    passes = info_o["n_passes"]
    cap = (n_imf - 1) * (S + 1) + 1
    imfs_g, info_g = ctx.run(_dev(y), _dev(V), n_imf, S, window=window, use_graph=use_graph, mode=mode,
                             trace_passes=cap)
    got = imfs_g.cpu().numpy()
    assert info_g["n_imf_out"] == info_o["n_imf_out"]
    assert np.array_equal(info_g["sift_counts"], info_o["sift_counts"])
    kc = info_g["knot_counts"].cpu().numpy()[:passes]
    assert np.array_equal(kc, info_o["knot_counts"]), "knot counts differ"
    used = info_o["n_imf_out"]
    got, imfs_o = got[:used], imfs_o[:used]
    assert _rel_l2(got, imfs_o) <= 1e-9
    mism = np.argwhere(got != imfs_o)
    assert mism.size == 0, f"{len(mism)} samples differ, first at {mism[0].tolist()}"
    return got


def test_directions_bit_exact(ctx, oracle):
    for n, K in [(1, 4), (8, 16), (32, 64), (128, 256)]:
        got = ctx.directions(n, K).cpu().numpy()
        assert np.array_equal(got, oracle.directions(n, K))
    P = np.array([2, 3, 5, 7, 11, 13, 17, 19], np.int32)
    assert np.array_equal(ctx.directions(8, 16, primes=P).cpu().numpy(), oracle.directions(8, 16, P))


@pytest.mark.parametrize("n,T,K,n_imf,S,seed", [
    (8, 1000, 16, 9, 10, 0),      # the 8 x 1000, K=16 host case, reference schedule
    (3, 300, 6, 4, 3, 1),
    (1, 500, 4, 5, 4, 2),         # n = 1 reduces to univariate EMD
    (5, 777, 9, 3, 3, 3),         # ragged sizes
    (33, 2049, 7, 3, 2, 4),       # channel count not a power of two
    (2, 5000, 3, 3, 2, 5),        # many solver blocks (E ~ 1600 knots)
])
def test_run_matches_oracle_sines(ctx, oracle, n, T, K, n_imf, S, seed):
    import synth
    _check_epoch(ctx, oracle, synth.sines_epoch(n, T, seed), K, n_imf, S)


def test_headline_shape_full_schedule(ctx, oracle):
    # 32 ch x 4096 samples, K = 64, the reference's 9 slots x 10 sifts
    import synth
    y = synth.sines_epoch(32, 4096, 7)
    got = _check_epoch(ctx, oracle, y, 64, 9, 10)
    np.testing.assert_allclose(got.sum(0), y, rtol=0, atol=1e-12 * np.abs(y).max())


@pytest.mark.parametrize("n,T,K,n_imf,S,seed", [
    (128, 2048, 256, 3, 2, 60),   # the 128 x 16384, K=256 configuration's channels and directions
    (200, 700, 16, 3, 2, 61),     # above 128 channels, ragged evaluation groups
    (65, 1500, 130, 3, 2, 62),    # 4 full 16-channel chunks and 1 channel; 130 directions: a partial last group
    (23, 900, 40, 3, 2, 63),      # a 7-channel chunk (4 + 3 channels); 40 directions: groups of 32 and 8
])
def test_high_density_matches_oracle(ctx, oracle, n, T, K, n_imf, S, seed):
    # wide epochs: the projection/extrema pass takes the channels 16 at a time
    import synth
    _check_epoch(ctx, oracle, synth.eeg_epoch(n, T, seed, gamma40=True), K, n_imf, S)


@pytest.mark.parametrize("gen,n,T,K,n_imf,S,seed", [
    ("sines", 8, 1000, 16, 12, 50, 20),   # the 8 x 1000, K=16 host case with [0.075, 0.75, 0.075]
    ("sines", 1, 700, 4, 10, 30, 21),     # univariate
    ("eeg", 4, 2000, 8, 12, 20, 22),
    ("sines", 3, 64, 6, 8, 10, 23),       # short: the decomposition ends early (too few extrema)
    ("eeg", 32, 4096, 64, 12, 10, 24),    # headline shape
    ("eeg", 128, 1024, 256, 4, 6, 25),    # stop mode at its 128-channel limit
])
def test_stop_mode_matches_oracle(ctx, oracle, gen, n, T, K, n_imf, S, seed):
    import synth
    y = synth.sines_epoch(n, T, seed) if gen == "sines" else synth.eeg_epoch(n, T, seed)
    _check_epoch(ctx, oracle, y, K, n_imf, S, mode=1)


def test_stop_mode_batch_matches_alone(ctx, oracle):
    import synth
    eps = []
    for i, (n, T, K) in enumerate([(8, 1500, 16), (4, 900, 8), (16, 1200, 32)]):
        y = synth.eeg_epoch(n, T, 300 + i)
        eps.append({"y": _dev(y), "V": _dev(oracle.directions(n, K)), "n_imf": 10, "S": 15})
    outs, info = ctx.run_batch(eps, mode=1)
    for i, (ep, out) in enumerate(zip(eps, outs)):
        alone, ainfo = ctx.run(ep["y"], ep["V"], 10, 15, mode=1)
        used = ainfo["n_imf_out"]
        assert info["n_imf_out"][i] == used
        assert np.array_equal(info["sift_counts"][i], ainfo["sift_counts"])
        assert np.array_equal(out.cpu().numpy()[:used], alone.cpu().numpy()[:used])


def test_full_window(ctx, oracle):
    import synth
    _check_epoch(ctx, oracle, synth.sines_epoch(4, 4096, 8), 8, 3, 3, window=1)


def test_graph_and_direct_launch_agree(ctx, oracle):
    import synth
    y = synth.eeg_epoch(8, 3000, 9)
    a = _check_epoch(ctx, oracle, y, 16, 4, 3, use_graph=True)
    b = _check_epoch(ctx, oracle, y, 16, 4, 3, use_graph=False)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n_imf,S", [(2, 17), (4, 6), (3, 16), (5, 4)])
def test_graph_steps_per_replay(ctx, oracle, n_imf, S):
    # fix mode captures the largest divisor (up to 16) of the (n_imf - 1) * S steps per
    # graph: 17 steps -> 1 per replay, 18 -> 9, 32 -> 16, 16 -> 16; each against the
    # oracle and against direct launches
    import synth
    y = synth.sines_epoch(8, 2000, 90 + S)
    print(f"graph steps n_imf={n_imf} S={S}", flush=True)
    a = _check_epoch(ctx, oracle, y, 16, n_imf, S, use_graph=True)
    b = _check_epoch(ctx, oracle, y, 16, n_imf, S, use_graph=False)
    assert np.array_equal(a, b)


def test_graph_replays_only_for_an_equal_plan(ctx, oracle):
    # consecutive single-epoch calls with the same launch geometry (block counts,
    # grids, list capacities) but different workspace layouts: T = 1000 and 1010 move
    # every region after the signal, 10 and 70 output slots move the plan region.
    # Each call must capture its own step instead of replaying one whose kernel
    # arguments point into the previous layout.
    import synth
    V = oracle.directions(8, 16)
    for T, n_imf in [(1000, 10), (1010, 10), (1010, 70), (1000, 10)]:
        y = synth.sines_epoch(8, T, 70 + T)
        print(f"graph case T={T} n_imf={n_imf}", flush=True)
        imfs_o, info_o = oracle.memd(y, V, n_imf, 2)
        imfs_g, info_g = ctx.run(_dev(y), _dev(V), n_imf, 2, use_graph=True)
        assert np.array_equal(info_g["sift_counts"], info_o["sift_counts"])
        assert np.array_equal(imfs_g.cpu().numpy(), imfs_o, equal_nan=True)


def test_eeg_sample_reduced_schedule(ctx, oracle):
    # the reference's own input file, 32 x 30504, K = 64 (2 IMFs x 2 sifts to keep the oracle fast)
    import synth
    y = synth.load_eeg_sample(os.path.join(GOLD_DIR, "eeg_sample_32x30504.npz")).astype(np.float64)
    _check_epoch(ctx, oracle, y, 64, 3, 2)


def test_constant_epoch_gives_nan_like_reference(ctx, oracle):
    # no interior extrema -> no valid direction -> 0/0 (memReducedEpochs_dev.cu:862)
    y = np.ones((4, 256))
    V = oracle.directions(4, 8)
    imfs_o, _ = oracle.memd(y, V, 3, 2)
    imfs_g, _ = ctx.run(_dev(y), _dev(V), 3, 2)
    got = imfs_g.cpu().numpy()
    assert np.array_equal(np.isnan(got), np.isnan(imfs_o))
    assert np.isnan(got[0]).all()


def test_f32_entry_point(ctx, oracle):
    import synth
    torch = _torch()
    y32 = synth.sines_epoch(8, 1000, 10).astype(np.float32)
    V32 = oracle.directions(8, 16).astype(np.float32)
    imfs_o, _ = oracle.memd(y32.astype(np.float64), V32.astype(np.float64), 5, 4)
    out, _ = ctx.run(_dev(y32), _dev(V32), 5, 4)
    assert out.dtype == torch.float32
    assert np.array_equal(out.cpu().numpy(), imfs_o.astype(np.float32))


def test_reference_named_entry_points(oracle):
    import torch
    import memd_amd
    import synth
    n, T, K, N_imf, S = 8, 1000, 16, 4, 3
    y = synth.sines_epoch(n, T, 11)
    d_V = torch.empty((K, n), dtype=torch.float64, device="cuda:0")
    memd_amd.GenerateDirectionVector(n, K, d_V)
    d_x = torch.arange(T, dtype=torch.int32, device="cuda:0")
    d_IMFs = torch.empty((N_imf, n, T), dtype=torch.float64, device="cuda:0")
    secs = memd_amd.memd(n, T, K, N_imf, S, d_x, _dev(y), d_V, d_IMFs)
    assert secs > 0
    imfs_o, _ = oracle.memd(y, oracle.directions(n, K), N_imf, S)
    assert np.array_equal(d_IMFs.cpu().numpy(), imfs_o)


def test_mex_layout(ctx, oracle):
    # assistMatlabCuda.m: single data, int32 index 0..T-1, primes(2000)(1:n), output T x n x N_imf
    import torch
    import synth
    n, T, K, N_imf, S = 6, 900, 12, 4, 3
    y32 = synth.sines_epoch(n, T, 12).astype(np.float32)
    primes = np.array([2, 3, 5, 7, 11, 13], np.int32)
    out = torch.zeros((N_imf, n, T), dtype=torch.float32, device="cuda:0")
    idx = torch.arange(T, dtype=torch.int32, device="cuda:0")
    ctx.mex_compat(n, K, N_imf, S, _dev(y32), idx, _dev(primes), out)
    V = oracle.directions(n, K, primes)
    imfs_o, _ = oracle.memd(y32.astype(np.float64), V, N_imf, S)
    assert np.array_equal(out.cpu().numpy(), imfs_o.astype(np.float32))


def test_bad_coordinates_rejected(ctx, oracle):
    import torch
    import memd_amd
    y = np.random.default_rng(0).standard_normal((2, 100))
    V = oracle.directions(2, 4)
    x = torch.arange(100, dtype=torch.int32, device="cuda:0")
    x[5] = 7
    with pytest.raises(memd_amd.MemdError) as ei:
        ctx.run(_dev(y), _dev(V), 3, 2, x=x)
    assert ei.value.status == 4


def test_batch_bit_identical_to_alone(ctx, oracle):
    import synth
    eps = []
    for i, (n, T, K) in enumerate([(8, 1200, 16), (16, 1000, 32), (4, 2500, 8), (8, 1100, 16)]):
        y = synth.eeg_epoch(n, T, 100 + i)
        eps.append({"y": _dev(y), "V": _dev(oracle.directions(n, K)), "n_imf": 4, "S": 3})
    outs, info = ctx.run_batch(eps)
    for ep, out in zip(eps, outs):
        alone, _ = ctx.run(ep["y"], ep["V"], 4, 3)
        assert np.array_equal(out.cpu().numpy(), alone.cpu().numpy())
    # and the first one against the oracle
    y0 = eps[0]["y"].cpu().numpy()
    imfs_o, _ = oracle.memd(y0, eps[0]["V"].cpu().numpy(), 4, 3)
    assert np.array_equal(outs[0].cpu().numpy(), imfs_o)


def test_batch_many_epochs_xcd_groups(ctx, oracle):
    # more than 8 epochs: blocks and work items are pinned to XCD groups
    import synth
    rng = np.random.default_rng(11)
    eps = []
    for i in range(11):
        n = int(rng.choice([2, 4, 8]))
        T = int(rng.integers(600, 1800))
        y = synth.sines_epoch(n, T, 400 + i)
        eps.append({"y": _dev(y), "V": _dev(oracle.directions(n, 2 * n)), "n_imf": int(rng.integers(3, 6)),
                    "S": int(rng.integers(2, 4))})
    outs, info = ctx.run_batch(eps)
    for ep, out in zip(eps, outs):
        alone, _ = ctx.run(ep["y"], ep["V"], ep["n_imf"], ep["S"])
        assert np.array_equal(out.cpu().numpy(), alone.cpu().numpy())


@pytest.mark.parametrize("mode", [0, 1])
def test_batch_mixing_wide_and_narrow_epochs(ctx, oracle, mode):
    # epochs above and below the 16-channel limit of the whole-tile projection path in
    # one batch: each takes its own path (the batch's widest epoch must not decide it)
    import synth
    eps = []
    for i, (n, T, K) in enumerate([(96, 700, 12), (16, 1300, 32), (64, 900, 16), (8, 1100, 16), (72, 600, 8)]):
        eps.append({"y": _dev(synth.eeg_epoch(n, T, 520 + i)), "V": _dev(oracle.directions(n, K)),
                    "n_imf": 4, "S": 3})
    outs, _ = ctx.run_batch(eps, mode=mode)
    for ep, out in zip(eps, outs):
        alone, _ = ctx.run(ep["y"], ep["V"], 4, 3, mode=mode)
        assert np.array_equal(out.cpu().numpy(), alone.cpu().numpy())
    y1, V1 = eps[1]["y"].cpu().numpy(), eps[1]["V"].cpu().numpy()
    imfs_o, info_o = oracle.memd(y1, V1, 4, 3, mode=mode)
    used = info_o["n_imf_out"]
    assert np.array_equal(outs[1].cpu().numpy()[:used], imfs_o[:used])


def test_directions_f32_is_rounded_f64(ctx, oracle):
    # GenerateDirectionVector<float>, as the reference's main() instantiates it
    import torch
    out = torch.empty((64, 32), dtype=torch.float32, device="cuda:0")
    ctx.directions(32, 64, out=out)
    assert np.array_equal(out.cpu().numpy(), oracle.directions(32, 64).astype(np.float32))


@pytest.mark.parametrize("gen,n,T,K,n_imf,S,seed,mode", [
    ("sines", 8, 1000, 16, 5, 4, 30, 0),
    ("eeg", 4, 5000, 8, 3, 3, 31, 0),      # many solver blocks per direction
    ("sines", 1, 600, 4, 4, 3, 32, 0),
    ("eeg", 32, 4096, 64, 3, 2, 33, 0),
    ("eeg", 8, 1500, 16, 8, 12, 34, 1),    # stop mode reads the record array
])
def test_not_a_knot_matches_oracle(ctx, oracle, gen, n, T, K, n_imf, S, seed, mode):
    import synth
    y = synth.sines_epoch(n, T, seed) if gen == "sines" else synth.eeg_epoch(n, T, seed)
    V = oracle.directions(n, K)
    imfs_o, info_o = oracle.memd(y, V, n_imf, S, mode=mode, spline=oracle.SPLINE_NOT_A_KNOT)
    imfs_g, info_g = ctx.run(_dev(y), _dev(V), n_imf, S, mode=mode, spline=1)
    used = info_o["n_imf_out"]
    assert info_g["n_imf_out"] == used
    assert np.array_equal(info_g["sift_counts"], info_o["sift_counts"])
    got = imfs_g.cpu().numpy()[:used]
    mism = np.argwhere(got != imfs_o[:used])
    assert mism.size == 0, f"{len(mism)} samples differ, first at {mism[0].tolist()}"


def test_not_a_knot_batch_matches_alone(ctx, oracle):
    import synth
    eps = []
    for i, (n, T, K) in enumerate([(8, 1200, 16), (4, 2600, 8), (16, 900, 32)]):
        eps.append({"y": _dev(synth.eeg_epoch(n, T, 40 + i)), "V": _dev(oracle.directions(n, K)),
                    "n_imf": 4, "S": 3})
    outs, _ = ctx.run_batch(eps, spline=1)
    for ep, out in zip(eps, outs):
        alone, _ = ctx.run(ep["y"], ep["V"], 4, 3, spline=1)
        assert np.array_equal(out.cpu().numpy(), alone.cpu().numpy())


def test_stop_mode_beyond_the_record_path_cap(ctx, oracle):
    # stop mode with the reference's ends evaluates in the fused kernel: no 128-channel
    # cap (160 channels: 10 blocks of 16, the last a partial group of 4)
    import synth
    y = synth.eeg_epoch(158, 700, 26)
    _check_epoch(ctx, oracle, y, 24, 3, 4, mode=1)


def test_stop_mode_ignores_wider_tile_tunings(ctx, oracle):
    # the amplitude's 16-channel blocks take 64-sample tiles whatever the tuning says
    import synth
    y = synth.eeg_epoch(8, 1500, 27)
    V = oracle.directions(8, 16)
    want, winfo = oracle.memd(y, V, 6, 8, mode=1)
    ctx.set_tuning(8, 4, 2)
    try:
        got, info = ctx.run(_dev(y), _dev(V), 6, 8, mode=1)
    finally:
        ctx.set_tuning(8, 4, 1)
    used = winfo["n_imf_out"]
    assert info["n_imf_out"] == used and np.array_equal(info["sift_counts"], winfo["sift_counts"])
    assert np.array_equal(got.cpu().numpy()[:used], want[:used])


def test_batch_with_a_spread_epoch_equals_alone(ctx, oracle):
    # one epoch with more than an XCD's share of the batch's work is spread over all
    # XCDs (its directions dealt over the sub-lists); placement changes no result
    import synth
    shapes = [(32, 20000, 64)] + [(4, 900 + 50 * i, 8) for i in range(9)]
    eps = [{"y": _dev(synth.eeg_epoch(n, T, 120 + i)), "V": _dev(oracle.directions(n, K)), "n_imf": 4, "S": 3}
           for i, (n, T, K) in enumerate(shapes)]
    outs, _ = ctx.run_batch(eps)
    for ep, out in zip(eps, outs):
        alone, _ = ctx.run(ep["y"], ep["V"], 4, 3)
        assert np.array_equal(out.cpu().numpy(), alone.cpu().numpy())
