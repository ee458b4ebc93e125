"""Mirrored ends (memd.m's boundary_conditions, nbsym = 2) on the GPU (needs an MI355X).

The option is not in the reference; its oracle restatement is pinned on the CPU
(tests/test_oracle_mirror.py).  Bar, as for every other path: knot counts per
direction per pass, sift counts, IMF count and every IMF sample identical to the
oracle's, in fix and stop mode, with natural and not-a-knot ends, alone, batched
and served.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIRROR = 1


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _gen(gen, n, T, seed):
    import synth
    return synth.sines_epoch(n, T, seed) if gen == "sines" else synth.eeg_epoch(n, T, seed)


@pytest.mark.parametrize("gen,n,T,K,n_imf,S,seed,mode,spline", [
    ("sines", 8, 1000, 16, 5, 4, 50, 0, 0),
    ("sines", 1, 600, 4, 4, 3, 51, 0, 1),        # univariate, not-a-knot
    ("eeg", 4, 5000, 8, 3, 3, 52, 0, 0),         # many solver blocks per direction
    ("eeg", 32, 4096, 64, 3, 2, 53, 0, 0),       # the 32 x 4096 K=64 shape
    ("eeg", 8, 1500, 16, 8, 12, 54, 1, 1),       # memd.m's setting: stop criterion, not-a-knot
    ("sines", 3, 64, 6, 8, 10, 55, 1, 0),        # short: ends by stop_emd
    ("eeg", 128, 1024, 256, 3, 2, 56, 0, 0),     # the record path's 128-channel limit
    ("eeg", 16, 30504, 32, 3, 2, 57, 0, 0),      # long epoch: the extrema window's gaps near the end
])
def test_mirror_matches_oracle(ctx, oracle, gen, n, T, K, n_imf, S, seed, mode, spline):
    y = _gen(gen, n, T, seed)
    V = oracle.directions(n, K)
    print(f"mirror {gen} n={n} T={T} K={K} {n_imf}x{S} mode={mode} spline={spline}", flush=True)
    imfs_o, info_o = oracle.memd(y, V, n_imf, S, mode=mode, spline=spline, boundary=oracle.BOUNDARY_MIRROR,
                                 threads=16)
    cap = (n_imf - 1) * (S + 1) + 1
    imfs_g, info_g = ctx.run(_dev(y), _dev(V), n_imf, S, mode=mode, spline=spline, boundary=MIRROR,
                             trace_passes=cap)
    used = info_o["n_imf_out"]
    assert info_g["n_imf_out"] == used
    assert np.array_equal(info_g["sift_counts"], info_o["sift_counts"])
    passes = info_o["n_passes"]
    assert np.array_equal(info_g["knot_counts"].cpu().numpy()[:passes], info_o["knot_counts"])
    got = imfs_g.cpu().numpy()[:used]
    mism = np.argwhere(got != imfs_o[:used])
    assert mism.size == 0, f"{len(mism)} samples differ, first at {mism[0].tolist()}"
    # the option changes the decomposition (the ends are not the reference's)
    clamp, _ = oracle.memd(y, V, n_imf, S, mode=mode, spline=spline, threads=16)
    assert not np.array_equal(clamp[:used], imfs_o[:used])


@pytest.mark.parametrize("mode", [0, 1])
def test_mirror_batch_matches_alone(ctx, oracle, mode):
    import synth
    eps = []
    for i, (n, T, K) in enumerate([(8, 1500, 16), (4, 900, 8), (16, 1200, 32), (2, 3000, 4)]):
        eps.append({"y": _dev(synth.eeg_epoch(n, T, 60 + i)), "V": _dev(oracle.directions(n, K)),
                    "n_imf": 6, "S": 5})
    outs, info = ctx.run_batch(eps, mode=mode, boundary=MIRROR)
    for i, (ep, out) in enumerate(zip(eps, outs)):
        alone, ainfo = ctx.run(ep["y"], ep["V"], 6, 5, mode=mode, boundary=MIRROR)
        used = ainfo["n_imf_out"]
        assert info["n_imf_out"][i] == used
        assert np.array_equal(info["sift_counts"][i], ainfo["sift_counts"])
        assert np.array_equal(out.cpu().numpy()[:used], alone.cpu().numpy()[:used])


def test_mirror_served_matches_alone(ctx, oracle):
    import memd_amd
    import synth
    shapes = [(8, 1500, 16, 4, 3), (4, 900, 8, 3, 2), (16, 1200, 32, 3, 3), (2, 2000, 4, 5, 2)]
    data = [(_dev(synth.eeg_epoch(n, T, 70 + i)), _dev(oracle.directions(n, K)), n_imf, S)
            for i, (n, T, K, n_imf, S) in enumerate(shapes)]
    srv = memd_amd.MemdServer(ctx, pool_bytes=1 << 30, boundary=MIRROR)
    try:
        ids = [srv.submit(*d) for d in data[:2]]
        later = [lambda d=d: srv.submit(*d) for d in data[2:]]
        for _ in range(5000):
            if later:
                ids.append(later.pop(0)())
            active, queued = srv.step(1)
            srv.completed()
            if not later and active == 0 and queued == 0:
                break
        for rid, (y, V, n_imf, S) in zip(ids, data):
            out, info = srv.result(rid)
            alone, ainfo = ctx.run(y, V, n_imf, S, boundary=MIRROR)
            assert np.array_equal(info["sift_counts"], ainfo["sift_counts"])
            assert np.array_equal(out.cpu().numpy(), alone.cpu().numpy())
    finally:
        srv.close()


def test_mirror_limits(ctx, oracle):
    import memd_amd
    y = _dev(np.zeros((129, 512)))
    V = _dev(oracle.directions(129, 4))
    with pytest.raises(memd_amd.MemdError) as ei:
        ctx.run(y, V, 2, 1, boundary=MIRROR)
    assert ei.value.status == 5   # MEMD_ERR_UNSUPPORTED: the record path holds 128 channels
    with pytest.raises(memd_amd.MemdError) as ei:
        ctx.run(_dev(np.zeros((2, 512))), _dev(oracle.directions(2, 4)), 2, 1, boundary=7)
    assert ei.value.status == 1   # MEMD_ERR_ARG: unknown end treatment
