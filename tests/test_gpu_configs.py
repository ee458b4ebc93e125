"""The measured workloads at their full sizes, through the C ABI (needs an MI355X).

* 128 ch x 16384 samples, K = 256 (about 85 solver blocks per direction and
  W = 256 flag words) against the FP64 oracle;
* the reference demo's own run (memReducedEpochs_dev.cu:1473-1476, :1522): the
  shipped 32 x 30504 EEG sample, K = 64, 9 slots x 10 sifts, against the oracle;
* the 512 uneven epochs of the batched benchmark (n in {8,16,32,64}, T in
  [1000, 8000], K = 2n, 9 x 10) as one batched call, every epoch bit-identical to
  its own plain call, and one epoch of each channel count against the oracle;
* a served mix with 32 x 60000 epochs (window gap 478 samples) joining and
  leaving among short ones, each result bit-identical to the epoch alone, one
  long epoch against the oracle;
* the headline shape (32 x 4096, K = 64, 9 x 10) at the tile shapes the
  benchmark's autotuner picked, against the oracle, and every tile shape at 32
  and 64 channels bit-identical to the built-in one.

Bar: bit-exact (knot counts per pass, sift counts, every IMF sample), with the
1e-9 relative L2 tolerance asserted as a second line.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD_DIR = os.path.join(os.path.dirname(__file__), "golden")
THREADS = max(1, min(16, os.cpu_count() or 1))


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _vs_oracle(ctx, oracle, y, K, n_imf, S, V=None):
    n, T = y.shape
    if V is None:
        V = oracle.directions(n, K)
    print(f"oracle run n={n} T={T} K={K} {n_imf}x{S}", flush=True)
    imfs_o, info_o = oracle.memd(y, V, n_imf, S, threads=THREADS)
    print(f"gpu run n={n} T={T} K={K} {n_imf}x{S}", flush=True)
    cap = (n_imf - 1) * (S + 1) + 1
    imfs_g, info_g = ctx.run(_dev(y), _dev(V), n_imf, S, trace_passes=cap)
    got = imfs_g.cpu().numpy()
    assert info_g["n_imf_out"] == info_o["n_imf_out"]
    assert np.array_equal(info_g["sift_counts"], info_o["sift_counts"])
    kc = info_g["knot_counts"].cpu().numpy()[:info_o["n_passes"]]
    assert np.array_equal(kc, info_o["knot_counts"]), "knot counts differ"
    assert _rel_l2(got, imfs_o) <= 1e-9
    mism = np.argwhere(got != imfs_o)
    assert mism.size == 0, f"{len(mism)} samples differ, first at {mism[0].tolist()}"
    return got


def test_wide_128x16384_k256(ctx, oracle):
    import synth
    y = synth.eeg_epoch(128, 16384, 4000, gamma40=True)
    _vs_oracle(ctx, oracle, y, 256, 3, 2)


def test_reference_demo_run_on_eeg_sample(ctx, oracle):
    # main(): EEGsample_32CH_30504.bin, SignalDim 32, NumDirVector 64, num_IMFs 9, max_iter 10
    import synth
    y = synth.load_eeg_sample(os.path.join(GOLD_DIR, "eeg_sample_32x30504.npz")).astype(np.float64)
    got = _vs_oracle(ctx, oracle, y, 64, 9, 10)
    np.testing.assert_allclose(got.sum(0), y, rtol=0, atol=1e-12 * np.abs(y).max())


def test_uneven_512_batched_equals_plain(ctx, oracle):
    import torch
    import synth
    shapes = synth.batch_shapes(512, 77)
    Vs, eps = {}, []
    for i, (n, T, K) in enumerate(shapes):
        if (n, K) not in Vs:
            Vs[(n, K)] = ctx.directions(n, K)
        eps.append({"y": _dev(synth.eeg_epoch(n, T, 78 + i)), "V": Vs[(n, K)], "n_imf": 9, "S": 10})
    print("batched call over 512 epochs", flush=True)
    outs, info = ctx.run_batch(eps)
    print("plain calls", flush=True)
    for i, (ep, out) in enumerate(zip(eps, outs)):
        alone, ainfo = ctx.run(ep["y"], ep["V"], 9, 10)
        assert np.array_equal(info["sift_counts"][i], ainfo["sift_counts"])
        assert torch.equal(out, alone), f"epoch {i} {shapes[i]} differs from its plain call"
        del alone
    # one epoch of each channel count against the oracle (the one with the median length)
    for n in (8, 16, 32, 64):
        idx = [i for i, s in enumerate(shapes) if s[0] == n]
        i = sorted(idx, key=lambda j: shapes[j][1])[len(idx) // 2]
        print(f"oracle check of epoch {i} {shapes[i]}", flush=True)
        y = eps[i]["y"].cpu().numpy()
        V = eps[i]["V"].cpu().numpy()
        imfs_o, _ = oracle.memd(y, V, 9, 10, threads=THREADS, trace=False)
        got = outs[i].cpu().numpy()
        assert _rel_l2(got, imfs_o) <= 1e-9
        assert np.array_equal(got, imfs_o), f"epoch {i} {shapes[i]} differs from the oracle"


def test_served_mix_with_long_epochs(ctx, oracle):
    # 90% short uneven epochs, 10% 32 x 60000 (K = 64), joining while others run
    import torch
    import memd_amd
    import synth
    rng = np.random.default_rng(12)
    short = synth.batch_shapes(18, 55)
    reqs = []
    for i, (n, T, K) in enumerate(short):
        reqs.append((synth.sines_epoch(n, T, 800 + i), K, 9, 10))
    reqs.insert(3, (synth.sines_epoch(32, 60000, 900), 64, 9, 10))
    reqs.insert(11, (synth.sines_epoch(32, 60000, 901), 64, 3, 2))   # also checked against the oracle
    Vs = {}
    data = []
    for y, K, n_imf, S in reqs:
        n = y.shape[0]
        if (n, K) not in Vs:
            Vs[(n, K)] = ctx.directions(n, K)
        data.append((_dev(y), Vs[(n, K)], n_imf, S))
    srv = memd_amd.MemdServer(ctx, pool_bytes=8 << 30, max_active=8)
    try:
        ids, pending = [], list(range(len(data)))
        calls = 0
        while True:
            # submit 0..3 requests per call: epochs join while others are mid-sift
            for _ in range(int(rng.integers(0, 4))):
                if pending:
                    ids.append(srv.submit(*data[pending.pop(0)]))
            active, queued = srv.step(int(rng.integers(1, 4)))
            srv.completed()
            calls += 1
            if not pending and active == 0 and queued == 0:
                break
            assert calls < 20000, "server did not drain"
        print(f"served {len(ids)} requests in {calls} step calls", flush=True)
        for rid, (y, V, n_imf, S) in zip(ids, data):
            out, info = srv.result(rid)
            alone, ainfo = ctx.run(y, V, n_imf, S)
            assert np.array_equal(info["sift_counts"], ainfo["sift_counts"])
            assert torch.equal(out, alone), f"request {rid} differs from the epoch alone"
        y, V, n_imf, S = data[11]
        imfs_o, _ = oracle.memd(y.cpu().numpy(), V.cpu().numpy(), n_imf, S, threads=THREADS, trace=False)
        assert np.array_equal(srv.result(ids[11])[0].cpu().numpy(), imfs_o)
    finally:
        srv.close()


@pytest.fixture
def tctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import memd_amd
    c = memd_amd.MemdContext(0)
    yield c
    c.close()


def test_headline_shape_at_tuned_tiles(tctx, oracle):
    # the autotuner's choice in the benchmark run: block_groups 4, eval_tile_words 2
    import synth
    tctx.set_tuning(32, 4, 2)
    y = synth.sines_epoch(32, 4096, 1000)
    _vs_oracle(tctx, oracle, y, 64, 9, 10)


@pytest.mark.parametrize("n,T,K", [(32, 4096, 64), (64, 3000, 128)])
def test_every_tile_shape_wide(tctx, oracle, n, T, K):
    import torch
    import memd_amd
    import synth
    y = _dev(synth.eeg_epoch(n, T, 31 + n))
    V = tctx.directions(n, K)
    ref, _ = tctx.run(y, V, 3, 3)
    for bg in (1, 2, 4, 8):
        for tw in (1, 2, 4):
            try:
                tctx.set_tuning(n, bg, tw)
            except memd_amd.MemdError:
                continue   # block_groups that do not fit the LDS at this channel count
            got, _ = tctx.run(y, V, 3, 3)
            assert torch.equal(got, ref), f"block_groups={bg} eval_tile_words={tw} differs"
