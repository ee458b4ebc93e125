"""Tile shapes and the autotuner, through the C ABI (needs an MI355X).

The block solve's row groups (block_groups) and the evaluation tile
(eval_tile_words) change only how work is cut into workgroups, never the
arithmetic: every shape must give bit-identical IMFs, and the default shape is
pinned against the FP64 oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# channel counts around the evaluation tile widths (4, 8, 16 channels) and ragged
# channel groups; T long enough for several solver blocks per direction
SHAPES = [(3, 2600, 6), (8, 1500, 16), (12, 1800, 24), (16, 1200, 32), (24, 1000, 48)]


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


@pytest.fixture
def tctx():
    # a context of its own: installed tunings stay with it
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import memd_amd
    c = memd_amd.MemdContext(0)
    yield c
    c.close()


def _batch(oracle, n_imf=4, S=3):
    import synth
    return [{"y": _dev(synth.eeg_epoch(n, T, 700 + i)), "V": _dev(oracle.directions(n, K)), "n_imf": n_imf, "S": S}
            for i, (n, T, K) in enumerate(SHAPES)]


@pytest.mark.parametrize("mode,spline", [(0, 0), (0, 1), (1, 0)])
def test_every_tile_shape_is_bit_identical(tctx, oracle, mode, spline):
    eps = _batch(oracle)
    ref, _ = tctx.run_batch(eps, mode=mode, spline=spline)
    ref = [r.cpu().numpy() for r in ref]
    # the default shapes against the oracle (first epoch)
    y0, V0 = eps[0]["y"].cpu().numpy(), eps[0]["V"].cpu().numpy()
    imfs_o, info_o = oracle.memd(y0, V0, 4, 3, mode=mode,
                                 spline=oracle.SPLINE_NOT_A_KNOT if spline else oracle.SPLINE_NATURAL)
    used = info_o["n_imf_out"]
    assert np.array_equal(ref[0][:used], imfs_o[:used])
    for bg in (1, 2, 4, 8):
        for tw in (1, 2, 4):
            for n, _, _ in SHAPES:
                tctx.set_tuning(n, bg, tw)
                got = tctx.get_tuning(n)
                assert (got["block_groups"], got["eval_tile_words"]) == (bg, tw)
            outs, _ = tctx.run_batch(eps, mode=mode, spline=spline)
            for i, (o, r) in enumerate(zip(outs, ref)):
                assert np.array_equal(o.cpu().numpy(), r), f"epoch {i} differs at block_groups={bg} eval_tile_words={tw}"


def test_autotune_installs_a_valid_choice(tctx, oracle):
    import synth
    t = tctx.autotune(16, 2000, 32)
    assert t["n"] == 16 and 1 <= t["block_groups"] <= 8 and t["eval_tile_words"] in (1, 2, 4)
    assert t["block_ms"] > 0 and t["eval_ms"] > 0
    got = tctx.get_tuning(16)
    assert (got["block_groups"], got["eval_tile_words"]) == (t["block_groups"], t["eval_tile_words"])
    # other channel counts keep the built-in shapes
    assert tctx.get_tuning(24)["block_ms"] == 0.0
    y = synth.eeg_epoch(16, 1700, 77)
    V = oracle.directions(16, 32)
    imfs_g, _ = tctx.run(_dev(y), _dev(V), 4, 3)
    imfs_o, info_o = oracle.memd(y, V, 4, 3)
    used = info_o["n_imf_out"]
    assert np.array_equal(imfs_g.cpu().numpy()[:used], imfs_o[:used])


def test_set_tuning_rejects_bad_shapes(tctx):
    import memd_amd
    for args in [(0, 1, 1), (300, 1, 1), (16, 0, 1), (16, 9, 1), (16, 2, 3), (256, 8, 1)]:
        with pytest.raises(memd_amd.MemdError) as ei:
            tctx.set_tuning(*args)
        assert ei.value.status == 1   # MEMD_ERR_ARG
    before = tctx.get_tuning(16)
    with pytest.raises(memd_amd.MemdError):
        tctx.set_tuning(16, 2, 5)
    assert tctx.get_tuning(16) == before
