"""The allocation-free entry point, memd_run_noalloc_f64, through the C ABI
(needs an MI355X).

It must give memd_run_f64's results bit for bit on two calls in a row, the
second with a smaller epoch (the per-device array keeps the first call's
contents), in fix and stop mode; allocate no device memory; and refuse a shape
whose workspace does not fit its array.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MiB = 1 << 20


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


@pytest.mark.parametrize("mode,cases", [
    (0, [(32, 4096, 64, 9, 10, 1), (8, 1000, 16, 5, 4, 2)]),   # the headline shape, then a smaller epoch
    (1, [(16, 2000, 32, 8, 12, 3), (4, 900, 8, 6, 10, 4)]),
])
def test_noalloc_matches_allocating_entry_point(mode, cases):
    import torch
    import memd_amd
    import synth
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    ref_ctx = memd_amd.MemdContext(0)
    refs, ins = [], []
    for n, T, K, n_imf, S, seed in cases:
        y = _dev(synth.sines_epoch(n, T, seed))
        V = ref_ctx.directions(n, K)
        print(f"allocating entry point n={n} T={T} K={K} {n_imf}x{S} mode={mode}", flush=True)
        # direct launches, as the allocation-free path makes them: the HIP runtime sets up its
        # own launch resources (kernel-argument space) on a process's first direct launches
        ref, info = ref_ctx.run(y, V, n_imf, S, mode=mode, use_graph=False)
        refs.append((ref, info))
        ins.append((y, V, n_imf, S, torch.full((n_imf, n, T), float("nan"), dtype=torch.float64, device="cuda:0")))
    ref_ctx.close()
    ctx = memd_amd.MemdContext(0)   # a context with no workspace of its own
    try:
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        gots = []
        for (y, V, n_imf, S, out), (n, T, K, _, _, _) in zip(ins, cases):
            print(f"allocation-free entry point n={n} T={T} K={K} mode={mode}", flush=True)
            gots.append(ctx.run(y, V, n_imf, S, out=out, mode=mode, noalloc=True))
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        assert free0 - free1 < 16 * MiB, f"device free memory fell by {(free0 - free1) / MiB:.1f} MiB"
        for (got, ginfo), (ref, rinfo) in zip(gots, refs):
            used = rinfo["n_imf_out"]
            assert ginfo["n_imf_out"] == used
            assert np.array_equal(ginfo["sift_counts"], rinfo["sift_counts"])
            assert torch.equal(got[:used], ref[:used])
    finally:
        ctx.close()


def test_noalloc_refuses_shapes_beyond_its_array(ctx, oracle):
    import torch
    import memd_amd
    lib = memd_amd.load_library()
    n, T, K = 64, 8000, 128
    assert lib.memd_workspace_bytes(n, T, K) > memd_amd.NOALLOC_SCRATCH_BYTES
    y = torch.zeros((n, T), dtype=torch.float64, device="cuda:0")
    V = ctx.directions(n, K)
    out = torch.zeros((3, n, T), dtype=torch.float64, device="cuda:0")
    with pytest.raises(memd_amd.MemdError) as ei:
        ctx.run(y, V, 3, 2, out=out, noalloc=True)
    assert ei.value.status == 5   # MEMD_ERR_UNSUPPORTED
    assert "allocation-free" in str(ei.value)
