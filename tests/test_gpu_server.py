"""Served continuous batching through the C ABI (needs an MI355X).

Epochs join and leave the active set between sift steps; each result must be
bit-identical to the same epoch decomposed alone (and, for one of them, to the
FP64 oracle), whatever else was active while it ran.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _drain(srv, submit_later=(), max_calls=5000, steps=1):
    """Step until every request is done; submit_later[i] is submitted before call i."""
    later = list(submit_later)
    ids = []
    for call in range(max_calls):
        if later:
            ids.append(later.pop(0)())
        active, queued = srv.step(steps)
        ids_done = srv.completed()
        assert all(srv.poll(i) == 2 for i in ids_done)
        if not later and active == 0 and queued == 0:
            return ids
    raise AssertionError("server did not drain")


def _check_alone(ctx, srv, rid, y, V, n_imf, S, mode=0):
    out, info = srv.result(rid)
    alone, ainfo = ctx.run(y, V, n_imf, S, mode=mode)
    used = ainfo["n_imf_out"]
    assert info["n_imf_out"] == used
    assert np.array_equal(info["sift_counts"], ainfo["sift_counts"])
    assert np.array_equal(out.cpu().numpy()[:used], alone.cpu().numpy()[:used])


def test_server_join_retire_bit_identical(ctx, oracle):
    import memd_amd
    import synth
    shapes = [(8, 1500, 16, 5, 3), (4, 900, 8, 3, 2), (16, 1200, 32, 4, 4), (8, 2000, 16, 6, 2),
              (2, 700, 4, 3, 5), (32, 1100, 64, 3, 2)]
    data = []
    for i, (n, T, K, n_imf, S) in enumerate(shapes):
        y = _dev(synth.eeg_epoch(n, T, 500 + i))
        V = _dev(oracle.directions(n, K))
        data.append((y, V, n_imf, S))
    ws = [memd_amd.load_library().memd_workspace_bytes(n, T, K) for (n, T, K, _, _) in shapes]
    # room for about three of them at once: admissions wait for retirements
    srv = memd_amd.MemdServer(ctx, pool_bytes=sorted(ws)[-1] + sorted(ws)[-2] + sorted(ws)[-3])
    try:
        ids = [srv.submit(*data[0]), srv.submit(*data[1])]
        ids += _drain(srv, [lambda d=d: srv.submit(*d) for d in data[2:]])
        assert srv.pool_used() == 0
        for rid, d in zip(ids, data):
            _check_alone(ctx, srv, rid, *d)
        # and one against the oracle
        y0, V0, n_imf0, S0 = data[0]
        imfs_o, _ = oracle.memd(y0.cpu().numpy(), V0.cpu().numpy(), n_imf0, S0)
        assert np.array_equal(srv.result(ids[0])[0].cpu().numpy(), imfs_o)
        for rid in ids:
            srv.release(rid)
    finally:
        srv.close()


def test_server_many_epochs_xcd_groups(ctx, oracle):
    # more than 8 active epochs: the plan pins epochs to XCD groups
    import memd_amd
    import synth
    rng = np.random.default_rng(7)
    data = []
    for i in range(14):
        n = int(rng.choice([2, 4, 8]))
        T = int(rng.integers(600, 1600))
        y = _dev(synth.sines_epoch(n, T, 600 + i))
        V = _dev(oracle.directions(n, 2 * n))
        data.append((y, V, int(rng.integers(3, 6)), int(rng.integers(2, 4))))
    srv = memd_amd.MemdServer(ctx, pool_bytes=1 << 30, max_active=10)
    try:
        ids = [srv.submit(*d) for d in data]
        _drain(srv, steps=2)
        for rid, d in zip(ids, data):
            _check_alone(ctx, srv, rid, *d)
    finally:
        srv.close()


def test_server_stop_mode(ctx, oracle):
    import memd_amd
    import synth
    data = []
    for i, (n, T, K) in enumerate([(8, 1500, 16), (4, 900, 8), (16, 1200, 32)]):
        y = _dev(synth.eeg_epoch(n, T, 700 + i))
        data.append((y, _dev(oracle.directions(n, K)), 8, 12))
    srv = memd_amd.MemdServer(ctx, pool_bytes=1 << 30, mode=memd_amd.MODE_STOP)
    try:
        ids = [srv.submit(*data[0])]
        ids += _drain(srv, [lambda d=d: srv.submit(*d) for d in data[1:]])
        for rid, d in zip(ids, data):
            _check_alone(ctx, srv, rid, *d, mode=1)
    finally:
        srv.close()


def test_server_admission_and_ids(ctx, oracle):
    import memd_amd
    import synth
    lib = memd_amd.load_library()
    small = lib.memd_workspace_bytes(4, 800, 8)
    srv = memd_amd.MemdServer(ctx, pool_bytes=small)
    try:
        yb = _dev(synth.sines_epoch(16, 4000, 1))
        with pytest.raises(memd_amd.MemdError) as ei:
            srv.submit(yb, _dev(oracle.directions(16, 32)), 3, 2)
        assert ei.value.status == 6   # can never fit the pool
        with pytest.raises(memd_amd.MemdError) as ei:
            srv.poll(12345)
        assert ei.value.status == 7
        y = _dev(synth.sines_epoch(4, 800, 2))
        V = _dev(oracle.directions(4, 8))
        a = srv.submit(y, V, 3, 2)
        b = srv.submit(y, V, 3, 2)   # queued until a retires (the pool holds one)
        srv.step(1)
        assert srv.poll(a) == 1 and srv.poll(b) == 0
        with pytest.raises(memd_amd.MemdError) as ei:
            srv.release(a)
        assert ei.value.status == 1
        _drain(srv)
        assert srv.poll(a) == 2 and srv.poll(b) == 2
        assert np.array_equal(srv.result(a)[0].cpu().numpy(), srv.result(b)[0].cpu().numpy())
        srv.release(a)
        with pytest.raises(memd_amd.MemdError):
            srv.poll(a)
    finally:
        srv.close()
