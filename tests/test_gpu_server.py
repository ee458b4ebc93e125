"""Served continuous batching through the C ABI (needs an MI355X).

Epochs join and leave the active set between sift steps; each result must be
bit-identical to the same epoch decomposed alone (and, for one of them, to the
FP64 oracle), whatever else was active while it ran.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _drain(srv, submit_later=(), max_calls=5000, steps=1):
    """Step until every request has ended; submit_later[i] is submitted before call i."""
    later = list(submit_later)
    ids = []
    for call in range(max_calls):
        if later:
            ids.append(later.pop(0)())
        active, queued = srv.step(steps)
        ids_done = srv.completed()
        assert all(srv.poll(i) >= 2 for i in ids_done)   # done, cancelled or timed out
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


def test_server_cancel_queued_and_running(ctx, oracle):
    import memd_amd
    import synth
    lib = memd_amd.load_library()
    ws = lib.memd_workspace_bytes(8, 3000, 16)
    srv = memd_amd.MemdServer(ctx, pool_bytes=ws)   # one epoch at a time
    try:
        data = [(_dev(synth.eeg_epoch(8, 3000, 900 + i)), _dev(oracle.directions(8, 16)), 6, 8) for i in range(3)]
        a, b, c = (srv.submit(*d) for d in data)
        srv.step(1)
        assert srv.poll(a) == memd_amd.REQ_RUNNING and srv.poll(b) == memd_amd.REQ_QUEUED
        srv.cancel(b)                                   # queued: removed at once
        assert srv.poll(b) == memd_amd.REQ_CANCELLED
        srv.cancel(a)                                   # running: the device skips it from here on
        assert srv.poll(a) == memd_amd.REQ_CANCELLED and srv.pool_used() == 0
        out_a = srv._req[a][2]
        frozen = out_a.cpu().numpy().copy()
        ids = _drain(srv)
        assert ids == []
        assert srv.poll(c) == memd_amd.REQ_DONE
        assert np.array_equal(out_a.cpu().numpy(), frozen)   # untouched while c ran
        _check_alone(ctx, srv, c, *data[2])
        with pytest.raises(memd_amd.MemdError) as ei:
            srv.cancel(c)
        assert ei.value.status == 1                      # already ended
        for rid in (a, b, c):
            srv.release(rid)
        info = srv.info()
        assert info["withdrawn"] == 2 and info["retired"] == 1
    finally:
        srv.close()


def test_server_deadlines(ctx, oracle):
    import memd_amd
    import synth
    srv = memd_amd.MemdServer(ctx, pool_bytes=8 << 30, max_active=1)
    try:
        long_y = _dev(synth.sines_epoch(32, 60000, 950))
        V = _dev(oracle.directions(32, 64))
        slow = srv.submit(long_y, V, 9, 10)
        waiting = srv.submit(long_y, V, 9, 10)           # queued behind it (one active at a time)
        short_y = _dev(synth.sines_epoch(4, 800, 951))
        Vs = _dev(oracle.directions(4, 8))
        quick = srv.submit(short_y, Vs, 3, 2)
        srv.step(1)                                      # one step of `slow`; the others wait
        assert srv.poll(slow) == memd_amd.REQ_RUNNING and srv.poll(waiting) == memd_amd.REQ_QUEUED
        srv.set_deadline(slow, 1e-6)                     # both expire before the next call
        srv.set_deadline(waiting, 1e-6)
        srv.set_deadline(quick, 600.0)
        time.sleep(0.01)
        _drain(srv)
        assert srv.poll(slow) == memd_amd.REQ_TIMED_OUT
        assert srv.poll(waiting) == memd_amd.REQ_TIMED_OUT
        assert srv.poll(quick) == memd_amd.REQ_DONE
        _check_alone(ctx, srv, quick, short_y, Vs, 3, 2)
        assert srv.pool_used() == 0
    finally:
        srv.close()


def test_server_replays_one_graph_per_plan(ctx, oracle):
    import memd_amd
    import synth
    data = [(_dev(synth.sines_epoch(4, 1000 + 100 * i, 960 + i)), _dev(oracle.directions(4, 8)), 5, 4)
            for i in range(6)]
    srv = memd_amd.MemdServer(ctx, pool_bytes=1 << 30)
    try:
        ids = [srv.submit(*d) for d in data]
        calls = 0
        while True:
            active, queued = srv.step(1)
            calls += 1
            if active == 0 and queued == 0:
                break
        info = srv.info()
        # everything joined at once: one plan (and graph) for the whole run, one probe
        # per call that queued steps, 4 x 4 = 16 sift steps
        assert info["rebuilds"] == 1
        assert info["steps"] == info["probes"] and info["steps"] >= 16
        assert info["retired"] == 6
        for rid, d in zip(ids, data):
            _check_alone(ctx, srv, rid, *d)
    finally:
        srv.close()


def test_direction_cache_evicts_least_recently_used(ctx, oracle):
    import memd_amd
    c = memd_amd.MemdContext(0)
    try:
        c.set_direction_cache(3 * 64 * 32 * 8)           # room for three 32 x 64 sets
        for K in (64, 65, 66, 67):
            got = c.directions(32, K).cpu().numpy()
            assert np.array_equal(got, oracle.directions(32, K))
            assert c.direction_cache_bytes() <= 3 * 64 * 32 * 8 + 3 * 32 * 8
        c.set_direction_cache(0)                         # evicts everything
        assert c.direction_cache_bytes() == 0
        assert np.array_equal(c.directions(8, 16).cpu().numpy(), oracle.directions(8, 16))
        assert c.direction_cache_bytes() == 8 * 16 * 8   # a set larger than the cap is still kept
    finally:
        c.close()


def test_injected_allocation_failures_report_nomem(ctx, oracle):
    import memd_amd
    import synth
    y = _dev(synth.sines_epoch(4, 900, 970))
    V = _dev(oracle.directions(4, 8))
    c = memd_amd.MemdContext(0)   # a fresh context: its workspace is not allocated yet
    try:
        c.fail_allocations(1)
        with pytest.raises(memd_amd.MemdError) as ei:
            c.run(y, V, 3, 2)
        assert ei.value.status == 3                      # MEMD_ERR_NOMEM
        got, _ = c.run(y, V, 3, 2)                       # the next call allocates normally
        want, _ = ctx.run(y, V, 3, 2)
        assert np.array_equal(got.cpu().numpy(), want.cpu().numpy())
        c.fail_allocations(1)
        with pytest.raises(memd_amd.MemdError) as ei:
            c.directions(5, 9)
        assert ei.value.status == 3
        assert np.array_equal(c.directions(5, 9).cpu().numpy(), oracle.directions(5, 9))
        c.fail_allocations(1)
        with pytest.raises(memd_amd.MemdError) as ei:
            memd_amd.MemdServer(c, pool_bytes=1 << 28)
        assert ei.value.status == 3
        # a server whose plan region cannot be allocated keeps the request queued
        srv = memd_amd.MemdServer(c, pool_bytes=1 << 28)
        try:
            rid = srv.submit(y, V, 3, 2)
            c.fail_allocations(1)
            with pytest.raises(memd_amd.MemdError) as ei:
                srv.step(1)
            assert ei.value.status == 3 and srv.poll(rid) == memd_amd.REQ_QUEUED and srv.pool_used() == 0
            _drain(srv)
            assert srv.poll(rid) == memd_amd.REQ_DONE
            assert np.array_equal(srv.result(rid)[0].cpu().numpy(), want.cpu().numpy())
        finally:
            srv.close()
    finally:
        c.close()
