"""Lanes (inspire_ctx_create_lane): batches on two lanes from two threads run
concurrently over the owner's resident database and tables, and every
response equals the same batch on the owner alone; writers are refused on a
lane."""
from __future__ import annotations

import sys
import threading

import numpy as np
import pytest

import oracle as O
from helpers import Client, encode_db, random_db

pytestmark = pytest.mark.gpu

W_SEED = bytes(range(32))


def say(*a):
    print("[gpu-lanes]", *a, flush=True)
    sys.stdout.flush()


def test_two_lanes_match_the_owner():
    import inspire_mi355x as M

    say("lanes")
    o = O.Oracle(O.baseline_params())
    d, K = o.params.d, o.params.K
    shards = encode_db(o, random_db(4 * d, 32, seed=81), 32)
    cl = Client(o, 82)
    yb, _ = cl.packing_keys(16, W_SEED, with_y_all=False)
    g = np.random.default_rng(83)
    n, rounds = 6, 4
    batches = []
    for _ in range(2):
        sids = g.integers(0, 4, n).astype(np.uint32)
        rows, seeds = [], []
        for i in range(n):
            s, b = cl.seeded_query(int(g.integers(0, d)))
            rows.append(b[3:6].astype(np.uint32))
            seeds.append(np.frombuffer(s[3:6].tobytes(), dtype="<u4").reshape(3, 8))
        batches.append((sids, np.stack(rows), np.stack(seeds)))
    per = 2 * K * d
    with M.Context(M.InspireParams(), W_SEED) as ctx:
        ctx.load_database(0, shards)
        dev = [(s, M.DeviceArray.from_numpy(r), M.DeviceArray.from_numpy(sd)) for s, r, sd in batches]
        dy = M.DeviceArray.from_numpy(np.stack([yb.astype(np.uint32)] * n))
        want = []
        for sids, dr, ds in dev:  # each batch on the owner alone
            out = M.DeviceArray(n * per * 4)
            assert (ctx.respond_batch_device(0, M.TWO_PACKING, sids, dr.ptr, out.ptr, seeds_ptr=ds.ptr,
                                             y_body_ptr=dy.ptr) == 0).all()
            want.append(out.download(np.zeros((n, per), np.uint32)))
        lanes = [ctx.lane(), ctx.lane()]
        # two outputs per lane, taken in turn, read after the threads end (a
        # copy on the null stream is refused while another lane captures a graph)
        outs = [[M.DeviceArray(n * per * 4) for _ in range(2)] for _ in lanes]
        errors = []

        def run(j):
            try:
                sids, dr, ds = dev[j]
                for r in range(rounds):
                    st = lanes[j].respond_batch_device(0, M.TWO_PACKING, sids, dr.ptr, outs[j][r % 2].ptr,
                                                       seeds_ptr=ds.ptr, y_body_ptr=dy.ptr)
                    assert (st == 0).all()
            except Exception as e:  # reported below
                errors.append(e)

        th = [threading.Thread(target=run, args=(j,)) for j in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errors, errors
        for j in range(2):
            for r in range(2):
                assert np.array_equal(outs[j][r].download(np.zeros((n, per), np.uint32)), want[j]), (j, r)
        # and against the oracle, one query
        sids, rows, seeds = batches[0]
        full = np.zeros((6, 2, K, d), dtype=np.uint64)
        for r in range(3):
            full[3 + r, 0] = o.poly_from_seed(seeds[0, r].astype("<u4").tobytes())
            full[3 + r, 1] = rows[0, r]
        _, ct = o.respond_inspiring(W_SEED, shards[int(sids[0])], full, y_body=yb)
        assert np.array_equal(want[0][0].astype(np.uint64), ct.reshape(-1))
        with pytest.raises(M.PirError) as e:
            lanes[0].load_database(1, shards)
        assert e.value.code == M.ERR_INVALID_ARGUMENT
        for ln in lanes:
            ln.close()


@pytest.mark.parametrize("window_us", [10_000_000, 0])
def test_chain_window_keeps_every_response(window_us):
    """inspire_ctx_set_chain_window: with a window wider than the run every
    chain launch of three lanes waits, on the GPU, for the chain submitted
    before it by another lane (and with 0 none does); over graph replays and
    eager launches alike every response equals the same batch on the owner
    alone, and lanes that close leave the owner's ordering usable."""
    import inspire_mi355x as M

    say("chain window", window_us)
    o = O.Oracle(O.baseline_params())
    d, K = o.params.d, o.params.K
    shards = encode_db(o, random_db(4 * d, 32, seed=91), 32)
    cl = Client(o, 92)
    yb, _ = cl.packing_keys(16, W_SEED, with_y_all=False)
    g = np.random.default_rng(93)
    n, rounds, L = 5, 8, 3
    per = 2 * K * d
    with M.Context(M.InspireParams(), W_SEED) as ctx:
        ctx.load_database(0, shards)
        assert ctx.chain_window() == 60  # the default
        ctx.set_chain_window(window_us)
        assert ctx.chain_window() == window_us
        dy = M.DeviceArray.from_numpy(np.stack([yb.astype(np.uint32)] * n))
        dev, want = [], []
        for _ in range(L):
            sids = g.integers(0, 4, n).astype(np.uint32)
            rows = (g.integers(0, 2**63, (n, 3, K, d), dtype=np.uint64)
                    % np.array([o.params.moduli[0], o.params.moduli[1]], dtype=np.uint64)[:, None]).astype(np.uint32)
            seeds = g.integers(0, 2**32, (n, 3, 8), dtype=np.uint64).astype(np.uint32)
            dr, ds = M.DeviceArray.from_numpy(rows), M.DeviceArray.from_numpy(seeds)
            out = M.DeviceArray(n * per * 4)
            assert (ctx.respond_batch_device(0, M.TWO_PACKING, sids, dr.ptr, out.ptr, seeds_ptr=ds.ptr,
                                             y_body_ptr=dy.ptr) == 0).all()
            want.append(out.download(np.zeros((n, per), np.uint32)))
            dev.append((sids, dr, ds))
        for rep in range(2):  # the second time with new lanes, the first ones closed
            lanes = [ctx.lane() for _ in range(L)]
            # two outputs per lane, taken in turn (each one's batch runs eagerly, is
            # captured, then replayed), read after the threads end
            outs = [[M.DeviceArray(n * per * 4) for _ in range(2)] for _ in lanes]
            errors = []
            go = threading.Barrier(L)

            def run(j):
                try:
                    sids, dr, ds = dev[j]
                    go.wait()
                    for r in range(rounds):
                        st = lanes[j].respond_batch_device(0, M.TWO_PACKING, sids, dr.ptr, outs[j][r % 2].ptr,
                                                           seeds_ptr=ds.ptr, y_body_ptr=dy.ptr)
                        assert (st == 0).all()
                except Exception as e:  # reported below
                    errors.append(e)

            th = [threading.Thread(target=run, args=(j,)) for j in range(L)]
            for t in th:
                t.start()
            for t in th:
                t.join()
            assert not errors, errors
            for j in range(L):
                for r in range(2):  # the last writes: a captured launch and replays
                    got = outs[j][r].download(np.zeros((n, per), np.uint32))
                    assert np.array_equal(got, want[j]), (rep, j, r)
                assert lanes[j].graph_stats()[1] >= 1
            assert lanes[1].chain_window() == window_us  # a lane reports its owner's
            with pytest.raises(M.PirError) as e:
                lanes[0].set_chain_window(0)
            assert e.value.code == M.ERR_INVALID_ARGUMENT
            for ln in lanes:
                ln.close()
        out = M.DeviceArray(n * per * 4)  # the owner alone after its lanes closed
        sids, dr, ds = dev[0]
        assert (ctx.respond_batch_device(0, M.TWO_PACKING, sids, dr.ptr, out.ptr, seeds_ptr=ds.ptr,
                                         y_body_ptr=dy.ptr) == 0).all()
        assert np.array_equal(out.download(np.zeros((n, per), np.uint32)), want[0])
