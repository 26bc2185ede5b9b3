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
        outs = [M.DeviceArray(n * per * 4) for _ in lanes]
        got = [[], []]
        errors = []

        def run(j):
            try:
                sids, dr, ds = dev[j]
                for _ in range(rounds):
                    st = lanes[j].respond_batch_device(0, M.TWO_PACKING, sids, dr.ptr, outs[j].ptr, seeds_ptr=ds.ptr,
                                                       y_body_ptr=dy.ptr)
                    assert (st == 0).all()
                    got[j].append(outs[j].download(np.zeros((n, per), np.uint32)))
            except Exception as e:  # reported below
                errors.append(e)

        th = [threading.Thread(target=run, args=(j,)) for j in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errors, errors
        for j in range(2):
            for r in got[j]:
                assert np.array_equal(r, want[j]), j
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
