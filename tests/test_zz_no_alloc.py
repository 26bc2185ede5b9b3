"""inspire_respond_batch_device_noalloc (library-held scratch, no allocation
in the call) against inspire_respond_batch_device on the same inputs, bit for
bit: InsPIRe^2 (seeded rows, y_body keys, fused front end), InsPIRe^2 over a
17-column database (general front end, Horner step 1) and InsPIRe^0 (full
rows), each on two calls in a row, the second with fewer queries than the
first, so the second reads scratch the first left behind.  Then the refusals:
a batch beyond the scratch, tree packing, and a column count without tables."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle as O
from helpers import random_db

pytestmark = pytest.mark.gpu

W_SEED = bytes(range(32))
Q0, Q1 = 268369921, 249561089
D, K, ELL = 2048, 2, 3


def say(*a):
    print("[gpu-noalloc]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    import inspire_mi355x as M

    return M


def residues(g, *shape):
    m = np.array([Q0, Q1], dtype=np.uint64)[:, None]
    return (g.integers(0, 2**63, shape + (K, D), dtype=np.uint64) % m).astype(np.uint32)


def batch_inputs(M, g, n, nshards, full_rows=False):
    shards = g.integers(0, nshards, n, dtype=np.uint64).astype(np.uint32)
    shards[-1] = nshards + 5  # a missing shard: SHARD_NOT_FOUND, the others unaffected
    rows = residues(g, n, ELL, 2) if full_rows else residues(g, n, ELL)
    seeds = g.integers(0, 2**32, (n, ELL, 8), dtype=np.uint64).astype(np.uint32)
    y = residues(g, n, ELL)
    return shards, [M.DeviceArray.from_numpy(a) for a in (rows, seeds, y)]


def run_both(M, ctx, variant, shards, dev, out_words, seeded=True):
    drows, dseeds, dy = dev
    n = len(shards)
    res = []
    for noalloc in (False, True):
        out = M.DeviceArray(n * out_words * 4)
        st = ctx.respond_batch_device(0, variant, shards, drows.ptr, out.ptr, seeds_ptr=dseeds.ptr if seeded else 0,
                                      y_body_ptr=dy.ptr, seeded=seeded, noalloc=noalloc)
        got = out.download(np.zeros((n, out_words), dtype=np.uint32))
        out.free()
        res.append((st, got))
    (st0, a), (st1, b) = res
    assert (st0 == st1).all()
    assert st0[-1] == M.ERR_SHARD_NOT_FOUND and (st0[:-1] == 0).all()
    ok = st0 == 0
    assert (a[ok] == b[ok]).all()
    return a


@pytest.mark.parametrize("entry_size,variant,counts", [(32, 2, (48, 20)), (34, 2, (24, 10)), (32, 0, (48, 20))])
def test_noalloc_equals_batch_device_twice(M, entry_size, variant, counts):
    g = np.random.default_rng(7000 + entry_size + variant)
    db = random_db(16 * D, entry_size, seed=entry_size)
    with M.Context(M.InspireParams(), W_SEED) as ctx:
        ctx.load_entries(0, db, entry_size)
        cols = (entry_size * 8 + 15) // 16
        out_words = 2 * K * D if variant else (1 + cols) * 2 * K * D
        for n in counts:  # the second call smaller than the first (17 columns: 0.9 MiB of scratch per query)
            say("entry", entry_size, "variant", variant, "queries", n)
            shards, dev = batch_inputs(M, g, n, 16, full_rows=variant == 0)
            got = run_both(M, ctx, variant, shards, dev, out_words, seeded=variant != 0)
            if variant == 2 and n == counts[1]:  # one response against the oracle as well
                o = O.Oracle(O.baseline_params())
                drows, dseeds, dy = dev
                rows = drows.download(np.zeros((n, ELL, K, D), dtype=np.uint32))
                seeds = dseeds.download(np.zeros((n, ELL, 8), dtype=np.uint32))
                y = dy.download(np.zeros((n, ELL, K, D), dtype=np.uint32))
                s = int(shards[0])
                polys = o.encode_shard(np.frombuffer(db, dtype=np.uint8), entry_size, s)
                rg = np.zeros((2 * ELL, 2, K, D), dtype=np.uint64)
                for r in range(ELL):
                    rg[ELL + r, 0] = o.poly_from_seed(seeds[0, r].astype("<u4").tobytes())
                    rg[ELL + r, 1] = rows[0, r]
                rc, want = o.respond_inspiring(W_SEED, polys, rg, y_body=y[0].astype(np.uint64))
                assert rc == 0
                assert (got[0].reshape(2, K, D) == want).all()
            for a in dev:
                a.free()


def test_noalloc_refusals(M):
    g = np.random.default_rng(7100)
    db = random_db(4 * D, 32, seed=9)
    with M.Context(M.InspireParams(), W_SEED) as ctx:
        ctx.load_entries(0, db, 32)
        shards, dev = batch_inputs(M, g, 4, 4)
        drows, dseeds, dy = dev
        out = M.DeviceArray(4 * 3 * K * D * 4)
        say("refusals: no tables yet")
        with pytest.raises(M.PirError) as e:  # no InspiRING tables for 16 columns on this context yet
            ctx.respond_batch_device(0, 2, shards, drows.ptr, out.ptr, seeds_ptr=dseeds.ptr, y_body_ptr=dy.ptr,
                                     noalloc=True)
        assert e.value.code == M.ERR_UNSUPPORTED
        say("refusals: tree packing")
        with pytest.raises(M.PirError) as e:
            ctx.set_galois_keys(residues(g, 11, ELL, 2).astype(np.uint64))
            ctx.respond_batch_device(0, 2, shards, drows.ptr, out.ptr, seeds_ptr=dseeds.ptr,
                                     packing_mode=M.PACKING_TREE, noalloc=True)
        assert e.value.code == M.ERR_UNSUPPORTED
        # build the tables with an allocating call, then a batch beyond the scratch
        st = ctx.respond_batch_device(0, 2, shards, drows.ptr, out.ptr, seeds_ptr=dseeds.ptr, y_body_ptr=dy.ptr)
        assert (st[:-1] == 0).all()
        say("refusals: beyond the scratch")
        big = np.zeros(100, dtype=np.uint32)
        with pytest.raises(M.PirError) as e:  # 100 packed queries need about 35 MiB
            ctx.respond_batch_device(0, 2, big, drows.ptr, out.ptr, seeds_ptr=dseeds.ptr, y_body_ptr=dy.ptr,
                                     noalloc=True)
        assert e.value.code == M.ERR_CAPACITY and "bytes of scratch" in str(e.value)
        out.free()
        for a in dev:
            a.free()
