"""The serving configuration at full size on the GPU: 2^28 x 32 B in 64
databases of log-uniform size (2^20..2^26 entries before scaling), resident
as NTT-domain digits (96 GiB), 16 clients with registered packing keys, and
InsPIRe^0/^1/^2 at 10/30/60 % arriving open-loop at Poisson rates of 10^4 and
10^5 queries/s through the continuous-batching loop (inspire_server_*).
Every response equals the oracle's respond / respond_inspiring of the same
query (bit-exact, tolerance 0), and a dozen queries answered alone equal the
same queries answered inside one large mixed submission.  The reference's
counterpart is its HTTP handlers calling respond_with_variant /
respond_seeded_with_variant per request (src/bin/server.rs:120-236,
src/pir/respond.rs:129-183)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W_SEED = bytes(range(32))


def say(*a):
    print("[gpu-config5]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    import inspire_mi355x as M

    return M


def test_serving_2e28_in_64_uneven_databases(M):
    import bench

    rng = np.random.default_rng(5)
    W = bench.serve_world(M, rng, 28, 64, W_SEED, say, pool_size=128)
    ctx, pool, sizes = W["ctx"], W["pool"], W["sizes"]
    try:
        assert len(sizes) == 64 and int(sizes.sum()) == 1 << 28 and (sizes % 2048 == 0).all()
        assert sizes.min() < 2**21 and sizes.max() > 2**24  # uneven: log-uniform before scaling
        variants = np.bincount([q["variant"] for q in pool], minlength=3)
        say("pool variants", variants.tolist(), "device GiB", round(ctx.device_bytes / 2**30, 1))
        assert (variants > 0).all()
        answers = bench.serve_cpu_answers(pool, W["raw"], W["offs"], W["ybodies"], W_SEED, say)
        structs, keep = bench.request_structs(M, pool)

        def check(j, st, r):
            q, want = pool[j], answers[j]
            assert st == M.OK, (j, st)
            assert np.array_equal(r.ciphertext, want[0]), j
            if q["variant"] == 0:
                assert np.array_equal(r.column_ciphertexts, want[1]), j
            else:
                assert len(r.column_ciphertexts) == 0

        with M.Server(ctx, lanes=2, max_batch=256) as srv:
            say("alone")
            alone = {}
            for j in range(12):
                arr = (M.RequestStruct * 1)(structs[j])
                t = srv.submit_structs(arr, 1)[0]
                st, r, _ = srv.wait(int(t), cols=16)
                check(j, st, r)
                alone[j] = r
            for lam, n in ((1e4, 2000), (1e5, 20000)):
                say("poisson", lam, n)
                res, bad = bench.poisson_run(M, srv, pool, structs, answers, lam, n, rng)
                say(res)
                assert bad == 0, res
            say("the whole pool in one submission")
            arr = (M.RequestStruct * len(pool))(*structs)
            tickets = srv.submit_structs(arr, len(pool))
            for j, t in enumerate(tickets):
                st, r, _ = srv.wait(int(t), cols=16)
                check(j, st, r)
                if j in alone:  # alone == in the mix
                    assert np.array_equal(r.ciphertext, alone[j].ciphertext)
                    assert np.array_equal(r.column_ciphertexts, alone[j].column_ciphertexts)
            stats = srv.stats()
            say("stats", stats)
            assert stats["requests"] == 12 + 2000 + 20000 + len(pool)
            assert stats["max_batch_seen"] > 1  # queries were batched as they came
    finally:
        ctx.close()
