"""The serving loop (inspire_server_*) on the GPU: every response of a mixed
InsPIRe^0/^1/^2 stream over several databases equals the oracle's respond /
respond_inspiring for the same query, and equals the response the same query
gets alone; batches overlap on the device (two lanes); admission by the HBM
budget; the reference's error cases.  Bit-exact (integer outputs, tolerance 0)."""
from __future__ import annotations

import sys
import threading
import time

import numpy as np
import pytest

import oracle as O
from helpers import Client, encode_db, random_db

pytestmark = pytest.mark.gpu

W_SEED = bytes(range(32))


def say(*a):
    print("[gpu-serve]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    import inspire_mi355x as M

    return M


@pytest.fixture(scope="module")
def world(M):
    """Three databases, four clients with registered keys, and 36 queries
    (with the oracle's answers) mixing variants, databases and clients."""
    o = O.Oracle(O.baseline_params())
    d, ell = o.params.d, o.params.gadget_len
    dbs = {0: random_db(3 * d, 32, seed=81), 1: random_db(2 * d - 7, 32, seed=82), 2: random_db(d, 32, seed=83)}
    shards = {k: encode_db(o, v, 32) for k, v in dbs.items()}
    clients = [Client(o, 90 + i) for i in range(4)]
    keys = [cl.packing_keys(16, W_SEED, with_y_all=False)[0] for cl in clients]
    g = np.random.default_rng(84)
    queries = []
    for i in range(36):
        db = int(g.integers(0, 3))
        shard = int(g.integers(0, shards[db].shape[0]))
        local = int(g.integers(0, d))
        variant = [0, 1, 2, 2, 2][i % 5]  # InsPIRe^0 / ^1 / ^2 at 20 / 20 / 60 %
        c = i % 4
        cl = clients[c]
        if variant == 1:  # full query rows
            rows = cl.query_rows(local)
            seeds, b = None, None
            full = rows
        else:  # seeded
            seeds, b = cl.seeded_query(local)
            full = o.seeded_rgsw_expand(seeds, b)
        if variant == 0:
            want, want_cols = o.respond(shards[db][shard], full)
        else:
            rc, want = o.respond_inspiring(W_SEED, shards[db][shard], full, y_body=keys[c])
            assert rc == 0
            want_cols = None
        queries.append(dict(db=db, shard=shard, variant=variant, client=c, seeds=seeds, b=b, full=full,
                            want=want, want_cols=want_cols))
    say("oracle answers ready")
    return dict(o=o, dbs=dbs, shards=shards, keys=keys, queries=queries)


def make_ctx(M, world):
    o = world["o"]
    p = o.params
    ctx = M.Context(M.InspireParams(ring_dim=p.d, crt_moduli=p.moduli, p=p.p, gadget_base=p.gadget_base,
                                    gadget_len=p.gadget_len), W_SEED)
    for k, v in world["dbs"].items():
        ctx.load_entries(k, v, 32)
    for c, yb in enumerate(world["keys"]):
        ctx.register_client_keys(c, yb)
    return ctx


def request(M, q, ell):
    if q["variant"] == 1:
        return M.Request(q["db"], q["shard"], 1, rows=q["full"][ell:].astype(np.uint32), client_id=q["client"])
    return M.Request(q["db"], q["shard"], q["variant"], rows=q["b"][ell:].astype(np.uint32),
                     seeds=q["seeds"][ell:], client_id=q["client"])


def check(M, q, res):
    st, r, _ = res
    assert st == M.OK
    assert (r.ciphertext == q["want"]).all()
    if q["variant"] == 0:
        assert (r.column_ciphertexts == q["want_cols"]).all()
    else:
        assert len(r.column_ciphertexts) == 0


def test_serving_mix_equals_oracle_and_alone(M, world):
    ell = world["o"].params.gadget_len
    qs = world["queries"]
    with make_ctx(M, world) as ctx, M.Server(ctx, lanes=2, max_batch=8) as srv:
        say("alone")
        alone = []
        for q in qs[:12]:
            t = srv.submit([request(M, q, ell)])[0]
            alone.append(srv.wait(int(t), cols=16))
            check(M, q, alone[-1])
        say("mixed: all at once")
        tickets = srv.submit([request(M, q, ell) for q in qs])
        mixed = [srv.wait(int(t), cols=16) for t in tickets]
        for q, res in zip(qs, mixed):
            check(M, q, res)
        for a, m in zip(alone, mixed[:12]):  # alone == in the mix
            assert (a[1].ciphertext == m[1].ciphertext).all()
            assert (a[1].column_ciphertexts == m[1].column_ciphertexts).all()
        say("mixed: arrivals while batches run")
        results = {}

        def submitter():
            for i in range(0, len(qs), 5):
                for j, t in enumerate(srv.submit([request(M, q, ell) for q in qs[i:i + 5]])):
                    results[i + j] = int(t)
                time.sleep(0.002)

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        for i, q in enumerate(qs):
            check(M, q, srv.wait(results[i], cols=16))
        st = srv.stats()
        say("stats", st)
        assert st["requests"] == 12 + 2 * len(qs)
        assert st["batches"] >= len(qs) // 8
        assert st["max_batch_seen"] <= 8
        # one column count (16): every batch, whatever its variants, is one launch sequence
        assert st["launch_groups"] == st["batches"]


def test_serving_variants_share_one_launch(M, world):
    """InsPIRe^0, ^1 and ^2 queries of one column count in one batch run as
    one launch sequence (expansion, fused front end with per-query roles,
    chain kernel whose InsPIRe^0 workgroups only transform), and every
    response equals the oracle's."""
    ell = world["o"].params.gadget_len
    qs = world["queries"]
    assert {q["variant"] for q in qs} == {0, 1, 2}
    with make_ctx(M, world) as ctx, M.Server(ctx, lanes=1, max_batch=64) as srv:
        tickets = srv.submit([request(M, q, ell) for q in qs])
        for q, t in zip(qs, tickets):
            check(M, q, srv.wait(int(t), cols=16))
        st = srv.stats()
        say("one-launch stats", st)
        assert st["requests"] == len(qs)
        assert st["launch_groups"] == st["batches"]
        assert st["max_batch_seen"] > 1


def test_serving_budget_limits_batches(M, world):
    """An HBM budget of two packed queries keeps batches at two or fewer."""
    ell = world["o"].params.gadget_len
    qs = [q for q in world["queries"] if q["variant"] != 0][:10]
    o = world["o"]
    K, d = o.params.K, o.params.d
    KN = K * d
    per = (16 + ell * 32 + ell * 2 * KN * 4) + ell * 2 * KN * 4 + 16 * KN * 4 + (d // 64) * 16 * K * 8 + 2 * KN * 4
    with make_ctx(M, world) as ctx, M.Server(ctx, lanes=2, max_batch=64, hbm_budget=2 * per) as srv:
        tickets = srv.submit([request(M, q, ell) for q in qs])
        for q, t in zip(qs, tickets):
            check(M, q, srv.wait(int(t)))
        st = srv.stats()
        say("budget stats", st)
        assert st["max_batch_seen"] <= 2
        assert st["peak_inflight_bytes"] <= 2 * per


def test_serving_errors_match_reference(M, world):
    ell = world["o"].params.gadget_len
    q = [q for q in world["queries"] if q["variant"] == 2][0]
    with make_ctx(M, world) as ctx, M.Server(ctx, lanes=1, max_batch=4) as srv:
        bad_shard = M.Request(q["db"], 999, 2, rows=q["b"][ell:].astype(np.uint32), seeds=q["seeds"][ell:],
                              client_id=q["client"])
        bad_client = M.Request(q["db"], q["shard"], 2, rows=q["b"][ell:].astype(np.uint32), seeds=q["seeds"][ell:],
                               client_id=77)
        bad_db = M.Request(9, q["shard"], 2, rows=q["b"][ell:].astype(np.uint32), seeds=q["seeds"][ell:],
                           client_id=q["client"])
        good = request(M, q, ell)
        t = srv.submit([bad_shard, bad_client, bad_db, good])
        assert srv.wait(int(t[0]))[0] == M.ERR_SHARD_NOT_FOUND
        assert srv.wait(int(t[1]))[0] == M.ERR_PACKING_KEYS_MISSING
        assert srv.wait(int(t[2]))[0] == M.ERR_UNKNOWN_DATABASE
        check(M, q, srv.wait(int(t[3])))
        with pytest.raises(M.PirError):
            srv.wait(int(t[3]))  # released


def test_serving_other_column_counts(M):
    """Databases whose column count is off the fused front end's shape:
    entry sizes 34 (gamma 17) and 256 (gamma 128) serve packed queries equal
    to the oracle's; entry size 6 (gamma 3) has an even generator, whose
    rotations of the registered y_body are not NTT-slot permutations, so its
    packed queries are refused (INSPIRE_ERR_UNSUPPORTED) while its InsPIRe^0
    queries are served."""
    o = O.Oracle(O.baseline_params())
    d, ell = o.params.d, o.params.gadget_len
    sizes = {0: 34, 1: 256, 2: 6}
    dbs = {k: random_db(d + 3, es, seed=100 + k) for k, es in sizes.items()}
    shards = {k: encode_db(o, v, sizes[k]) for k, v in dbs.items()}
    cl = Client(o, 120)
    yb = cl.packing_keys(16, W_SEED, with_y_all=False)[0]
    p = o.params
    reqs, wants = [], []
    for k in sizes:
        for variant, sh, local in ((2, 1, 5), (0, 0, 9)):
            seeds, b = cl.seeded_query(local)
            full = o.seeded_rgsw_expand(seeds, b)
            reqs.append(M.Request(k, sh, variant, rows=b[ell:].astype(np.uint32), seeds=seeds[ell:], client_id=3))
            if variant == 0:
                wants.append(o.respond(shards[k][sh], full))
            elif sizes[k] == 6:
                wants.append(None)
            else:
                rc, ct = o.respond_inspiring(W_SEED, shards[k][sh], full, y_body=yb)
                assert rc == 0
                wants.append((ct, None))
    say("other column counts: oracle answers ready")
    ctx = M.Context(M.InspireParams(ring_dim=p.d, crt_moduli=p.moduli, p=p.p, gadget_base=p.gadget_base,
                                    gadget_len=p.gadget_len), W_SEED)
    with ctx:
        for k, v in dbs.items():
            ctx.load_entries(k, v, sizes[k])
        ctx.register_client_keys(3, yb)
        with M.Server(ctx, lanes=2, max_batch=4) as srv:
            tickets = srv.submit(reqs)
            for r, t, want in zip(reqs, tickets, wants):
                st, res, _ = srv.wait(int(t), cols=128)
                if want is None:
                    assert st == M.ERR_UNSUPPORTED
                    continue
                assert st == M.OK
                assert (res.ciphertext == want[0]).all()
                if want[1] is not None:
                    assert (res.column_ciphertexts == want[1]).all()
            # none of these column counts has the fused front end: each keeps a
            # launch sequence per variant class
            assert srv.stats()["launch_groups"] >= srv.stats()["batches"]
