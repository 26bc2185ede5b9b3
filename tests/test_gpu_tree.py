"""gfx950 tree packing (respond_one_packing / respond_seeded_packed,
src/pir/respond.rs:213-268, 428-435; pack_lwes, src/inspiring/automorph_pack.rs)
against the oracle: the golden cases, the reference's two-modulus b layout
and the corrected residue-wise one, every column count from one LWE (no tree)
to the full set, other ring degrees and a single modulus, the dispatch by
packing_mode, error paths, mixed batches and device-resident batches.
Bit-exact throughout."""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

import oracle as O
from helpers import Client, encode_db, phase_error, random_db

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "oracle_vectors.json")
W_SEED = bytes(range(32))


def say(*a):
    print("[gpu-tree]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    import inspire_mi355x as M

    return M


@pytest.fixture(scope="module")
def inp():
    from golden.make_golden import inputs

    return inputs()


@pytest.fixture(scope="module")
def want():
    return json.load(open(GOLDEN))["digests"]


def make_ctx(M, params: O.Params, flags=0):
    p = M.InspireParams(ring_dim=params.d, crt_moduli=params.moduli, p=params.p, gadget_base=params.gadget_base,
                        gadget_len=params.gadget_len, flags=flags)
    return M.Context(p, W_SEED)


def words(a, b):
    return np.concatenate([np.asarray(a).reshape(-1), np.asarray(b).reshape(-1)])


def test_tree_golden_full_and_seeded(M, inp, want):
    from golden.make_golden import digest

    say("tree golden")
    crt = inp["crt"]
    d = crt.params.d
    shards = encode_db(crt, inp["db"], 32)
    with make_ctx(M, crt.params) as ctx:
        ctx.load_database(0, shards)
        ctx.set_galois_keys(inp["gk"])
        r = M.respond_one_packing(ctx, 0, M.ClientQuery(shard_id=0, rgsw=inp["rows"]))
        assert r.b_dim == 2 * d and r.b.shape == (2, 2 * d) and r.a.shape == (2, d)
        assert len(r.column_ciphertexts) == 0
        assert digest(r.ciphertext) == want["tree_full"]
        a, b = crt.respond_one_packing(shards[0], inp["rows"], inp["gk"])
        assert np.array_equal(r.a, a) and np.array_equal(r.b, b)
        rs = M.respond_seeded_packed(ctx, 0, M.ClientQuery(shard_id=0, rgsw=inp["b"], seeds=inp["seeds"]))
        assert digest(rs.ciphertext) == want["tree_seeded"]


def test_tree_residue_wise_golden(M, inp, want):
    from golden.make_golden import digest

    say("tree residue-wise")
    crt = inp["crt"]
    small = encode_db(crt, inp["db_small"], 32)
    with make_ctx(M, crt.params, flags=M.FLAG_RESIDUE_WISE_AUTOMORPHISM) as ctx:
        ctx.load_database(1, small)
        ctx.set_galois_keys(inp["gk"])
        r = M.respond_one_packing(ctx, 1, M.ClientQuery(shard_id=0, rgsw=inp["rows"]))
        assert r.b_dim == crt.params.d and r.ciphertext.shape == (2, 2, crt.params.d)
        assert digest(r.ciphertext) == want["tree_residue_wise"]


def test_tree_dispatch_by_packing_mode(M, inp):
    """respond_with_variant / respond_seeded_with_variant with packing_mode =
    Tree take the tree path for both packed variants and need no client
    packing keys (respond.rs:135-149, 167-181)."""
    say("tree dispatch")
    crt = inp["crt"]
    shards = encode_db(crt, inp["db"], 32)
    with make_ctx(M, crt.params) as ctx:
        ctx.load_database(0, shards)
        ctx.set_galois_keys(inp["gk"])
        ref = M.respond_one_packing(ctx, 0, M.ClientQuery(shard_id=1, rgsw=inp["rows"]))
        for v in (M.ONE_PACKING, M.TWO_PACKING):
            q = M.ClientQuery(shard_id=1, rgsw=inp["rows"], packing_mode=M.PACKING_TREE)
            r = M.respond_with_variant(ctx, 0, q, v)
            assert np.array_equal(r.ciphertext, ref.ciphertext), v
        qs = M.ClientQuery(shard_id=1, rgsw=inp["b"], seeds=inp["seeds"], packing_mode=M.PACKING_TREE)
        r = M.respond_seeded_with_variant(ctx, 0, qs, M.TWO_PACKING)
        rows_x = crt.seeded_rgsw_expand(inp["seeds"], inp["b"])
        assert np.array_equal(r.ciphertext, words(*crt.respond_one_packing(shards[1], rows_x, inp["gk"])))
        # NoPacking ignores packing_mode (respond.rs:136)
        r0 = M.respond_with_variant(ctx, 0, M.ClientQuery(shard_id=1, rgsw=inp["rows"], packing_mode=M.PACKING_TREE),
                                    M.NO_PACKING)
        ct, _ = crt.respond(shards[1], inp["rows"])
        assert np.array_equal(r0.ciphertext, ct)


@pytest.mark.parametrize("entry_size", [1, 2, 3, 6, 10, 32, 34, 64])
def test_tree_every_column_count(M, inp, entry_size):
    """m = ceil(8 e / 16) columns: one LWE (no tree, b_dim = d), counts that
    are not powers of two (nodes with and without an odd child in one level)
    and more than 16."""
    say("tree columns", entry_size)
    crt = inp["crt"]
    d = crt.params.d
    db = random_db(d, entry_size, seed=100 + entry_size)
    shards = encode_db(crt, db, entry_size)
    rows = inp["cl"].query_rows(7 * entry_size)
    a, b = crt.respond_one_packing(shards[0], rows, inp["gk"])
    with make_ctx(M, crt.params) as ctx:
        ctx.load_database(0, shards)
        ctx.set_galois_keys(inp["gk"])
        r = M.respond_one_packing(ctx, 0, M.ClientQuery(shard_id=0, rgsw=rows))
        assert r.b_dim == b.shape[1], (entry_size, r.b_dim, b.shape)
        assert np.array_equal(r.a, a) and np.array_equal(r.b, b), entry_size


# the last two: moduli above 2^28 take the unsigned (Harvey) arithmetic, lazy
# (3.0e8) and reduced (4.5e8) multiply-accumulates
@pytest.mark.parametrize("d,moduli,entry_size", [(256, (O.Q0, O.Q1), 32), (1024, (O.Q0, O.Q1), 12),
                                                 (256, (O.Q0,), 32), (2048, (O.Q1,), 8),
                                                 (2048, (300122113, 300244993), 32),
                                                 (2048, (450011137, 450109441), 32)])
def test_tree_other_degrees_and_one_modulus(M, d, moduli, entry_size):
    say("tree shape", d, moduli, entry_size)
    params = O.Params(d=d, moduli=moduli)
    o = O.Oracle(params)
    cl = Client(o, d + entry_size)
    gk = o.galois_keys(cl.sk, o.rng(bytes(range(5, 37))))
    db = random_db(d + 3, entry_size, seed=d)
    shards = encode_db(o, db, entry_size)
    rows = cl.query_rows(d // 5)
    with make_ctx(M, params) as ctx:
        ctx.load_database(0, shards)
        ctx.set_galois_keys(gk)
        for sid in (0, 1):
            r = M.respond_one_packing(ctx, 0, M.ClientQuery(shard_id=sid, rgsw=rows))
            a, b = o.respond_one_packing(shards[sid], rows, gk)
            assert np.array_equal(r.a, a) and np.array_equal(r.b, b), (d, moduli, sid)


def test_tree_residue_wise_phase(M, inp):
    """In the corrected mode the GPU's packed response encrypts column c of
    the entry times d (mod p) at coefficient c: its phase stays within 4 Delta
    of that message everywhere (the tree's key-switching noise at d = 2048 is
    about Delta; a wrong index, sign or key leaves errors near q/2)."""
    say("tree phase")
    crt = inp["crt"]
    d, p = crt.params.d, crt.params.p
    db = random_db(d, 32, seed=77, values_below_256=True)
    shards = encode_db(crt, db, 32)
    cl = inp["cl"]
    with make_ctx(M, crt.params, flags=M.FLAG_RESIDUE_WISE_AUTOMORPHISM) as ctx:
        ctx.load_database(0, shards)
        ctx.set_galois_keys(inp["gk"])
        for idx in (0, 1234):
            r = M.respond_one_packing(ctx, 0, M.ClientQuery(shard_id=0, rgsw=cl.query_rows(idx)))
            entry = np.frombuffer(db, dtype=np.uint8).reshape(d, 32)[idx]
            cols = [int(entry[2 * c]) | (int(entry[2 * c + 1]) << 8) for c in range(16)]
            assert phase_error(crt, cl.sk, r.a, r.b, [v * d % p for v in cols] + [0] * (d - 16)) < 4.0


def test_tree_errors_and_edges(M, inp):
    say("tree errors")
    crt = inp["crt"]
    shards = encode_db(crt, inp["db"], 32)
    rows = inp["rows"]
    K, d = crt.params.K, crt.params.d
    with make_ctx(M, crt.params) as ctx:
        ctx.load_database(0, shards)
        ctx.load_database(2, np.zeros((1, 0, K, d), dtype=np.uint64))  # one empty shard
        with pytest.raises(M.PirError) as e:  # no galois keys in the CRS
            M.respond_one_packing(ctx, 0, M.ClientQuery(shard_id=0, rgsw=rows))
        assert e.value.code == M.ERR_INVALID_ARGUMENT and "galois" in str(e.value)
        with pytest.raises(M.PirError) as e:  # log2(d) matrices
            ctx.set_galois_keys(inp["gk"][:5])
        assert e.value.code == M.ERR_INVALID_ARGUMENT
        ctx.set_galois_keys(inp["gk"])
        with pytest.raises(M.PirError) as e:  # respond.rs:224-228
            M.respond_one_packing(ctx, 0, M.ClientQuery(shard_id=999, rgsw=rows))
        assert str(e.value) == "Shard 999 not found"
        # empty shard: zero ciphertext, column_ciphertexts = [zero] (respond.rs:230-236)
        r = M.respond_one_packing(ctx, 2, M.ClientQuery(shard_id=0, rgsw=rows))
        assert r.b_dim == d and not r.ciphertext.any() and len(r.column_ciphertexts) == 1
        # a caller buffer of 2 K d words cannot hold b of 2d coefficients
        keep: list = []
        cq = ctx._query(0, M.ClientQuery(shard_id=0, rgsw=rows), keep)
        ct = np.zeros(2 * K * d, dtype=np.uint64)
        rr = M._Response()
        rr.ciphertext = M._ptr64(ct)
        rc = M.lib().inspire_respond_one_packing(ctx._h, M.C.byref(cq), M.C.byref(rr))
        assert rc == M.ERR_CAPACITY


def test_tree_in_mixed_batch(M, inp):
    """Tree-packed queries batched with InsPIRe^0 and InspiRING ones, over two
    databases: each equals its response alone and the oracle's."""
    say("tree mixed batch")
    crt = inp["crt"]
    d = crt.params.d
    sa = encode_db(crt, random_db(2 * d, 32, seed=61), 32)
    sb = encode_db(crt, random_db(d, 32, seed=62), 32)
    cl = inp["cl"]
    yb = inp["yb"]
    g = np.random.default_rng(63)
    qs, dbs, vs = [], [], []
    for i in range(9):
        db = i % 2
        sid = int(g.integers(0, 2)) if db == 0 else 0
        local = int(g.integers(0, d))
        mode = M.PACKING_TREE if i % 3 != 2 else M.PACKING_INSPIRING
        v = [M.ONE_PACKING, M.TWO_PACKING, M.NO_PACKING, M.TWO_PACKING][i % 4]
        if i % 2:
            s, b = cl.seeded_query(local)
            q = M.ClientQuery(shard_id=sid, rgsw=b, seeds=s, packing_mode=mode, has_packing_keys=True, y_body=yb)
        else:
            q = M.ClientQuery(shard_id=sid, rgsw=cl.query_rows(local), packing_mode=mode, has_packing_keys=True,
                              y_body=yb)
        qs.append(q)
        dbs.append(db)
        vs.append(v)
    with make_ctx(M, crt.params) as ctx:
        ctx.load_database(0, sa)
        ctx.load_database(1, sb)
        ctx.set_galois_keys(inp["gk"])
        res = ctx.respond_batch(dbs, qs, vs)
        for i, (q, db, v) in enumerate(zip(qs, dbs, vs)):
            alone = M.respond_with_variant(ctx, db, q, v)
            assert np.array_equal(res[i].ciphertext, alone.ciphertext), i
            shards = sa if db == 0 else sb
            rows = crt.seeded_rgsw_expand(q.seeds, q.rgsw) if q.seeded else q.rgsw
            if v == M.NO_PACKING:
                want_ct, _ = crt.respond(shards[q.shard_id], rows)
            elif q.packing_mode == M.PACKING_TREE:
                want_ct = words(*crt.respond_one_packing(shards[q.shard_id], rows, inp["gk"]))
            else:
                _, want_ct = crt.respond_inspiring(W_SEED, shards[q.shard_id], rows, y_body=yb)
            assert np.array_equal(res[i].ciphertext.reshape(-1), np.asarray(want_ct).reshape(-1)), i


def test_tree_device_batch(M, inp):
    """inspire_respond_batch_device with packing_mode = Tree (inputs in HBM,
    seeded rows, no client keys) == the host API, per query."""
    say("tree device batch")
    crt = inp["crt"]
    d, K = crt.params.d, crt.params.K
    shards = encode_db(crt, random_db(3 * d, 32, seed=71), 32)
    cl = inp["cl"]
    g = np.random.default_rng(72)
    n = 5
    sids = g.integers(0, 3, n).astype(np.uint32)
    sids[2] = 50  # unknown shard: reported per query
    rows, seeds, qs = [], [], []
    for i in range(n):
        s, b = cl.seeded_query(int(g.integers(0, d)))
        rows.append(b[3:6].astype(np.uint32))
        seeds.append(np.frombuffer(s[3:6].tobytes(), dtype="<u4").reshape(3, 8))
        qs.append(M.ClientQuery(shard_id=int(sids[i]), rgsw=b, seeds=s))
    per = K * d + K * 2 * d
    with make_ctx(M, crt.params) as ctx:
        ctx.load_database(0, shards)
        ctx.set_galois_keys(inp["gk"])
        drows = M.DeviceArray.from_numpy(np.stack(rows))
        dseeds = M.DeviceArray.from_numpy(np.stack(seeds))
        dout = M.DeviceArray(n * per * 4)
        st = ctx.respond_batch_device(0, M.TWO_PACKING, sids, drows.ptr, dout.ptr, seeds_ptr=dseeds.ptr,
                                      packing_mode=M.PACKING_TREE)
        out = dout.download(np.zeros((n, per), dtype=np.uint32))
        for i in range(n):
            if sids[i] == 50:
                assert st[i] == M.ERR_SHARD_NOT_FOUND
                continue
            assert st[i] == M.OK
            ref = M.respond_seeded_packed(ctx, 0, qs[i])
            assert np.array_equal(out[i].astype(np.uint64), ref.ciphertext.reshape(-1)), i
