"""gfx950 respond path vs the oracle: InsPIRe^0 (full and seeded), InspiRING
(InsPIRe^1/^2) with y_body, y_all and seeded queries, the corrected
residue-wise mode, the reference's error paths, and batches that mix
variants and databases.  Bit-exact throughout."""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

import oracle as O
from helpers import Client, encode_db, random_db

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "oracle_vectors.json")
W_SEED = bytes(range(32))


def say(*a):
    print("[gpu-respond]", *a, flush=True)
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


def test_respond0_full_and_seeded_golden(M, inp, want):
    from golden.make_golden import digest

    say("respond0 golden")
    crt = inp["crt"]
    shards = encode_db(crt, inp["db"], 32)
    with make_ctx(M, crt.params) as ctx:
        ctx.load_database(0, shards)
        r = M.respond(ctx, 0, M.ClientQuery(shard_id=0, rgsw=inp["rows"]))
        assert digest(r.ciphertext) == want["respond0_full_ct"]
        assert digest(r.column_ciphertexts) == want["respond0_full_cols"]
        r = M.respond_seeded(ctx, 0, M.ClientQuery(shard_id=0, rgsw=inp["b"], seeds=inp["seeds"]))
        assert digest(r.ciphertext) == want["respond0_seeded_ct"]


def test_respond0_extracts_entries_via_entries_loader(M):
    say("respond0 e2e extract")
    o = O.Oracle(O.baseline_params())
    d = o.params.d
    db = random_db(2 * d + 50, 32, seed=31)
    cl = Client(o, 3)
    with make_ctx(M, o.params) as ctx:
        ctx.load_entries(7, db, 32)
        for idx in [0, 5, d - 1, d, 2 * d + 49]:
            s, k = divmod(idx, d)
            rows = cl.query_rows(k)
            r = M.respond(ctx, 7, M.ClientQuery(shard_id=s, rgsw=rows))
            ct, cols = o.respond(encode_db(o, db, 32)[s], rows)
            assert (r.ciphertext == ct).all() and (r.column_ciphertexts == cols).all()
            vals = [int(o.decrypt(r.column_ciphertexts[c], cl.sk)[0]) for c in range(16)]
            assert O.reconstruct_entry(vals, 32) == db[idx * 32:(idx + 1) * 32]


@pytest.mark.parametrize("case", ["inspiring_y_body", "inspiring_y_all", "inspiring_seeded_y_body"])
def test_inspiring_golden(M, inp, want, case):
    from golden.make_golden import digest

    say("inspiring golden", case)
    crt = inp["crt"]
    small = encode_db(crt, inp["db_small"], 32)
    with make_ctx(M, crt.params) as ctx:
        ctx.load_database(1, small)
        if case == "inspiring_seeded_y_body":
            q = M.ClientQuery(shard_id=0, rgsw=inp["b"], seeds=inp["seeds"], has_packing_keys=True, y_body=inp["yb"])
            r = M.respond_seeded_inspiring(ctx, 1, q)
        else:
            q = M.ClientQuery(shard_id=0, rgsw=inp["rows"], has_packing_keys=True, y_body=inp["yb"],
                              y_all=inp["ya"] if case == "inspiring_y_all" else None)
            r = M.respond_inspiring(ctx, 1, q)
        assert len(r.column_ciphertexts) == 0
        assert digest(r.ciphertext) == want[case]


def test_inspiring_residue_wise_golden_and_decrypts(M, inp, want):
    from golden.make_golden import digest

    say("inspiring residue-wise")
    crt = inp["crt"]
    small = encode_db(crt, inp["db_small"], 32)
    with make_ctx(M, crt.params, flags=M.FLAG_RESIDUE_WISE_AUTOMORPHISM) as ctx:
        ctx.load_database(1, small)
        q = M.ClientQuery(shard_id=0, rgsw=inp["rows"], has_packing_keys=True, y_body=inp["yb"])
        r = M.respond_inspiring(ctx, 1, q)
        assert digest(r.ciphertext) == want["inspiring_residue_wise"]
        m = crt.decrypt(r.ciphertext, inp["cl"].sk)
        assert O.reconstruct_entry([int(m[c]) for c in range(16)], 32) == inp["db_small"][100 * 32:101 * 32]


@pytest.mark.parametrize("d,entry_size", [(256, 32), (1024, 8), (2048, 2), (2048, 64), (4096, 32)])
def test_inspiring_matches_oracle_other_shapes(M, d, entry_size):
    say("inspiring shape", d, entry_size)
    params = O.Params(d=d)
    o = O.Oracle(params)
    db = random_db(d, entry_size, seed=d + entry_size)
    shards = encode_db(o, db, entry_size)
    gamma = shards.shape[1]
    cl = Client(o, d)
    yb, ya = cl.packing_keys(gamma, W_SEED)
    rows = cl.query_rows(d // 3)
    with make_ctx(M, params) as ctx:
        ctx.load_database(0, shards)
        for use_all in (False, True):
            q = M.ClientQuery(shard_id=0, rgsw=rows, has_packing_keys=True, y_body=yb,
                              y_all=ya if use_all else None)
            r = M.respond_inspiring(ctx, 0, q)
            _, ct = o.respond_inspiring(W_SEED, shards[0], rows, y_body=yb, y_all=ya if use_all else None)
            assert (r.ciphertext == ct).all(), (d, entry_size, use_all)


# Moduli above 2^28 take the packing chain's unsigned (Harvey) arithmetic: its
# signed variant needs the digit transforms to stay inside int32, which holds
# for q below about 2.8e8.  4 ell q < 2^32 then selects the lazy
# multiply-accumulate (3.0e8) or the reduced one (4.5e8).
WIDE_MODULI = [(300122113, 300244993), (450011137, 450109441)]


@pytest.mark.parametrize("moduli", WIDE_MODULI)
def test_inspiring_matches_oracle_wide_moduli(M, moduli):
    say("inspiring moduli", moduli)
    params = O.Params(d=2048, moduli=moduli)
    o = O.Oracle(params)
    db = random_db(2048, 32, seed=moduli[0] % 1000)
    shards = encode_db(o, db, 32)
    gamma = shards.shape[1]
    cl = Client(o, 77)
    yb, ya = cl.packing_keys(gamma, W_SEED)
    rows = cl.query_rows(1234)
    seeds, sb = cl.seeded_query(99)
    with make_ctx(M, params) as ctx:
        ctx.load_database(0, shards)
        for use_all in (False, True):
            q = M.ClientQuery(shard_id=0, rgsw=rows, has_packing_keys=True, y_body=yb,
                              y_all=ya if use_all else None)
            r = M.respond_inspiring(ctx, 0, q)
            _, ct = o.respond_inspiring(W_SEED, shards[0], rows, y_body=yb, y_all=ya if use_all else None)
            assert (r.ciphertext == ct).all(), (moduli, use_all)
        q = M.ClientQuery(shard_id=0, rgsw=sb, seeds=seeds, has_packing_keys=True, y_body=yb)
        r = M.respond_seeded_inspiring(ctx, 0, q)
        _, ct = o.respond_inspiring(W_SEED, shards[0], o.seeded_rgsw_expand(seeds, sb), y_body=yb)
        assert (r.ciphertext == ct).all(), (moduli, "seeded")


def test_error_paths_match_reference(M, inp):
    say("error paths")
    crt = inp["crt"]
    shards = encode_db(crt, inp["db"], 32)
    rows = inp["rows"]
    with make_ctx(M, crt.params) as ctx:
        ctx.load_database(0, shards)
        ctx.load_database(2, np.zeros((1, 0) + crt.P, dtype=np.uint64))  # one empty shard
        with pytest.raises(M.PirError) as e:  # respond.rs:78-82
            M.respond(ctx, 0, M.ClientQuery(shard_id=999, rgsw=rows))
        assert str(e.value) == "Shard 999 not found" and e.value.code == M.ERR_SHARD_NOT_FOUND
        with pytest.raises(M.PirError) as e:  # respond.rs:296-299
            M.respond_inspiring(ctx, 0, M.ClientQuery(shard_id=0, rgsw=rows))
        assert str(e.value) == "InspiRING client packing keys missing from query"
        with pytest.raises(M.PirError) as e:  # respond.rs:141-143
            M.respond_with_variant(ctx, 0, M.ClientQuery(shard_id=0, rgsw=rows), M.ONE_PACKING)
        assert str(e.value) == "InspiRING packing keys missing (set packing_mode=tree to use tree packing)"
        with pytest.raises(M.PirError) as e:  # respond.rs:366-369
            M.respond_inspiring(ctx, 0, M.ClientQuery(shard_id=0, rgsw=rows, has_packing_keys=True))
        assert str(e.value) == "InspiRING packing keys invalid: y_all and y_body are both empty"
        # keys check precedes the shard lookup in respond_inspiring
        with pytest.raises(M.PirError) as e:
            M.respond_inspiring(ctx, 0, M.ClientQuery(shard_id=999, rgsw=rows))
        assert e.value.code == M.ERR_PACKING_KEYS_MISSING
        # empty shard: zero ciphertext and one zero column (respond.rs:84-90, 307-313)
        r = M.respond(ctx, 2, M.ClientQuery(shard_id=0, rgsw=rows))
        assert not r.ciphertext.any() and len(r.column_ciphertexts) == 1
        r = M.respond_inspiring(ctx, 2, M.ClientQuery(shard_id=0, rgsw=rows, has_packing_keys=True))
        assert not r.ciphertext.any() and len(r.column_ciphertexts) == 1
        with pytest.raises(M.PirError) as e:
            M.respond(ctx, 5, M.ClientQuery(shard_id=0, rgsw=rows))
        assert e.value.code == M.ERR_UNKNOWN_DATABASE
        with pytest.raises(M.PirError) as e:
            M.respond(ctx, 0, M.ClientQuery(shard_id=0, rgsw=inp["b"], seeds=inp["seeds"]))
        assert e.value.code == M.ERR_INVALID_ARGUMENT


def test_batch_mixes_variants_and_databases(M, inp):
    """Each query's response in a mixed batch equals its response alone."""
    say("mixed batch")
    crt = inp["crt"]
    d = crt.params.d
    db_a = random_db(3 * d, 32, seed=41)
    db_b = random_db(d + 5, 32, seed=42)
    sa, sb = encode_db(crt, db_a, 32), encode_db(crt, db_b, 32)
    cl = Client(crt, 44)
    yb, ya = cl.packing_keys(16, W_SEED)
    qs, dbs, vs = [], [], []
    g = np.random.default_rng(9)
    for i in range(12):
        which = i % 2
        shard = int(g.integers(0, 3 if which == 0 else 2))
        local = int(g.integers(0, d))
        v = [M.NO_PACKING, M.ONE_PACKING, M.TWO_PACKING][i % 3]
        if v == M.TWO_PACKING or i % 4 == 1:
            seeds, b = cl.seeded_query(local)
            q = M.ClientQuery(shard_id=shard, rgsw=b, seeds=seeds, has_packing_keys=True, y_body=yb)
        else:
            q = M.ClientQuery(shard_id=shard, rgsw=cl.query_rows(local), has_packing_keys=True, y_body=yb,
                              y_all=ya if i % 5 == 0 else None)
        qs.append(q)
        dbs.append(which)
        vs.append(v)
    qs.append(M.ClientQuery(shard_id=77, rgsw=cl.query_rows(0)))  # fails alone, must not poison the rest
    dbs.append(0)
    vs.append(M.NO_PACKING)
    with make_ctx(M, crt.params) as ctx:
        ctx.load_database(0, sa)
        ctx.load_database(1, sb)
        res = ctx.respond_batch(dbs, qs, vs)
        assert isinstance(res[-1], M.PirError) and res[-1].code == M.ERR_SHARD_NOT_FOUND
        for i, (q, db, v) in enumerate(zip(qs[:-1], dbs[:-1], vs[:-1])):
            alone = M.respond_with_variant(ctx, db, q, v)
            assert (res[i].ciphertext == alone.ciphertext).all(), i
            assert (res[i].column_ciphertexts == alone.column_ciphertexts).all(), i
            # and against the oracle
            shards = sa if db == 0 else sb
            rows = crt.seeded_rgsw_expand(q.seeds, q.rgsw) if q.seeded else q.rgsw
            if v == M.NO_PACKING:
                ct, _ = crt.respond(shards[q.shard_id], rows)
            else:
                _, ct = crt.respond_inspiring(W_SEED, shards[q.shard_id], rows, y_body=q.y_body, y_all=q.y_all)
            assert (res[i].ciphertext == ct).all(), i


@pytest.mark.parametrize("variant", ["two", "zero", "one_full"])
def test_device_batch_matches_host_api(M, variant):
    """inspire_respond_batch_device (inputs resident in HBM) == the host API."""
    say("device batch", variant)
    o = O.Oracle(O.baseline_params())
    d, K = o.params.d, o.params.K
    db = random_db(3 * d, 32, seed=51)
    shards = encode_db(o, db, 32)
    cl = Client(o, 52)
    yb, _ = cl.packing_keys(16, W_SEED, with_y_all=False)
    g = np.random.default_rng(53)
    n = 7
    sids = g.integers(0, 3, n).astype(np.uint32)
    sids[3] = 99  # unknown shard: reported per query, the others still run
    locals_ = g.integers(0, d, n)
    qs, rows, seeds, ys = [], [], [], []
    for i in range(n):
        if variant == "one_full":
            full = cl.query_rows(int(locals_[i]))
            rows.append(full[3:6].astype(np.uint32))
            qs.append(M.ClientQuery(shard_id=int(sids[i]), rgsw=full, has_packing_keys=True, y_body=yb))
        else:
            s, b = cl.seeded_query(int(locals_[i]))
            rows.append(b[3:6].astype(np.uint32))
            seeds.append(np.frombuffer(s[3:6].tobytes(), dtype="<u4").reshape(3, 8))
            qs.append(M.ClientQuery(shard_id=int(sids[i]), rgsw=b, seeds=s, has_packing_keys=True, y_body=yb))
        ys.append(yb.astype(np.uint32))
    v = {"two": M.TWO_PACKING, "zero": M.NO_PACKING, "one_full": M.ONE_PACKING}[variant]
    cols = 16
    per = (2 if v != M.NO_PACKING else 2 * (1 + cols)) * K * d
    with make_ctx(M, o.params) as ctx:
        ctx.load_database(0, shards)
        drows = M.DeviceArray.from_numpy(np.stack(rows))
        dseeds = M.DeviceArray.from_numpy(np.stack(seeds)) if seeds else None
        dy = M.DeviceArray.from_numpy(np.stack(ys))
        dout = M.DeviceArray(n * per * 4)
        st = ctx.respond_batch_device(0, v, sids, drows.ptr, dout.ptr, seeds_ptr=dseeds.ptr if dseeds else 0,
                                      y_body_ptr=dy.ptr if v != M.NO_PACKING else 0, seeded=bool(seeds))
        out = dout.download(np.zeros((n, per), dtype=np.uint32))
        for i in range(n):
            if sids[i] == 99:
                assert st[i] == M.ERR_SHARD_NOT_FOUND
                continue
            assert st[i] == M.OK
            ref = M.respond_with_variant(ctx, 0, qs[i], v)
            if v == M.NO_PACKING:
                want = np.concatenate([ref.ciphertext.reshape(-1), ref.column_ciphertexts.reshape(-1)])
            else:
                want = ref.ciphertext.reshape(-1)
            assert np.array_equal(out[i].astype(np.uint64), want), i


@pytest.mark.parametrize("variant,n", [("two", 48), ("zero", 24)])
def test_host_batch_many_queries_matches_device_batch(M, variant, n):
    """A host-buffer batch large enough to be staged and handed back on several
    host threads equals the same queries as a device-resident batch, and a
    sample equals the oracle's responses."""
    say("host batch", variant, n)
    o = O.Oracle(O.baseline_params())
    d, K = o.params.d, o.params.K
    db = random_db(4 * d, 32, seed=61)
    shards = encode_db(o, db, 32)
    cl = Client(o, 62)
    yb, _ = cl.packing_keys(16, W_SEED, with_y_all=False)
    g = np.random.default_rng(63)
    sids = g.integers(0, 4, n).astype(np.uint32)
    locals_ = g.integers(0, d, n)
    qs, rows, seeds = [], [], []
    for i in range(n):
        s, b = cl.seeded_query(int(locals_[i]))
        rows.append(b[3:6].astype(np.uint32))
        seeds.append(np.frombuffer(s[3:6].tobytes(), dtype="<u4").reshape(3, 8))
        qs.append(M.ClientQuery(shard_id=int(sids[i]), rgsw=b, seeds=s, has_packing_keys=True, y_body=yb))
    v = {"two": M.TWO_PACKING, "zero": M.NO_PACKING}[variant]
    cols = 16
    per = (2 if v != M.NO_PACKING else 2 * (1 + cols)) * K * d
    with make_ctx(M, o.params) as ctx:
        ctx.load_database(0, shards)
        res = ctx.respond_batch([0] * n, qs, [v] * n)
        drows = M.DeviceArray.from_numpy(np.stack(rows))
        dseeds = M.DeviceArray.from_numpy(np.stack(seeds))
        dy = M.DeviceArray.from_numpy(np.stack([yb.astype(np.uint32)] * n))
        dout = M.DeviceArray(n * per * 4)
        st = ctx.respond_batch_device(0, v, sids, drows.ptr, dout.ptr, seeds_ptr=dseeds.ptr,
                                      y_body_ptr=dy.ptr if v != M.NO_PACKING else 0, seeded=True)
        out = dout.download(np.zeros((n, per), dtype=np.uint32))
    for i in range(n):
        assert st[i] == M.OK and not isinstance(res[i], M.PirError), i
        if v == M.NO_PACKING:
            got = np.concatenate([res[i].ciphertext.reshape(-1), res[i].column_ciphertexts.reshape(-1)])
        else:
            got = res[i].ciphertext.reshape(-1)
        assert np.array_equal(out[i].astype(np.uint64), got), i
    for i in (0, n // 2, n - 1):
        full = o.seeded_rgsw_expand(qs[i].seeds, qs[i].rgsw)
        if v == M.NO_PACKING:
            ct, colct = o.respond(shards[qs[i].shard_id], full)
            assert np.array_equal(res[i].ciphertext, ct), i
        else:
            _, ct = o.respond_inspiring(W_SEED, shards[qs[i].shard_id], full, y_body=yb)
            assert np.array_equal(res[i].ciphertext, ct), i


@pytest.mark.parametrize("entry_size", [6, 34, 256])
def test_inspiring_general_column_counts(M, entry_size):
    """Column counts off the fused front end's shape, through the host-buffer
    batch API: gamma = 3 (generator 1366, even: the y_body rotations are
    coefficient maps computed as generate_rotations writes them), 17 (odd
    generator, not a power of two) and 128 (gamma^2 does not divide 2d).
    These run the general first dimension, b sums and Horner step 1.  y_body,
    y_all and a seeded query, each against the oracle; the device-resident
    batch matches for the odd generators and refuses the even one."""
    o = O.Oracle(O.baseline_params())
    d, ell = o.params.d, o.params.gadget_len
    db = random_db(d + 40, entry_size, seed=entry_size)
    shards = encode_db(o, db, entry_size)
    gamma = shards.shape[1]
    say("inspiring general gamma", gamma)
    cl = Client(o, 40 + entry_size)
    yb, ya = cl.packing_keys(gamma, W_SEED)
    rows = cl.query_rows(d // 5)
    seeds, b = cl.seeded_query(33)
    full_seeded = o.seeded_rgsw_expand(seeds, b)
    cases = [
        (0, M.ClientQuery(shard_id=0, rgsw=rows, has_packing_keys=True, y_body=yb), rows, None),
        (1, M.ClientQuery(shard_id=1, rgsw=rows, has_packing_keys=True, y_body=yb, y_all=ya), rows, ya),
        (0, M.ClientQuery(shard_id=0, rgsw=b, seeds=seeds, has_packing_keys=True, y_body=yb), full_seeded, None),
    ]
    wants = []
    for sh, _, rg, y_all in cases:
        rc, ct = o.respond_inspiring(W_SEED, shards[sh], rg, y_body=yb, y_all=y_all)
        assert rc == 0
        wants.append(ct)
    with make_ctx(M, o.params) as ctx:
        ctx.load_database(0, shards)
        res = ctx.respond_batch([0] * len(cases), [c[1] for c in cases], [M.TWO_PACKING] * len(cases))
        for r, want in zip(res, wants):
            assert not isinstance(r, M.PirError), r
            assert (r.ciphertext == want).all(), gamma
        # device-resident batch of the seeded query (y_body rotated on the device)
        drows = M.DeviceArray.from_numpy(np.ascontiguousarray(b[ell:], dtype=np.uint32))
        dseeds = M.DeviceArray.from_numpy(np.ascontiguousarray(seeds[ell:]).view(np.uint32))
        dy = M.DeviceArray.from_numpy(yb.astype(np.uint32))
        dout = M.DeviceArray(2 * 2 * d * 4)
        try:
            if gamma == 3:
                with pytest.raises(M.PirError) as e:
                    ctx.respond_batch_device(0, M.TWO_PACKING, [0], drows.ptr, dout.ptr, seeds_ptr=dseeds.ptr,
                                             y_body_ptr=dy.ptr)
                assert e.value.code == M.ERR_UNSUPPORTED
            else:
                st = ctx.respond_batch_device(0, M.TWO_PACKING, [0], drows.ptr, dout.ptr, seeds_ptr=dseeds.ptr,
                                              y_body_ptr=dy.ptr)
                assert (st == 0).all()
                got = dout.download(np.zeros((2, 2, d), dtype=np.uint32))
                assert (got == wants[2]).all(), gamma
        finally:
            for a in (drows, dseeds, dy, dout):
                a.free()
