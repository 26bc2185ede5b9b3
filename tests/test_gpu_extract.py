"""gfx950 client-side extract (src/pir/extract.rs: extract,
extract_with_variant, extract_inspiring) against the oracle, bit for bit, on
responses of every kind: InsPIRe^0 with and without column ciphertexts,
InspiRING in the reference's mode and the corrected one, tree-packed
responses with the reference's two-modulus b of 2d coefficients, a single
modulus, ragged entry sizes, device-resident batches in the layouts
inspire_respond_batch_device writes, and the error paths."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle as O
from helpers import Client, encode_db, random_db

pytestmark = pytest.mark.gpu

W_SEED = bytes(range(32))


def say(*a):
    print("[gpu-extract]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    import inspire_mi355x as M

    return M


@pytest.fixture(scope="module")
def inp():
    from golden.make_golden import inputs

    return inputs()


def make_ctx(M, params: O.Params, flags=0):
    p = M.InspireParams(ring_dim=params.d, crt_moduli=params.moduli, p=params.p, gadget_base=params.gadget_base,
                        gadget_len=params.gadget_len, flags=flags)
    return M.Context(p, W_SEED)


def test_extract_inspire0(M, inp):
    say("extract inspire0")
    crt, cl = inp["crt"], inp["cl"]
    db = inp["db"]
    shards = encode_db(crt, db, 32)
    with make_ctx(M, crt.params) as ctx:
        ctx.load_database(0, shards)
        r = M.respond(ctx, 0, M.ClientQuery(shard_id=0, rgsw=inp["rows"]))
        assert M.extract(ctx, r, cl.sk, 32) == db[100 * 32:101 * 32]
        for es in (0, 1, 2, 31, 32, 33, 40, 64):
            want = crt.extract(r.ciphertext, r.column_ciphertexts, cl.sk, es)
            assert M.extract(ctx, r, cl.sk, es) == want, es
            assert M.extract_with_variant(ctx, r, cl.sk, es, M.NO_PACKING) == want, es
        # a response without column ciphertexts: coefficient 0 of the ciphertext for every column
        bare = M.ServerResponse(r.ciphertext)
        for es in (32, 1000):
            assert M.extract(ctx, bare, cl.sk, es) == crt.extract(r.ciphertext, None, cl.sk, es), es


def test_extract_inspiring_both_modes(M, inp):
    say("extract inspiring")
    crt, cl = inp["crt"], inp["cl"]
    small = encode_db(crt, inp["db_small"], 32)
    for flags, o in ((0, crt), (M.FLAG_RESIDUE_WISE_AUTOMORPHISM, O.Oracle(O.baseline_params(), residue_wise=True))):
        with make_ctx(M, crt.params, flags=flags) as ctx:
            ctx.load_database(0, small)
            q = M.ClientQuery(shard_id=0, rgsw=inp["rows"], has_packing_keys=True, y_body=inp["yb"])
            r = M.respond_inspiring(ctx, 0, q)
            for es in (2, 17, 32, 4096):
                want = o.extract_inspiring(r.ciphertext, cl.sk, es)
                assert M.extract_inspiring(ctx, r, cl.sk, es) == want, (flags, es)
                assert M.extract_with_variant(ctx, r, cl.sk, es, M.TWO_PACKING) == \
                    o.extract_packed(r.ciphertext, cl.sk, es), (flags, es)
            if flags:  # the corrected mode decrypts to the entry
                assert M.extract_inspiring(ctx, r, cl.sk, 32) == inp["db_small"][100 * 32:101 * 32]


def test_extract_tree_wide_b(M, inp):
    """Tree packing with two moduli leaves b with 2d coefficients per residue;
    decrypt reads its first 2d words (Poly's Add, src/math/poly.rs:710-736)."""
    say("extract tree")
    crt, cl = inp["crt"], inp["cl"]
    shards = encode_db(crt, inp["db"], 32)
    with make_ctx(M, crt.params) as ctx:
        ctx.load_database(0, shards)
        ctx.set_galois_keys(inp["gk"])
        r = M.respond_one_packing(ctx, 0, M.ClientQuery(shard_id=0, rgsw=inp["rows"]))
        assert r.b_dim == 2 * crt.params.d
        for v in (M.ONE_PACKING, M.TWO_PACKING):
            for es in (1, 32, 33):
                assert M.extract_with_variant(ctx, r, cl.sk, es, v) == \
                    crt.extract_packed(r.ciphertext, cl.sk, es), (v, es)


@pytest.mark.parametrize("p", [65536, 65537])
def test_extract_tree_corrected_mode(M, p):
    """With p odd, d has an inverse mod p and extract_packed returns the
    entry; with p = 2^16 the values come back times d."""
    say("extract tree corrected", p)
    params = O.Params(d=256, moduli=(O.Q0, O.Q1), p=p)
    o = O.Oracle(params, residue_wise=True)
    cl = Client(o, 6)
    gk = o.galois_keys(cl.sk, o.rng(bytes(range(7, 39))))
    db = random_db(256, 32, seed=p, values_below_256=True)
    shards = encode_db(o, db, 32)
    with make_ctx(M, params, flags=M.FLAG_RESIDUE_WISE_AUTOMORPHISM) as ctx:
        ctx.load_database(0, shards)
        ctx.set_galois_keys(gk)
        r = M.respond_one_packing(ctx, 0, M.ClientQuery(shard_id=0, rgsw=cl.query_rows(77)))
        got = M.extract_with_variant(ctx, r, cl.sk, 32, M.ONE_PACKING)
        assert got == o.extract_packed(r.ciphertext, cl.sk, 32)
        if p % 2:
            assert got == db[77 * 32:78 * 32]
        # 2d bytes is the most a packed ciphertext holds
        assert M.extract_with_variant(ctx, r, cl.sk, 512, M.ONE_PACKING) == o.extract_packed(r.ciphertext, cl.sk, 512)


def test_extract_single_modulus(M):
    """One 28-bit modulus: Delta = 2^12 is below the external product's noise
    here, so the entry does not come back; the bytes still match the oracle."""
    say("extract single modulus")
    params = O.Params(d=256, moduli=(O.Q0,))
    o = O.Oracle(params)
    cl = Client(o, 12)
    db = random_db(256, 24, seed=3)
    shards = encode_db(o, db, 24)
    with make_ctx(M, params) as ctx:
        ctx.load_database(0, shards)
        r = M.respond(ctx, 0, M.ClientQuery(shard_id=0, rgsw=cl.query_rows(200)))
        for es in (24, 5):
            assert M.extract(ctx, r, cl.sk, es) == o.extract(r.ciphertext, r.column_ciphertexts, cl.sk, es), es


@pytest.mark.parametrize("d,moduli", [(512, (O.Q0, O.Q1)), (1024, (O.Q1,)), (4096, (O.Q0, O.Q1))])
def test_extract_other_degrees(M, d, moduli):
    """Every ring degree and modulus count of the kernel (the key in LDS is
    64 KiB at d = 4096 with two moduli): InsPIRe^0 responses through extract,
    and their ciphertext through the packed modes."""
    say("extract shape", d, moduli)
    params = O.Params(d=d, moduli=moduli)
    o = O.Oracle(params)
    cl = Client(o, d + 1)
    db = random_db(d, 32, seed=d)
    shards = encode_db(o, db, 32)
    with make_ctx(M, params) as ctx:
        ctx.load_database(0, shards)
        r = M.respond(ctx, 0, M.ClientQuery(shard_id=0, rgsw=cl.query_rows(d // 3)))
        assert M.extract(ctx, r, cl.sk, 32) == o.extract(r.ciphertext, r.column_ciphertexts, cl.sk, 32)
        for es in (32, 2 * d):
            assert M.extract_inspiring(ctx, r, cl.sk, es) == o.extract_inspiring(r.ciphertext, cl.sk, es), es
            assert M.extract_with_variant(ctx, r, cl.sk, es, M.ONE_PACKING) == \
                o.extract_packed(r.ciphertext, cl.sk, es), es
        if len(moduli) == 2:  # two moduli decrypt the entry back
            assert M.extract(ctx, r, cl.sk, 32) == db[(d // 3) * 32:(d // 3 + 1) * 32]


def test_extract_batch_device(M, inp):
    """Responses straight from inspire_respond_batch_device (InspiRING
    [count][2][K][d] and InsPIRe^0 [count][1 + cols][2][K][d]) into the
    device batch extract, per response bit-exact."""
    say("extract device batch")
    crt = inp["crt"]
    K, d, ell = 2, crt.params.d, crt.params.gadget_len
    db = inp["db"]
    shards = encode_db(crt, db, 32)
    n = 24
    clients = [Client(crt, 300 + i) for i in range(n)]
    idx = [(37 * i + 5) % d for i in range(n)]
    rows = np.stack([c.query_rows(i) for c, i in zip(clients, idx)])  # [n][2ell][2][K][d]
    yb = np.stack([c.packing_keys(16, W_SEED, with_y_all=False)[0] for c in clients])
    sk = np.stack([c.sk for c in clients]).astype(np.uint32)
    with make_ctx(M, crt.params) as ctx:
        ctx.load_database(0, shards)
        drows = M.DeviceArray.from_numpy(np.ascontiguousarray(rows[:, ell:]).astype(np.uint32))
        dyb = M.DeviceArray.from_numpy(yb.astype(np.uint32))
        dsk = M.DeviceArray.from_numpy(sk)
        for variant, words in ((M.TWO_PACKING, 2 * K * d), (M.NO_PACKING, 17 * 2 * K * d)):
            out = M.DeviceArray(n * words * 4)
            st = ctx.respond_batch_device(0, variant, [0] * n, drows.ptr, out.ptr, y_body_ptr=dyb.ptr, seeded=False)
            assert (st == 0).all()
            resp = out.download(np.zeros((n, words), dtype=np.uint32))
            for mode, es in ((M.EXTRACT_COLUMNS, 32), (M.EXTRACT_INSPIRING, 32), (M.EXTRACT_PACKED, 33)):
                ncols = 16 if variant == M.NO_PACKING else 0
                ent = M.DeviceArray(n * es)
                ctx.extract_batch_device(n, mode, es, out.ptr, dsk.ptr, ent.ptr, num_column_ciphertexts=ncols)
                got = ent.download(np.zeros((n, es), dtype=np.uint8))
                for i in range(n):
                    ct = resp[i, :2 * K * d].astype(np.uint64).reshape(2, K, d)
                    cols = resp[i, 2 * K * d:].astype(np.uint64).reshape(-1, 2, K, d) if ncols else None
                    if mode == M.EXTRACT_COLUMNS:
                        want = crt.extract(ct, cols, clients[i].sk, es)
                    elif mode == M.EXTRACT_INSPIRING:
                        want = crt.extract_inspiring(ct, clients[i].sk, es)
                    else:
                        want = crt.extract_packed(ct, clients[i].sk, es)
                    assert got[i].tobytes() == want, (variant, mode, i)
                    if variant == M.NO_PACKING and mode == M.EXTRACT_COLUMNS:
                        assert want == db[idx[i] * 32:(idx[i] + 1) * 32]
                ent.free()
            # a stride wider than one response: every other response
            ent = M.DeviceArray(n // 2 * 32)
            ctx.extract_batch_device(n // 2, M.EXTRACT_INSPIRING, 32, out.ptr, dsk.ptr, ent.ptr,
                                     stride=2 * words)
            got = ent.download(np.zeros((n // 2, 32), dtype=np.uint8))
            for j in range(n // 2):
                ct = resp[2 * j, :2 * K * d].astype(np.uint64).reshape(2, K, d)
                # secret key j goes with response 2j here: keys are indexed by batch position
                assert got[j].tobytes() == crt.extract_inspiring(ct, clients[j].sk, 32), (variant, j)
            ent.free()
            out.free()


def test_extract_errors(M, inp):
    say("extract errors")
    crt, cl = inp["crt"], inp["cl"]
    shards = encode_db(crt, inp["db"], 32)
    with make_ctx(M, crt.params) as ctx:
        ctx.load_database(0, shards)
        r = M.respond(ctx, 0, M.ClientQuery(shard_id=0, rgsw=inp["rows"]))
        with pytest.raises(M.PirError) as e:  # coefficient beyond d of a packed ciphertext
            M.extract_inspiring(ctx, r, cl.sk, 2 * 2048 + 1)
        assert e.value.code == M.ERR_INVALID_ARGUMENT
        with pytest.raises(M.PirError) as e:
            M.extract_with_variant(ctx, r, cl.sk, 32, 7)
        assert e.value.code == M.ERR_INVALID_ARGUMENT
        bad = cl.sk.copy()
        bad[1, 5] = O.Q1
        with pytest.raises(M.PirError) as e:
            M.extract(ctx, r, bad, 32)
        assert e.value.code == M.ERR_INVALID_ARGUMENT
        with pytest.raises(M.PirError) as e:
            ctx.extract_batch_device(1, 3, 32, 1, 1, 1)
        assert e.value.code == M.ERR_INVALID_ARGUMENT
        with pytest.raises(M.PirError) as e:
            ctx.extract_batch_device(1, M.EXTRACT_PACKED, 32, 1, 1, 1, b_dim=100)
        assert e.value.code == M.ERR_INVALID_ARGUMENT
        assert M.extract(ctx, r, cl.sk, 0) == b""
