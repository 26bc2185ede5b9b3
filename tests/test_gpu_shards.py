"""Shard files (src/pir/mmap.rs) ingested straight into HBM: responses equal
those over the same shards loaded from encoded polynomials and the oracle's;
files that do not match the parameters are refused.  Bit-exact."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

import oracle as O
from helpers import Client, encode_db, random_db, save_shards_binary

pytestmark = pytest.mark.gpu

W_SEED = bytes(range(32))


def say(*a):
    print("[gpu-shards]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    import inspire_mi355x as M

    return M


def make_ctx(M, params: O.Params):
    p = M.InspireParams(ring_dim=params.d, crt_moduli=params.moduli, p=params.p, gadget_base=params.gadget_base,
                        gadget_len=params.gadget_len)
    return M.Context(p, W_SEED)


def test_shard_files_match_encoded_load_and_oracle(M, tmp_path):
    say("shard files vs encoded load")
    o = O.Oracle(O.baseline_params())
    d = o.params.d
    db = random_db(3 * d, 32, seed=121)
    shards = encode_db(o, db, 32)
    ids = [7, 3, 11]  # file ids need not follow the file order
    save_shards_binary(tmp_path, ids, shards, o.params.moduli, d)
    cl = Client(o, 122)
    yb, _ = cl.packing_keys(16, W_SEED, with_y_all=False)
    with make_ctx(M, o.params) as ctx:
        ctx.load_shard_files(0, str(tmp_path))
        ctx.load_database(1, shards, shard_ids=ids)
        for sid, polys in zip(ids, shards):
            rows = cl.query_rows(sid * 37 % d)
            r_file = M.respond(ctx, 0, M.ClientQuery(shard_id=sid, rgsw=rows))
            r_enc = M.respond(ctx, 1, M.ClientQuery(shard_id=sid, rgsw=rows))
            ct, cols = o.respond(polys, rows)
            assert (r_file.ciphertext == ct).all() and (r_file.column_ciphertexts == cols).all()
            assert (r_enc.ciphertext == r_file.ciphertext).all()
            q = M.ClientQuery(shard_id=sid, rgsw=rows, has_packing_keys=True, y_body=yb)
            p_file = M.respond_inspiring(ctx, 0, q)
            _, want = o.respond_inspiring(W_SEED, polys, rows, y_body=yb)
            assert (p_file.ciphertext == want).all()
        with pytest.raises(M.PirError) as e:
            M.respond(ctx, 0, M.ClientQuery(shard_id=0, rgsw=cl.query_rows(0)))
        assert e.value.code == M.ERR_SHARD_NOT_FOUND


def test_shard_files_refused_when_they_do_not_fit(M, tmp_path):
    say("mismatched shard files")
    o = O.Oracle(O.baseline_params())
    d = o.params.d
    shards = encode_db(o, random_db(2 * d, 32, seed=123), 32)
    with make_ctx(M, o.params) as ctx:
        bad_dim = tmp_path / "dim"
        save_shards_binary(bad_dim, [0], shards[:1], o.params.moduli, d // 2)
        with pytest.raises(M.PirError) as e:
            ctx.load_shard_files(0, str(bad_dim))
        assert e.value.code == M.ERR_INVALID_ARGUMENT and "ring dimension" in str(e.value)
        bad_mod = tmp_path / "moduli"
        save_shards_binary(bad_mod, [0], shards[:1], [o.params.moduli[1], o.params.moduli[0]], d)
        with pytest.raises(M.PirError) as e:
            ctx.load_shard_files(0, str(bad_mod))
        assert "moduli" in str(e.value)
        short = tmp_path / "short"
        save_shards_binary(short, [0], shards[:1], o.params.moduli, d)
        f = short / "shard-0000.bin"
        f.write_bytes(f.read_bytes()[:-8])
        with pytest.raises(M.PirError) as e:
            ctx.load_shard_files(0, str(short))
        assert "truncated" in str(e.value)
        mixed = tmp_path / "mixed"
        save_shards_binary(mixed, [0, 1], [shards[0], shards[1][:8]], o.params.moduli, d)
        with pytest.raises(M.PirError) as e:
            ctx.load_shard_files(0, str(mixed))
        assert e.value.code == M.ERR_UNSUPPORTED
        empty = tmp_path / "empty"
        save_shards_binary(empty, [4], [np.zeros((0,), dtype=np.uint64)], o.params.moduli, d)
        ctx.load_shard_files(2, str(empty))
        r = M.respond(ctx, 2, M.ClientQuery(shard_id=4, rgsw=Client(o, 124).query_rows(0)))
        assert not r.ciphertext.any() and len(r.column_ciphertexts) == 1  # respond.rs:84-90
