"""The reference server's request/response flow over the gfx950 path
(src/bin/server.rs:120-236): a serde_json (or bincode) ClientQuery /
SeededClientQuery in, respond_with_variant / respond_seeded_with_variant on
the device, the ServerResponse out as serde_json and bincode -- byte for byte
the encodings of the oracle's response (tests/wire_codec.py)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle as O
import wire_codec as W
from helpers import Client, encode_db, random_db

pytestmark = pytest.mark.gpu

W_SEED = bytes(range(32))


def say(*a):
    print("[gpu-wire]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    import inspire_mi355x as M

    return M


@pytest.fixture(scope="module")
def inp():
    from golden.make_golden import inputs

    return inputs()


def test_json_query_to_json_response(M, inp):
    say("wire flow")
    crt = inp["crt"]
    d, mod = crt.params.d, crt.params.moduli
    gadget = (crt.params.gadget_base, crt.params.gadget_len, crt.params.q)
    shards = encode_db(crt, random_db(2 * d, 32, seed=91), 32)
    cl = Client(crt, 92)
    yb, _ = cl.packing_keys(16, W_SEED, with_y_all=False)
    P = M.InspireParams()
    with M.Context(P, W_SEED) as ctx:
        ctx.load_database(0, shards)
        ctx.set_galois_keys(inp["gk"])
        # /query_seeded, InsPIRe^2: seeded query with y_body keys
        seeds, b = cl.seeded_query(1500)
        body = W.to_json(W.query_obj(1, b, mod, d, gadget, seeds=seeds, y_body=yb))
        q = M.query_from_json(P, body)
        r = M.respond_seeded_with_variant(ctx, 0, q, M.TWO_PACKING)
        rows = crt.seeded_rgsw_expand(seeds, b)
        _, want = crt.respond_inspiring(W_SEED, shards[1], rows, y_body=yb)
        assert M.response_to_json(P, r) == W.to_json(W.response_obj(want[0], want[1], [], mod, d))
        assert M.response_to_binary(P, r) == W.bin_response(W.response_obj(want[0], want[1], [], mod, d))
        # /query, InsPIRe^0: full query, bincode in and out
        full = cl.query_rows(33)
        qb = M.query_from_binary(P, W.bin_query(W.query_obj(0, full, mod, d, gadget, packing_mode="tree")), False)
        r0 = M.respond_with_variant(ctx, 0, qb, M.NO_PACKING)
        ct, cols = crt.respond(shards[0], full)
        assert M.response_to_binary(P, r0) == W.bin_response(W.response_obj(ct[0], ct[1], cols, mod, d))
        # tree packing requested on the wire: b of dimension 2d in both encodings
        qt = M.query_from_json(P, W.to_json(W.query_obj(0, full, mod, d, gadget, packing_mode="tree")))
        rt = M.respond_with_variant(ctx, 0, qt, M.ONE_PACKING)
        a, bb = crt.respond_one_packing(shards[0], full, inp["gk"])
        assert M.response_to_json(P, rt) == W.to_json(W.response_obj(a, bb, [], mod, d, b_dim=2 * d))
        back = M.response_from_binary(P, M.response_to_binary(P, rt))
        assert back.b_dim == 2 * d and np.array_equal(back.a, a) and np.array_equal(back.b, bb)
