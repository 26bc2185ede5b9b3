"""Wire formats of the respond path (bincode and serde_json of ServerResponse,
ClientQuery, SeededClientQuery; inspire-rs_amd/csrc/wire.cpp) against the
test-side codec in tests/wire_codec.py and Python's json, and their sizes
against the reference's measured payload sizes at d = 2048 with the two CRT
moduli (docs/COMMUNICATION_COSTS.md:210-219, measured with
benches/query_size_latency.rs).  CPU only: the codec touches no device."""
from __future__ import annotations

import json

import numpy as np
import pytest

import inspire_mi355x as M
import wire_codec as W

Q0, Q1 = M.Q0, M.Q1
MOD = (Q0, Q1)
D, K, ELL = 2048, 2, 3
GADGET = (1 << 20, 3, Q0 * Q1)


def rnd_polys(g, *lead, moduli=MOD, d=D):
    out = np.empty(tuple(lead) + (len(moduli), d), dtype=np.uint64)
    for k, m in enumerate(moduli):
        out[..., k, :] = g.integers(0, m, out[..., k, :].shape, dtype=np.uint64)
    return out


@pytest.fixture(scope="module")
def P():
    return M.InspireParams()


def kib(n):
    return round(n / 1024, 1)


def test_response_sizes_match_the_reference(P):
    """InsPIRe^0 1089.9 KB and packed 64.1 KB in bincode
    (COMMUNICATION_COSTS.md:217-219)."""
    g = np.random.default_rng(1)
    r0 = M.ServerResponse(rnd_polys(g, 2), rnd_polys(g, 16, 2), D)
    r2 = M.ServerResponse(rnd_polys(g, 2), np.zeros((0, 2, K, D), np.uint64), D)
    assert kib(len(M.response_to_binary(P, r0))) == 1089.9
    assert kib(len(M.response_to_binary(P, r2))) == 64.1


def test_query_sizes_match_the_reference(P):
    """Requests of 384.7 KB (full, no keys), 480.9 KB (full + keys) and
    288.7 KB (seeded + keys) in bincode (COMMUNICATION_COSTS.md:217-219),
    built by the test codec and decoded by the library."""
    g = np.random.default_rng(2)
    rows, yb = rnd_polys(g, 6, 2), rnd_polys(g, 3)
    seeds = g.integers(0, 256, (6, 32), dtype=np.uint8)
    q0 = W.query_obj(5, rows, MOD, D, GADGET, packing_mode="tree")
    q1 = W.query_obj(5, rows, MOD, D, GADGET, y_body=yb)
    q2 = W.query_obj(5, rows[:, 1], MOD, D, GADGET, seeds=seeds, y_body=yb)
    for o, want, seeded in ((q0, 384.7, False), (q1, 480.9, False), (q2, 288.7, True)):
        b = W.bin_query(o)
        assert kib(len(b)) == want
        q = M.query_from_binary(P, b, seeded)
        assert q.shard_id == 5 and q.seeded == seeded
        if seeded:
            assert np.array_equal(q.rgsw, rows[:, 1]) and np.array_equal(q.seeds, seeds)
        else:
            assert np.array_equal(q.rgsw, rows)
        assert q.has_packing_keys == ("inspiring_packing_keys" in o)
        if q.has_packing_keys:
            assert np.array_equal(q.y_body, yb)
    assert M.query_from_binary(P, W.bin_query(q0), False).packing_mode == M.PACKING_TREE


@pytest.mark.parametrize("cols", [0, 1, 16])
def test_response_bincode_is_the_reference_layout(P, cols):
    g = np.random.default_rng(3 + cols)
    ct, cc = rnd_polys(g, 2), rnd_polys(g, cols, 2)
    b = M.response_to_binary(P, M.ServerResponse(ct, cc, D))
    assert b == W.bin_response(W.response_obj(ct[0], ct[1], cc, MOD, D))
    o = W.read_bin_response(b)
    assert o["ciphertext"]["a"]["dim"] == D and o["ciphertext"]["a"]["crt_q0_inv_mod_q1"] == pow(Q0, -1, Q1)
    back = M.response_from_binary(P, b)
    assert np.array_equal(back.ciphertext, ct) and np.array_equal(back.column_ciphertexts, cc)


def test_tree_response_b_has_2d_coefficients(P):
    """A two-modulus tree-packed b is a Poly of dim 2d (automorph_pack.rs:476-484)."""
    g = np.random.default_rng(4)
    a, bb = rnd_polys(g), rnd_polys(g, d=2 * D)
    words = np.concatenate([a.reshape(-1), bb.reshape(-1)])
    r = M.ServerResponse(words, np.zeros((0, 2, K, D), np.uint64), 2 * D)
    b = M.response_to_binary(P, r)
    o = W.read_bin_response(b)
    assert o["ciphertext"]["b"]["dim"] == 2 * D and len(o["ciphertext"]["b"]["coeffs"]) == 4 * D
    assert b == W.bin_response(W.response_obj(a, bb, [], MOD, D, b_dim=2 * D))
    back = M.response_from_binary(P, b)
    assert back.b_dim == 2 * D and np.array_equal(back.a, a) and np.array_equal(back.b, bb)
    j = M.response_to_json(P, r)
    assert j == W.to_json(W.response_obj(a, bb, [], MOD, D, b_dim=2 * D))


@pytest.mark.parametrize("cols", [0, 16])
def test_response_json_is_serde_json(P, cols):
    """Byte-identical to the compact JSON of the same serde data model."""
    g = np.random.default_rng(5 + cols)
    ct, cc = rnd_polys(g, 2), rnd_polys(g, cols, 2)
    j = M.response_to_json(P, M.ServerResponse(ct, cc, D))
    assert j == W.to_json(W.response_obj(ct[0], ct[1], cc, MOD, D))
    assert list(json.loads(j)) == ["ciphertext", "column_ciphertexts"]


def test_response_json_size_is_near_the_reference(P):
    """The reference measured 1305.9 KB for an InsPIRe^0 response and
    76.8-76.9 KB for a packed one (COMMUNICATION_COSTS.md:217-219); the JSON
    size depends on the digits of the coefficients, so uniform residues land
    within 1 %."""
    g = np.random.default_rng(6)
    r0 = M.ServerResponse(rnd_polys(g, 2), rnd_polys(g, 16, 2), D)
    r2 = M.ServerResponse(rnd_polys(g, 2), np.zeros((0, 2, K, D), np.uint64), D)
    assert abs(kib(len(M.response_to_json(P, r0))) / 1305.9 - 1) < 0.01
    assert abs(kib(len(M.response_to_json(P, r2))) / 76.85 - 1) < 0.01


@pytest.mark.parametrize("seeded", [False, True])
@pytest.mark.parametrize("keys", [False, True])
def test_query_json_round_trip(P, seeded, keys):
    g = np.random.default_rng(7 + 2 * seeded + keys)
    rows, yb = rnd_polys(g, 6, 2), rnd_polys(g, 3)
    seeds = g.integers(0, 256, (6, 32), dtype=np.uint8) if seeded else None
    o = W.query_obj(77, rows[:, 1] if seeded else rows, MOD, D, GADGET, seeds=seeds, y_body=yb if keys else None)
    q = M.query_from_json(P, W.to_json(o))
    assert q.shard_id == 77 and q.seeded == seeded and q.packing_mode == M.PACKING_INSPIRING
    assert np.array_equal(q.rgsw, rows[:, 1] if seeded else rows)
    if seeded:
        assert np.array_equal(q.seeds, seeds)
    assert q.has_packing_keys == keys
    if keys:
        assert np.array_equal(q.y_body, yb)
    # the same query through bincode
    qb = M.query_from_binary(P, W.bin_query(o), seeded)
    assert np.array_equal(qb.rgsw, q.rgsw) and qb.has_packing_keys == keys


def test_query_json_defaults_and_field_order(P):
    """packing_mode defaults to inspiring (#[serde(default)]); keys may be
    null; fields may come in any order; whitespace is allowed."""
    g = np.random.default_rng(9)
    rows = rnd_polys(g, 6, 2)
    o = W.query_obj(3, rows, MOD, D, GADGET, with_mode=False)
    q = M.query_from_json(P, W.to_json(o))
    assert q.packing_mode == M.PACKING_INSPIRING and not q.has_packing_keys
    o2 = dict(reversed(list(W.query_obj(3, rows, MOD, D, GADGET, packing_mode="tree").items())))
    o2["inspiring_packing_keys"] = None
    q2 = M.query_from_json(P, json.dumps(o2, indent=1).encode())
    assert q2.packing_mode == M.PACKING_TREE and not q2.has_packing_keys
    assert np.array_equal(q2.rgsw, rows)


def test_query_decoding_refuses_malformed_input(P):
    g = np.random.default_rng(10)
    rows, yb = rnd_polys(g, 6, 2), rnd_polys(g, 3)
    good = W.query_obj(1, rows, MOD, D, GADGET, y_body=yb)

    def expect(code, obj=None, raw=None, binary=None):
        with pytest.raises(M.PirError) as e:
            if binary is not None:
                M.query_from_binary(P, binary, False)
            else:
                M.query_from_json(P, raw if raw is not None else W.to_json(obj))
        assert e.value.code == code, str(e.value)
        return str(e.value)

    bad = json.loads(W.to_json(good))
    bad["rgsw_ciphertext"]["rows"][0]["a"]["moduli"] = [Q0]
    expect(M.ERR_INVALID_ARGUMENT, bad)
    bad = json.loads(W.to_json(good))
    bad["rgsw_ciphertext"]["rows"][2]["b"]["coeffs"][5] = Q0  # not reduced
    expect(M.ERR_INVALID_ARGUMENT, bad)
    bad = json.loads(W.to_json(good))
    bad["rgsw_ciphertext"]["rows"][1]["b"]["dim"] = 1024
    expect(M.ERR_INVALID_ARGUMENT, bad)
    bad = json.loads(W.to_json(good))
    bad["rgsw_ciphertext"]["gadget"]["base"] = 1 << 16
    expect(M.ERR_UNSUPPORTED, bad)
    bad = json.loads(W.to_json(good))
    bad["inspiring_packing_keys"]["z_body"] = bad["inspiring_packing_keys"]["y_body"]
    expect(M.ERR_UNSUPPORTED, bad)
    bad = json.loads(W.to_json(good))
    bad["packing_mode"] = "full"
    expect(M.ERR_INVALID_ARGUMENT, bad)
    bad = json.loads(W.to_json(good))
    del bad["shard_id"]
    assert "shard_id" in expect(M.ERR_INVALID_ARGUMENT, bad)
    expect(M.ERR_INVALID_ARGUMENT, raw=W.to_json(good) + b"x")
    expect(M.ERR_INVALID_ARGUMENT, raw=W.to_json(good)[:-100])
    expect(M.ERR_INVALID_ARGUMENT, raw=W.to_json(good).replace(b'"shard_id":1', b'"shard_id":-1'))
    b = W.bin_query(good)
    expect(M.ERR_INVALID_ARGUMENT, binary=b[:-3])
    expect(M.ERR_INVALID_ARGUMENT, binary=b + b"\0")
    full = json.loads(W.to_json(good))
    full["inspiring_packing_keys"]["z_body"] = full["inspiring_packing_keys"]["y_body"][:1]
    expect(M.ERR_UNSUPPORTED, binary=W.bin_query(full))


def test_encoder_sizes_the_buffer(P):
    g = np.random.default_rng(11)
    r = M.ServerResponse(rnd_polys(g, 2), np.zeros((0, 2, K, D), np.uint64), D)
    keep: list = []
    rs = M._response_struct(P, r, keep)
    p = P._c()
    n = M.C.c_uint64()
    assert M.lib().inspire_response_to_binary(M.C.byref(p), M.C.byref(rs), None, 0, M.C.byref(n)) == M.ERR_CAPACITY
    assert n.value == 65658
    buf = M.C.create_string_buffer(n.value - 1)
    assert M.lib().inspire_response_to_binary(M.C.byref(p), M.C.byref(rs), buf, n.value - 1,
                                              M.C.byref(n)) == M.ERR_CAPACITY


def test_single_modulus_poly_has_no_crt_inverse():
    """crt_q0_inv_mod_q1 is 0 with one modulus (poly.rs:80-93)."""
    P1 = M.InspireParams(ring_dim=256, crt_moduli=(Q0,))
    g = np.random.default_rng(12)
    ct = rnd_polys(g, 2, moduli=(Q0,), d=256)
    b = M.response_to_binary(P1, M.ServerResponse(ct, np.zeros((0, 2, 1, 256), np.uint64), 256))
    o = W.read_bin_response(b)
    assert o["ciphertext"]["a"]["crt_q0_inv_mod_q1"] == 0 and o["ciphertext"]["a"]["q"] == Q0
    assert b == W.bin_response(W.response_obj(ct[0], ct[1], [], (Q0,), 256))


@pytest.mark.parametrize("ell", [0, 5])
def test_decoders_refuse_gadget_len_out_of_range(P, ell):
    """A parameter set with gadget_len outside [1, 4] is refused before any
    row is read: with gadget_len = 0 an empty row array would otherwise pass
    the 2 x gadget_len row count check."""
    bad = M.InspireParams(gadget_len=ell)
    empty = json.dumps({"shard_id": 1, "rgsw_ciphertext": {"rows": [], "gadget": {"base": 1 << 20, "len": ell,
                                                                                  "q": Q0 * Q1}}}).encode()
    with pytest.raises(M.PirError) as e:
        M.query_from_json(bad, empty)
    assert e.value.code == M.ERR_UNSUPPORTED
    b = np.zeros(8 + 8, dtype=np.uint8).tobytes()  # shard_id, then an empty row vector
    with pytest.raises(M.PirError) as e:
        M.query_from_binary(bad, b, False)
    assert e.value.code == M.ERR_UNSUPPORTED
    g = np.random.default_rng(13)
    with pytest.raises(M.PirError) as e:
        M.response_to_binary(bad, M.ServerResponse(rnd_polys(g, 2), np.zeros((0, 2, K, D), np.uint64), D))
    assert e.value.code == M.ERR_UNSUPPORTED
