"""The oracle's client-side extract (src/pir/extract.rs) pinned by what the
mathematics fixes, not by a retyped formula:

  * reconstruct_entry against the reference's own vector (extract.rs:411-425);
  * extract of an InsPIRe^0 response is the database entry (decryption is
    exact below the noise bound), at the CRT baseline and at the reference
    tests' single-modulus set;
  * without column ciphertexts every column reads coefficient 0 of the summed
    ciphertext, i.e. the sum of the entry's column values mod p;
  * extract_inspiring of a corrected-mode InspiRING response is the entry;
  * extract_packed of a tree-packed response: coefficient c holds d v_c mod p,
    so with p = 2^16 (d has no inverse, the multiplier is 1) the values come
    back times d, and with an odd p they come back exactly;
  * a b of 2d coefficients per residue is read through its first 2d words."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from helpers import Client, encode_db, random_db

W_SEED = bytes(range(32))


def col_values(entry: bytes):
    e = entry + b"\0"
    return [e[2 * c] | (e[2 * c + 1] << 8) for c in range((len(entry) + 1) // 2)]


def test_reconstruct_entry_reference_vector():
    vals = [i * 256 + i for i in range(16)]
    assert O.reconstruct_entry(vals, 32) == bytes(x for i in range(16) for x in (i, i))
    # odd sizes drop the last high byte; bits above 16 are dropped
    assert O.reconstruct_entry([0x1234, 0x15678], 3) == bytes([0x34, 0x12, 0x78])


@pytest.mark.parametrize("params", [O.baseline_params(), O.single_modulus_params(256)],
                         ids=["crt2048", "single256"])
def test_extract_inspire0_is_the_entry(params):
    o = O.Oracle(params)
    d = params.d
    cl = Client(o, 3)
    for es in (32, 7):
        db = random_db(d, es, seed=es)
        shards = encode_db(o, db, es)
        idx = 5 * es + 1
        ct, cols = o.respond(shards[0], cl.query_rows(idx))
        want = db[idx * es:(idx + 1) * es]
        assert o.extract(ct, cols, cl.sk, es) == want
        # a shorter entry reads fewer columns; a longer one leaves bytes 0
        assert o.extract(ct, cols, cl.sk, 2) == want[:2]
        longer = o.extract(ct, cols, cl.sk, 2 * len(cols) + 4)
        assert longer[:es] == want[:es] and longer[2 * len(cols):] == bytes(4)
        # no column ciphertexts: coefficient 0 of the sum for every column
        s = sum(col_values(want)) % params.p
        assert o.extract(ct, None, cl.sk, es) == O.reconstruct_entry([s] * len(col_values(want)), es)


def test_extract_inspiring_corrected_mode_is_the_entry():
    o = O.Oracle(O.baseline_params(), residue_wise=True)
    cl = Client(o, 4)
    db = random_db(2048, 32, seed=8, values_below_256=True)
    shards = encode_db(o, db, 32)
    yb, _ = cl.packing_keys(16, W_SEED, with_y_all=False)
    _, ct = o.respond_inspiring(W_SEED, shards[0], cl.query_rows(300), y_body=yb)
    assert o.extract_inspiring(ct, cl.sk, 32) == db[300 * 32:301 * 32]


@pytest.mark.parametrize("p", [65536, 65537])
def test_extract_packed_tree(p):
    params = O.Params(d=256, moduli=(O.Q0, O.Q1), p=p)
    o = O.Oracle(params, residue_wise=True)
    d = params.d
    cl = Client(o, 6)
    gk = o.galois_keys(cl.sk, o.rng(bytes(range(7, 39))))
    db = random_db(d, 32, seed=p, values_below_256=True)
    shards = encode_db(o, db, 32)
    a, b = o.respond_one_packing(shards[0], cl.query_rows(77), gk)
    ct = np.concatenate([a.reshape(-1), b.reshape(-1)])
    want = db[77 * 32:78 * 32]
    got = o.extract_packed(ct, cl.sk, 32)
    if p % 2:
        assert got == want
    else:
        assert got == O.reconstruct_entry([v * d % p for v in col_values(want)], 32)


def test_decrypt_reads_the_first_2d_words_of_a_wide_b():
    o = O.Oracle(O.baseline_params())
    K, d = 2, 2048
    g = np.random.default_rng(1)
    a = np.stack([g.integers(0, m, d, dtype=np.uint64) for m in (O.Q0, O.Q1)])
    b = np.stack([g.integers(0, O.Q1, d, dtype=np.uint64) for _ in range(K)])
    tail = g.integers(0, O.Q0, K * d, dtype=np.uint64)
    narrow = np.concatenate([a.reshape(-1), b.reshape(-1)])
    wide = np.concatenate([narrow, tail])
    sk = Client(o, 2).sk
    assert np.array_equal(o.decrypt(wide, sk), o.decrypt(narrow, sk))
