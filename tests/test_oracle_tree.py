"""Pins the oracle's tree packing (respond_one_packing / pack_lwes /
homomorphic_automorph, src/inspiring/automorph_pack.rs) by what fixes its
outputs without the reference: the reference's own decrypt tests
(automorph_pack.rs:554-893) at its d=256 single-modulus parameters, decryption
of d=2048 two-modulus responses in the corrected mode, identities that tie the
reference's two-modulus b layout (automorph_pack.rs:476-484) to the corrected
b, and the zero-subtree skip against the literal recursion.  CPU only."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from helpers import Client, coeff0_message, encode_db, phase_error, random_db, rlwe_encrypt


@pytest.fixture(scope="module")
def single():
    o = O.Oracle(O.single_modulus_params(256))
    r = o.rng(bytes(range(32)))
    sk = o.gaussian_poly(r)
    return o, sk, r, o.galois_keys(sk, r)


@pytest.fixture(scope="module")
def crt():
    o = O.Oracle(O.baseline_params())
    r = o.rng(bytes(range(1, 33)))
    sk = o.gaussian_poly(r)
    return o, sk, r, o.galois_keys(sk, r)


def lwes(o, sk, r, msgs):
    la, lb = [], []
    for m in msgs:
        a, b = o.sample_extract(rlwe_encrypt(o, sk, coeff0_message(o, m), r))
        la.append(a)
        lb.append(b)
    return np.array(la), np.array(lb, dtype=np.uint64)


def dec(o, sk, a, b):
    return o.decrypt(np.stack([a, b]), sk)


def test_galois_keys_are_the_crs_set(single):
    """setup.rs:167-173: log_d matrices, matrix i for t = d/2^i + 1, each
    ell rows that decrypt (under s) to tau_t(s) * z^i."""
    o, sk, r, gk = single
    d, ell = o.params.d, o.params.gadget_len
    assert gk.shape == (8, ell, 2, 1, d)
    q = o.params.q
    for i in range(8):
        ts = o.apply_automorphism(sk, (d >> i) + 1)
        for j in range(ell):
            a, b = gk[i, j]
            phase = (o.mul_ntt(a, sk).astype(object) + b.astype(object)) % q
            want = (ts.astype(object) * (o.params.gadget_base ** j)) % q
            err = [(int(x) - int(y)) % q for x, y in zip(phase.ravel(), want.ravel())]
            assert max(min(e, q - e) for e in err) < 64


def test_homomorphic_automorph_identity(single):  # automorph_pack.rs:554-585
    o, sk, r, _ = single
    ks1 = o.automorphism_ks_matrix(sk, 1, r)
    ct = rlwe_encrypt(o, sk, coeff0_message(o, 12345), r)
    assert dec(o, sk, *o.homomorphic_automorph(ct, 1, ks1))[0] == 12345


def test_homomorphic_automorph_applies_tau(single):
    """Decrypts to tau_t(m): m = X + 2X^3 under t = 5 gives X^5 + 2X^15."""
    o, sk, r, gk = single
    d = o.params.d
    t = 5
    ks = o.automorphism_ks_matrix(sk, t, r)
    m = o.zeros()
    m[0, 1], m[0, 3] = 1, 2
    out = dec(o, sk, *o.homomorphic_automorph(rlwe_encrypt(o, sk, m, r), t, ks))
    want = np.zeros(d, dtype=np.uint64)
    want[5], want[15] = 1, 2
    assert (out == want).all()


def test_pack_lwes_trivial(single):  # automorph_pack.rs:779-835
    o, sk, r, gk = single
    d, q, p = o.params.d, o.params.q, o.params.p
    msgs = [100, 200, 300, 400]
    la = np.zeros((4, d), dtype=np.uint64)
    lb = np.array([o.params.delta * m % q for m in msgs], dtype=np.uint64)
    a, b = o.pack_lwes(la, lb, gk)
    assert (a == 0).all()  # a = 0 packs to exactly 0
    assert [int(x) for x in dec(o, sk, a, b)[:5]] == [m * d % p for m in msgs] + [0]


@pytest.mark.parametrize("msgs", [[100, 200], [100, 200, 300, 400]])
def test_pack_lwes_real(single, msgs):  # automorph_pack.rs:721-777, 837-893
    o, sk, r, gk = single
    d, p = o.params.d, o.params.p
    a, b = o.pack_lwes(*lwes(o, sk, r, msgs), gk)
    out = dec(o, sk, a, b)
    assert [int(x) for x in out[:len(msgs)]] == [m * d % p for m in msgs]
    assert (out[len(msgs):] == 0).all()


def test_pack_lwes_single_is_unscaled(single):  # automorph_pack.rs:436-443, 676-719
    o, sk, r, gk = single
    la, lb = lwes(o, sk, r, [12345])
    a, b = o.pack_lwes(la, lb, gk)
    assert dec(o, sk, a, b)[0] == 12345
    assert b[0, 0] == lb[0] and (b[0, 1:] == 0).all()
    want_a = np.zeros_like(la[0])
    want_a[0], want_a[1:] = la[0][0], la[0][1:][::-1]  # invert_sample_extract
    assert (a[0] == want_a).all()


@pytest.mark.parametrize("n", [2, 5, 16, 33])
def test_skip_equals_literal_recursion(single, n):
    """Skipping zero subtrees leaves every output word unchanged."""
    o, sk, r, gk = single
    la, lb = lwes(o, sk, r, [(7 * i + 3) % 256 for i in range(n)])
    a1, b1 = o.pack_lwes(la, lb, gk, literal=True)
    a2, b2 = o.pack_lwes(la, lb, gk, literal=False)
    assert (a1 == a2).all() and (b1 == b2).all()


def test_crt_corrected_mode_decrypts(crt):
    o, sk, r, gk = crt
    d, p = o.params.d, o.params.p
    msgs = [(5 * i + 1) % 32 for i in range(16)]
    a, b = o.pack_lwes(*lwes(o, sk, r, msgs), gk, residue_wise=True)
    assert b.shape == (2, d)
    out = dec(o, sk, a, b)
    assert [int(x) for x in out[:16]] == [m * d % p for m in msgs]
    assert (out[16:] == 0).all()


def test_crt_reference_b_layout(crt):
    """The reference's two-modulus b (automorph_pack.rs:476-484): the 2d raw
    residue words of the packed b, the first d plus s_z = b_z * d mod q, are
    split again as composite coefficients.  So, against the corrected b:
    residue 0 of the first d equals the corrected residue 0, both residues of
    the last d equal the packed residue-1 words r1 (r1 < q1 < q0), and
    corrected residue 1 = r1 + s_z mod q1; a is the same in both modes."""
    o, sk, r, gk = crt
    d, q = o.params.d, o.params.q
    q0, q1 = o.params.moduli
    la, lb = lwes(o, sk, r, [3, 9, 27, 1, 4])
    a_ref, b_ref = o.pack_lwes(la, lb, gk, residue_wise=False)
    a_fix, b_fix = o.pack_lwes(la, lb, gk, residue_wise=True)
    assert b_ref.shape == (2, 2 * d)
    assert (a_ref == a_fix).all()
    assert (b_ref[0, :d] == b_fix[0]).all()
    r1 = b_ref[0, d:]
    assert (b_ref[1, d:] == r1).all()
    s = [int(lb[z]) * d % q if z < len(lb) else 0 for z in range(d)]
    assert all(int(b_fix[1, z]) == (int(r1[z]) + s[z]) % q1 for z in range(d))
    # b_ref residue 1 of the first d is the composite sum reduced mod q1
    v0 = [(int(b_fix[0, z]) - s[z] % q0) % q0 for z in range(d)]  # packed residue 0
    assert all(int(b_ref[1, z]) == ((v0[z] + s[z]) % q) % q1 for z in range(d))


def test_crt_skip_equals_literal_at_2048(crt):
    """The full 2047-node recursion at d = 2048 against the 127 nonzero nodes."""
    o, sk, r, gk = crt
    la, lb = lwes(o, sk, r, list(range(16)))
    a1, b1 = o.pack_lwes(la, lb, gk, literal=True)
    a2, b2 = o.pack_lwes(la, lb, gk, literal=False)
    assert (a1 == a2).all() and (b1 == b2).all()


@pytest.mark.parametrize("which", ["single", "crt"])
def test_respond_one_packing_e2e(which, single, crt):
    """respond_one_packing (respond.rs:213-268): the response encrypts column c
    of the entry times d (mod p) at coefficient c and 0 elsewhere.  At d=256
    with one 60-bit modulus it decrypts exactly; at d=2048 with the two moduli
    the tree's noise (11 levels of key switching, each level doubling what
    came before, with digits up to 2^20) reaches about Delta, so decryption
    there is not exact and the check is that the phase stays within 4 Delta of
    the message at every coefficient (a wrong index, sign or key leaves errors
    near q/2 = 32768 Delta)."""
    o, sk, r, gk = single if which == "single" else crt
    d, p = o.params.d, o.params.p
    db = random_db(d, 32, seed=77, values_below_256=True)
    shards = encode_db(o, db, 32)
    cl = Client(o, 5)
    cl.sk = sk
    for idx in (0, 17, d - 1):
        a, b = o.respond_one_packing(shards[0], cl.query_rows(idx), gk, residue_wise=True)
        entry = np.frombuffer(db, dtype=np.uint8).reshape(d, 32)[idx]
        cols = [int(entry[2 * c]) | (int(entry[2 * c + 1]) << 8) for c in range(16)]
        want = [v * d % p for v in cols] + [0] * (d - 16)
        if which == "single":
            assert [int(x) for x in dec(o, sk, a, b)] == want
        else:
            assert phase_error(o, sk, a, b, want) < 4.0


def test_respond_one_packing_matches_pack_of_columns(crt):
    """The same columns through respond (InsPIRe^0 columns) and pack_lwes."""
    o, sk, r, gk = crt
    db = random_db(300, 32, seed=78)
    shard = encode_db(o, db, 32)[0]
    rows = Client(o, 6).query_rows(7)
    _, cols = o.respond(shard, rows)
    ext = [o.sample_extract(c) for c in cols]
    want = o.pack_lwes(np.array([e[0] for e in ext]), np.array([e[1] for e in ext], dtype=np.uint64), gk)
    got = o.respond_one_packing(shard, rows, gk)
    assert (want[0] == got[0]).all() and (want[1] == got[1]).all()


def test_respond_one_packing_empty_shard(crt):  # respond.rs:230-236
    o, sk, r, gk = crt
    a, b = o.respond_one_packing(np.zeros((0,) + o.P, dtype=np.uint64), Client(o, 1).query_rows(0), gk)
    assert (a == 0).all() and (b == 0).all() and b.shape == o.P
