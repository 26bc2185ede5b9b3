"""Writes tests/golden/oracle_vectors.json: SHA-256 digests (little-endian u64)
of the oracle's outputs on seeded inputs at the baseline parameters
(d=2048, CRT moduli [268369921, 249561089], p=2^16, z=2^20, ell=3) and at the
reference tests' single-modulus d=256 set.

The same cases are rebuilt by tests (CPU: the oracle must still produce these
digests; GPU: the HIP path must produce them too).  Run from the repo root:
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.dirname(HERE)):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle as O  # noqa: E402
from helpers import Client, encode_db, random_db, uniform_polys  # noqa: E402

W_SEED = bytes(range(32))


def inputs():
    """The seeded inputs of every golden case (shared with the GPU tests)."""
    crt = O.Oracle(O.baseline_params())
    single = O.Oracle(O.single_modulus_params(256))
    cl = Client(crt, 9)
    db = random_db(2048 + 100, 32, seed=5)
    db_small = random_db(2048, 32, seed=15, values_below_256=True)
    rows = cl.query_rows(100)
    seeds, b = cl.seeded_query(1000)
    yb, ya = cl.packing_keys(16, W_SEED)
    cs = Client(single, 9)
    db_s = random_db(256, 32, seed=25, values_below_256=True)
    # ServerCrs.galois_keys under the client's key, from a stream of their own
    gk = crt.galois_keys(cl.sk, crt.rng(bytes(range(40, 72))))
    return dict(crt=crt, single=single, cl=cl, db=db, db_small=db_small, rows=rows, seeds=seeds, b=b,
                yb=yb, ya=ya, cs=cs, db_s=db_s, rows_s=cs.query_rows(42), keys_s=cs.packing_keys(16, W_SEED), gk=gk)


def tree_words(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """A tree-packed ciphertext as the flat words a.coeffs ++ b.coeffs."""
    return np.concatenate([np.ascontiguousarray(a).reshape(-1), np.ascontiguousarray(b).reshape(-1)])


def build_cases(inp=None):
    inp = inp or inputs()
    crt, single = inp["crt"], inp["single"]
    cases = []
    a = uniform_polys(crt, 8, 1)
    cases.append(("ntt_forward_8", crt.ntt_forward(a)))
    cases.append(("ntt_inverse_8", crt.ntt_inverse(a)))
    rl = uniform_polys(crt, 4, 2, lead=(2,))
    rg = uniform_polys(crt, 4, 3, lead=(6, 2))
    cases.append(("external_product_4", np.stack([crt.external_product(rl[i], rg[i]) for i in range(4)])))
    shards = encode_db(crt, inp["db"], 32)
    ct, cols = crt.respond(shards[0], inp["rows"])
    cases.append(("respond0_full_ct", ct))
    cases.append(("respond0_full_cols", cols))
    rows_x = crt.seeded_rgsw_expand(inp["seeds"], inp["b"])
    ct, cols = crt.respond(shards[0], rows_x)
    cases.append(("respond0_seeded_ct", ct))
    small = encode_db(crt, inp["db_small"], 32)
    _, ct = crt.respond_inspiring(W_SEED, small[0], inp["rows"], y_body=inp["yb"])
    cases.append(("inspiring_y_body", ct))
    _, ct = crt.respond_inspiring(W_SEED, small[0], inp["rows"], y_body=inp["yb"], y_all=inp["ya"])
    cases.append(("inspiring_y_all", ct))
    _, ct = crt.respond_inspiring(W_SEED, small[0], rows_x, y_body=inp["yb"])
    cases.append(("inspiring_seeded_y_body", ct))
    rw = O.Oracle(O.baseline_params(), residue_wise=True)
    _, ct = rw.respond_inspiring(W_SEED, small[0], inp["rows"], y_body=inp["yb"])
    cases.append(("inspiring_residue_wise", ct))
    cases.append(("w_all_ntt_16", crt.w_all_ntt(16, W_SEED)))
    cases.append(("poly_from_seed", crt.poly_from_seed(bytes(inp["seeds"][3]))))
    s_shards = encode_db(single, inp["db_s"], 32)
    _, ct = single.respond_inspiring(W_SEED, s_shards[0], inp["rows_s"], y_body=inp["keys_s"][0])
    cases.append(("single_inspiring", ct))
    cases.append(("tree_full", tree_words(*crt.respond_one_packing(shards[0], inp["rows"], inp["gk"]))))
    cases.append(("tree_seeded", tree_words(*crt.respond_one_packing(shards[0], rows_x, inp["gk"]))))
    cases.append(("tree_residue_wise", tree_words(*rw.respond_one_packing(small[0], inp["rows"], inp["gk"]))))
    return cases


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u8").tobytes()).hexdigest()


def main():
    cases = build_cases()
    out = {
        "params": {"d": 2048, "moduli": [O.Q0, O.Q1], "p": 65536, "gadget_base": 1 << 20, "gadget_len": 3},
        "w_seed": W_SEED.hex(),
        "digests": {name: digest(arr) for name, arr in cases},
        "head": {name: [int(x) for x in np.asarray(arr).reshape(-1)[:8]] for name, arr in cases},
    }
    with open(os.path.join(HERE, "oracle_vectors.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(f"wrote {len(cases)} cases")


if __name__ == "__main__":
    main()
