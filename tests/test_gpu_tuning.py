"""Kernel variants and their tuning on the GPU: the matrix-core (int8 MFMA)
forward NTT at d = 2048 equals the reference's NttContext::forward (the
oracle) and the vector-ALU kernel word for word; a respond whose database
digits and keys were transformed on the matrix cores equals the oracle's; the
packing chain with eight values per thread equals the one with four; and the
autotuner's report is complete, persists, and loads into another context."""
from __future__ import annotations

import json
import sys

import numpy as np
import pytest

import oracle as O
from helpers import Client, encode_db, random_db, uniform_polys

pytestmark = pytest.mark.gpu

W_SEED = bytes(range(32))


def say(*a):
    print("[gpu-tuning]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    import inspire_mi355x as M

    return M


def test_mfma_ntt_equals_oracle_and_valu(M):
    o = O.Oracle(O.baseline_params())
    a = uniform_polys(o, 37, 4242)
    with M.Context(M.InspireParams(), W_SEED) as ctx:
        ctx.set_kernel_choice("ntt_fwd_2048", 1)
        assert ctx.tuning()["ntt_fwd_2048"] == "mfma"
        say("mfma ntt")
        got = ctx.ntt_forward(a)
        ctx.set_kernel_choice("ntt_fwd_2048", 0)
        assert ctx.tuning()["ntt_fwd_2048"] == "valu"
        assert np.array_equal(got, ctx.ntt_forward(a))
        assert np.array_equal(got, o.ntt_forward(a))
        # edge values: zeros, q - 1 everywhere, a single 1 (the twiddle table itself)
        z = np.zeros_like(a[:3])
        for k, m in enumerate(o.params.moduli):
            z[1, k, :] = m - 1
            z[2, k, 5] = 1
        ctx.set_kernel_choice("ntt_fwd_2048", 1)
        assert np.array_equal(ctx.ntt_forward(z), o.ntt_forward(z))


@pytest.mark.parametrize("choice", [("ntt_fwd_2048", 1), ("chain_values_per_thread", 8)])
def test_respond_with_each_variant_equals_oracle(M, choice):
    """The variant is chosen before the database is loaded, so the database's
    digit transforms (and for the chain, the packing) run on it."""
    o = O.Oracle(O.baseline_params())
    db = random_db(2048, 32, seed=61)
    shards = encode_db(o, db, 32)
    cl = Client(o, 62)
    seeds, b = cl.seeded_query(777)
    y_body, _ = cl.packing_keys(16, W_SEED, with_y_all=False)
    rows = o.seeded_rgsw_expand(seeds, b)
    rc, want = o.respond_inspiring(W_SEED, shards[0], rows, y_body=y_body)
    assert rc == 0
    want0, want0_cols = o.respond(shards[0], rows)
    with M.Context(M.InspireParams(), W_SEED) as ctx:
        ctx.set_kernel_choice(*choice)
        say("respond with", choice)
        ctx.load_database(0, shards)
        q = M.ClientQuery(shard_id=0, rgsw=b, seeds=seeds, has_packing_keys=True, y_body=y_body)
        assert np.array_equal(M.respond_seeded_inspiring(ctx, 0, q).ciphertext, want)
        r0 = M.respond_seeded(ctx, 0, M.ClientQuery(shard_id=0, rgsw=b, seeds=seeds))
        assert np.array_equal(r0.ciphertext, want0) and np.array_equal(r0.column_ciphertexts, want0_cols)


def test_autotune_report_persists(M, tmp_path):
    path = str(tmp_path / "tuning.json")
    with M.Context(M.InspireParams(), W_SEED) as ctx:
        say("autotune")
        rep = ctx.autotune(path)
        say(rep)
        m = rep["measured"]
        assert m["ntt_mfma_bit_exact"] and m["chain_v8_bit_exact"]
        assert m["ntt_fwd_2048_valu_ms"] > 0 and m["ntt_fwd_2048_mfma_ms"] > 0
        assert m["chain_v4_ms"] > 0 and m["chain_v8_ms"] > 0
        want_ntt = "mfma" if m["ntt_fwd_2048_mfma_ms"] < m["ntt_fwd_2048_valu_ms"] else "valu"
        want_v = 8 if m["chain_v8_ms"] < m["chain_v4_ms"] else 4
        assert rep["ntt_fwd_2048"] == want_ntt and rep["chain_values_per_thread"] == want_v
        assert json.load(open(path))["ntt_fwd_2048"] == want_ntt
    with M.Context(M.InspireParams(), W_SEED) as ctx2:
        ctx2.set_kernel_choice("ntt_fwd_2048", 1 if want_ntt == "valu" else 0)
        ctx2.load_tuning(path)
        t = ctx2.tuning()
        assert t["ntt_fwd_2048"] == want_ntt and t["chain_values_per_thread"] == want_v


def test_measured_issue_rates_are_plausible(M):
    """The register-resident probes report device-wide rates: a Montgomery
    product is three multiply instructions, so it issues slower than one
    64-bit multiply-add, and both lie within what 256 CUs can issue.  The
    streaming read rate lies below the published 8 TB/s of HBM3E and above
    a quarter of it."""
    with M.Context(M.InspireParams(), W_SEED) as ctx:
        r = ctx.measure_rates()
        say("rates", r)
        assert 1.0 < r["mmul_tops"] < 100.0 and 1.0 < r["smmul_tops"] < 100.0
        assert r["mad64_tops"] > r["mmul_tops"]
        assert 2.0 < r["hbm_read_tbs"] < 8.0
