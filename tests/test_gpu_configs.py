"""The measured workloads at their full sizes, on the GPU, against the oracle.

  * 4096-polynomial NTT round trip and 1024 external products (d=2048, two
    CRT moduli): every 4th product compared with the oracle, all of them
    through a size-independent check (linearity in the RLWE operand);
  * InsPIRe^1 over a 2^20 x 32 B database: one query, bit-identical packed
    response;
  * InsPIRe^2 over a 2^26 x 32 B database resident in HBM: one 256-query
    device batch; 16 sampled queries compared with the oracle bit for bit, and
    three queries' responses equal to their responses in batches of their own.
Bit-exact throughout."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle as O
from helpers import Client, uniform_polys

pytestmark = pytest.mark.gpu

W_SEED = bytes(range(32))


def say(*a):
    print("[gpu-configs]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    import inspire_mi355x as M

    return M


def make_ctx(M, params: O.Params):
    p = M.InspireParams(ring_dim=params.d, crt_moduli=params.moduli, p=params.p, gadget_base=params.gadget_base,
                        gadget_len=params.gadget_len)
    return M.Context(p, W_SEED)


def test_external_products_1024_and_linearity(M):
    say("1024 external products")
    o = O.Oracle(O.baseline_params())
    rl = uniform_polys(o, 1024, 61, lead=(2,))
    rg = uniform_polys(o, 1024, 63, lead=(6, 2))
    mod = np.array(o.params.moduli, dtype=np.uint64)[:, None]
    with make_ctx(M, o.params) as ctx:
        out = ctx.external_product(rl, rg)
        for i in range(0, 1024, 4):
            assert (out[i] == o.external_product(rl[i], rg[i])).all(), i
        # The product is linear in the gadget digits.  For operands whose
        # composite coefficients are below the gadget base (the same small
        # value in both residues) the digits are (x, 0, 0), and x + y stays
        # below the base, so prod(x) + prod(y) == prod(x + y).
        g = np.random.default_rng(65)
        v1 = g.integers(0, 1 << 18, (1024, 2, 1, o.params.d), dtype=np.uint64)
        v2 = g.integers(0, 1 << 18, (1024, 2, 1, o.params.d), dtype=np.uint64)
        small = np.repeat(v1, o.params.K, axis=2)
        small2 = np.repeat(v2, o.params.K, axis=2)
        both = small + small2
        a = ctx.external_product(small, rg)
        b = ctx.external_product(small2, rg)
        c = ctx.external_product(both, rg)
        assert ((a + b) % mod == c).all()


def test_ntt_roundtrip_4096_polys_full(M):
    say("4096-poly NTT round trip")
    o = O.Oracle(O.baseline_params())
    a = uniform_polys(o, 4096, 64)
    with make_ctx(M, o.params) as ctx:
        A = ctx.ntt_forward(a)
        assert (ctx.ntt_inverse(A) == a).all()
        for i in range(0, 4096, 512):
            assert (A[i] == o.ntt_forward(a[i])).all()


def test_inspire1_2e20_single_query(M):
    say("InsPIRe^1, 2^20 x 32 B")
    o = O.Oracle(O.baseline_params())
    d = o.params.d
    g = np.random.default_rng(2020)
    db = np.frombuffer(g.bytes((1 << 20) * 32), dtype=np.uint8)
    cl = Client(o, 71)
    yb, _ = cl.packing_keys(16, W_SEED, with_y_all=False)
    with make_ctx(M, o.params) as ctx:
        ctx.load_entries(0, db, 32)
        for idx in (123456, (1 << 20) - 1):
            shard, local = divmod(idx, d)
            rows = cl.query_rows(local)
            q = M.ClientQuery(shard_id=shard, rgsw=rows, has_packing_keys=True, y_body=yb)
            r = M.respond_with_variant(ctx, 0, q, M.ONE_PACKING)
            rc, ct = o.respond_inspiring(W_SEED, o.encode_shard(db, 32, shard), rows, y_body=yb)
            assert rc == 0 and (r.ciphertext == ct).all(), idx
            assert len(r.column_ciphertexts) == 0


def test_inspire2_2e26_batch_sampled(M):
    say("InsPIRe^2, 2^26 x 32 B, 256-query device batch")
    o = O.Oracle(O.baseline_params())
    d, K, ell = o.params.d, o.params.K, o.params.gadget_len
    g = np.random.default_rng(2026)
    db = np.frombuffer(g.bytes((1 << 26) * 32), dtype=np.uint8)
    nshards = (1 << 26) // d
    B = 256
    mod = np.array(o.params.moduli, dtype=np.uint64)[:, None]
    b = (g.integers(0, 2**63, (B, ell, K, d), dtype=np.uint64) % mod).astype(np.uint32)
    y = (g.integers(0, 2**63, (B, ell, K, d), dtype=np.uint64) % mod).astype(np.uint32)
    seeds = g.integers(0, 2**32, (B, ell, 8), dtype=np.uint64).astype(np.uint32)
    shards = g.integers(0, nshards, B).astype(np.uint32)
    with make_ctx(M, o.params) as ctx:
        ctx.load_entries(0, db, 32)
        say("database resident", ctx.device_bytes / 2**30, "GiB")
        dr, ds, dy = (M.DeviceArray.from_numpy(x) for x in (b, seeds, y))
        dout = M.DeviceArray(B * 2 * K * d * 4)
        st = ctx.respond_batch_device(0, M.TWO_PACKING, shards, dr.ptr, dout.ptr, seeds_ptr=ds.ptr, y_body_ptr=dy.ptr)
        assert (st == 0).all()
        out = dout.download(np.zeros((B, 2, K, d), dtype=np.uint32))
        # alone == in the batch, for a few queries
        one = M.DeviceArray(2 * K * d * 4)
        per_q = ell * K * d * 4
        for i in (0, 77, 255):
            st1 = ctx.respond_batch_device(0, M.TWO_PACKING, shards[i:i + 1], dr.ptr + i * per_q, one.ptr,
                                           seeds_ptr=ds.ptr + i * ell * 32, y_body_ptr=dy.ptr + i * per_q)
            assert st1[0] == 0
            assert np.array_equal(one.download(np.zeros((2, K, d), dtype=np.uint32)), out[i]), i
        for x in (dr, ds, dy, dout, one):
            x.free()
    sample = list(range(0, B, 16))
    say("oracle on", len(sample), "sampled queries")
    o.pack(16, W_SEED)
    shard_polys, rgs, ys = [], [], []
    for i in sample:
        s = int(shards[i])
        shard_polys.append(o.encode_shard(db[s * d * 32:(s + 1) * d * 32], 32, 0))
        rg = np.zeros((2 * ell, 2, K, d), dtype=np.uint64)
        for r in range(ell):
            rg[ell + r, 0] = o.poly_from_seed(seeds[i, r].astype("<u4").tobytes())
            rg[ell + r, 1] = b[i, r]
        rgs.append(rg)
        ys.append(y[i].astype(np.uint64))
    rc, outs = o.respond_inspiring_many(W_SEED, shard_polys, rgs, ys)
    for j, i in enumerate(sample):
        assert rc[j] == 0
        assert np.array_equal(outs[j].astype(np.uint32), out[i]), i
