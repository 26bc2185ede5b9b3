"""gfx950 NTT and external product vs the oracle (bit-exact).

The NTT entry points return the reference's NTT-domain representation
(2^64 Montgomery form, src/math/ntt.rs:223-246), so they compare directly with
NttContext::forward / inverse as restated by the oracle."""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

import oracle as O
from helpers import uniform_polys

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "oracle_vectors.json")


def say(*a):
    print("[gpu-ntt]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    import inspire_mi355x as M

    return M


def ctx_for(M, params: O.Params, flags=0):
    p = M.InspireParams(ring_dim=params.d, crt_moduli=params.moduli, p=params.p, gadget_base=params.gadget_base,
                        gadget_len=params.gadget_len, flags=flags)
    return M.Context(p, bytes(range(32)))


@pytest.mark.parametrize("d", [256, 512, 1024, 2048, 4096])
def test_ntt_forward_inverse_match_oracle(M, d):
    say("ntt parity d =", d)
    params = O.Params(d=d)
    o = O.Oracle(params)
    a = uniform_polys(o, 6, 100 + d)
    with ctx_for(M, params) as ctx:
        A = ctx.ntt_forward(a)
        assert (A == o.ntt_forward(a)).all()
        B = uniform_polys(o, 6, 200 + d)
        assert (ctx.ntt_inverse(B) == o.ntt_inverse(B)).all()
        assert (ctx.ntt_inverse(A) == a).all()


def test_ntt_golden_digest(M):
    from golden.make_golden import digest

    say("ntt golden")
    want = json.load(open(GOLDEN))["digests"]
    o = O.Oracle(O.baseline_params())
    a = uniform_polys(o, 8, 1)
    with ctx_for(M, O.baseline_params()) as ctx:
        assert digest(ctx.ntt_forward(a)) == want["ntt_forward_8"]
        assert digest(ctx.ntt_inverse(a)) == want["ntt_inverse_8"]


def test_ntt_roundtrip_4096_polys(M):
    """Kernel microbench shape: 4096 polynomials per modulus, forward+inverse
    round trip bit-exact, forward spot-checked against the oracle."""
    say("ntt roundtrip 4096 polys")
    o = O.Oracle(O.baseline_params())
    a = uniform_polys(o, 4096, 7)
    with ctx_for(M, O.baseline_params()) as ctx:
        A = ctx.ntt_forward(a)
        assert (ctx.ntt_inverse(A) == a).all()
        for i in (0, 1, 2047, 4095):
            assert (A[i] == o.ntt_forward(a[i])).all()


def test_external_product_matches_oracle(M):
    from golden.make_golden import digest

    say("external product golden + 256 pairs")
    want = json.load(open(GOLDEN))["digests"]
    o = O.Oracle(O.baseline_params())
    rl = uniform_polys(o, 4, 2, lead=(2,))
    rg = uniform_polys(o, 4, 3, lead=(6, 2))
    with ctx_for(M, O.baseline_params()) as ctx:
        assert digest(ctx.external_product(rl, rg)) == want["external_product_4"]
        rl = uniform_polys(o, 256, 12, lead=(2,))
        rg = uniform_polys(o, 256, 13, lead=(6, 2))
        out = ctx.external_product(rl, rg)
        for i in range(0, 256, 37):
            assert (out[i] == o.external_product(rl[i], rg[i])).all()


@pytest.mark.parametrize("d", [256, 4096])
def test_external_product_other_degrees(M, d):
    say("external product d =", d)
    params = O.Params(d=d)
    o = O.Oracle(params)
    rl = uniform_polys(o, 3, d, lead=(2,))
    rg = uniform_polys(o, 3, d + 1, lead=(6, 2))
    with ctx_for(M, params) as ctx:
        out = ctx.external_product(rl, rg)
    for i in range(3):
        assert (out[i] == o.external_product(rl[i], rg[i])).all()
