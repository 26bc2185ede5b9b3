"""Graph replay of device-resident batches (inspire_ctx_set_graphs): the
same batch shape and buffers arriving again are captured into a HIP graph
and replayed.  Every replay, with new contents in the same buffers and the
program's own eager work (another shape, uploads, synchronisations) between
replays, equals the eager launches bit for bit, for InsPIRe^2 (fused front
end), InsPIRe^0 and tree packing.  A profiled batch is captured in segments
split at the profiled kernel (the chain launched between the graphs with its
own events; the first dimension and the tree leaves graphs of their own
between two events on the stream): it matches too, and its events time
about what the eager launch's do."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from helpers import random_db

pytestmark = pytest.mark.gpu

W_SEED = bytes(range(32))
Q0, Q1 = 268369921, 249561089
D, K, ELL = 2048, 2, 3


def say(*a):
    print("[gpu-graphs]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    import inspire_mi355x as M

    return M


def residues(g, *shape):
    m = np.array([Q0, Q1], dtype=np.uint64)[:, None]
    return (g.integers(0, 2**63, shape + (K, D), dtype=np.uint64) % m).astype(np.uint32)


@pytest.mark.parametrize("variant,mode", [(2, 0), (0, 0), (2, 1)])
def test_graph_replays_equal_eager(M, variant, mode):
    g = np.random.default_rng(300 + variant + 10 * mode)
    n, nshards = 24, 12
    db = random_db(nshards * D, 32, seed=71)
    seeded = variant != 0
    out_words = {0: 17 * 2 * K * D, 2: 2 * K * D}[variant] if mode == 0 else K * D + K * 2 * D
    with M.Context(M.InspireParams(), W_SEED) as ctx:
        ctx.load_entries(0, db, 32)
        if mode == 1:
            ctx.set_galois_keys(residues(g, 11, ELL, 2).astype(np.uint64))
        drows = M.DeviceArray(n * ELL * (1 if seeded else 2) * K * D * 4)
        dseeds = M.DeviceArray(n * ELL * 8 * 4)
        dy = M.DeviceArray(n * ELL * K * D * 4)
        dout = M.DeviceArray(n * out_words * 4)
        other = M.DeviceArray(4 * out_words * 4)

        def fill():
            shards = g.integers(0, nshards, n, dtype=np.uint64).astype(np.uint32)
            drows.upload(residues(g, n, ELL) if seeded else residues(g, n, ELL, 2))
            dseeds.upload(g.integers(0, 2**32, (n, ELL, 8), dtype=np.uint64).astype(np.uint32))
            dy.upload(residues(g, n, ELL))
            return shards

        def run(shards, out=dout, count=n):
            st = ctx.respond_batch_device(0, variant, shards[:count], drows.ptr, out.ptr, seeds_ptr=dseeds.ptr,
                                          y_body_ptr=dy.ptr, seeded=seeded, packing_mode=mode)
            assert (st == 0).all()
            return out.download(np.zeros((count, out_words), dtype=np.uint32)) if out is dout else None

        try:
            for rep in range(4):
                shards = fill()
                ctx.set_graphs(False)  # eager reference of these contents (drops the graphs)
                want = run(shards)
                ctx.set_graphs(True)
                got = [run(shards) for _ in range(3)]  # eager, captured, replayed
                for j, x in enumerate(got):
                    assert np.array_equal(x, want), (rep, j)
                run(shards, out=other, count=4)  # eager work of another shape between replays
                got2 = run(shards)
                assert np.array_equal(got2, want), rep
                caps, reps = ctx.graph_stats()
                say("variant", variant, "mode", mode, "rep", rep, "captures", caps, "replays", reps)
                assert caps == rep + 1 and reps == 2 * (rep + 1)
            # profiled batches: eager on the first sighting of the (new) key, then
            # captured with the profiled kernel as a segment of its own, then replayed
            name = "tree_leaves" if mode else ("pack_chain" if variant else "first_dim")
            ctx.set_profiling(True)
            ctx.set_profiling_kernels([name])
            ctx.profile(name, reset=True)
            assert np.array_equal(run(shards), want)
            ms_eager, n_eager = ctx.profile(name, reset=True)
            for _ in range(5):
                assert np.array_equal(run(shards), want)
            ms, launches = ctx.profile(name, reset=True)
            say(name, "eager ms", ms_eager, "captured + replayed ms", ms, "launches", launches,
                "stats", ctx.graph_stats())
            assert n_eager == 1 and launches == 5 and ms > 0 and ctx.graph_stats() == (5, 12)
            # the replayed kernel's events time about what the eager launch's do
            assert 0.25 * ms_eager < ms / launches < 2.0 * ms_eager + 0.05
        finally:
            ctx.set_profiling(False)
            for a in (drows, dseeds, dy, dout, other):
                a.free()
