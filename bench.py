#!/usr/bin/env python3
"""InsPIRe^2 respond throughput on one MI355X.

Workload (the headline configuration): InsPIRe^2 (seeded RGSW query +
InspiRING packing with y_body keys) at d=2048 over the CRT moduli
[268369921, 249561089], p=2^16, gadget 2^20 x 3, against a 2^26-entry x 32 B
database of uniform random bytes held resident in HBM.  A step is one batch of
256 queries with uniform shard ids; query material (seeds, b rows, y_body) is
uniform synthetic data of the real shape, already resident in HBM when the
timed region starts.

Prints one JSON line: queries/s over the timed steps, p50/p99 respond latency
(a query's latency is its batch's completion time), the speed-up over the same
path run one query per launch, the roofline of the dominant kernel (HIP events
inside the library over the timed region), and the CPU baseline (the oracle,
a C restatement of the reference respond, on this host's cores, on a bounded
sample) together with a bit-exact comparison of the sampled responses.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "inspire-rs_amd"))

METRIC = "PIR queries/sec + p50/p99 respond latency, InsPIRe^2, d=2048, 2^26x32B DB"
Q0, Q1 = 268369921, 249561089
D, K, ELL, GAMMA = 2048, 2, 3, 16

# Algorithmic modular multiplications per query in each kernel (one NTT of one
# residue = (d/2) log2 d = 11264; one pointwise product of one residue = d).
# The packing-chain kernel (k_pack_chain_crt) runs packing_offline step 2, the
# online sum, the a_hat and b inverse transforms, and the NTT of y_body.  Per
# step it transforms R (K INTTs) and ell-1 gadget digits (K NTTs each); the
# last digit's transform is recovered with 2 products per slot and residue
# (NTT(D_last) = z^-(ell-1) (R - sum_j z^j NTT(D_j)), exact because the digits
# sum to INTT(R) and the NTT is linear).
NTT = (D // 2) * 11
MODMUL_CHAIN = ((GAMMA - 1) * (2 + 2 * (ELL - 1)) * NTT + (GAMMA - 1) * 2 * K * D
                + (GAMMA - 1) * 2 * ELL * K * D + 4 * NTT + ELL * K * NTT)
MODMUL_QUERY = (
    ELL * (K + 1) * NTT          # RGSW rows ell..2ell-1: a of residue 0 and b of both (residue 1 of a is not read)
    + GAMMA * 2 * ELL * K * D    # first dimension MAC
    + GAMMA * GAMMA * D          # step 1 (residue 0)
    + MODMUL_CHAIN
)
# The same count for the reference's algorithm as written, with the database
# digits transformed per query (respond.rs:316-324 under external_product.rs:
# 122-150) and step 1 as the reference runs it (inspiring2.rs:952-982): RGSW
# rows 12 NTT; digit NTTs 16*3*K; column MAC 16*3*2*K*d; step 1 16*16*K*d plus
# the gamma^-1 scaling 16*K*d; step 2 15*(INTT + 3 NTT)*K plus its MAC
# 15*3*K*d; a_hat's INTT; online NTT(y_body) 3*K, MAC 45*K*d, INTT K.
MODMUL_QUERY_REF_ALGORITHM = (
    2 * ELL * K * NTT + GAMMA * ELL * K * NTT + GAMMA * ELL * 2 * K * D
    + GAMMA * GAMMA * K * D + GAMMA * K * D
    + (GAMMA - 1) * (1 + ELL) * K * NTT + (GAMMA - 1) * ELL * K * D
    + K * NTT
    + ELL * K * NTT + (GAMMA - 1) * ELL * K * D + K * NTT
)
# Algorithmic HBM bytes per query of the fused front end (first dimension +
# step 1 + b sums): the shard's NTT-domain digits, the RGSW rows it reads (b in
# both residues, a in residue 0, which is all the reference's residue handling
# reads), r_pre written for residue 0 and the per-block b sums.
FRONT_BYTES = (GAMMA * ELL * K * D * 4 + ELL * (K + 1) * D * 4 + GAMMA * D * 4
               + (D // 256) * GAMMA * K * 8)
# Published MI355X figures used for the roofs: HBM3E 8 TB/s; vector FP32
# 157.3 TFLOP/s = 78.6 T lane-FMA/s; 32-bit integer multiplies issue at half
# that rate and a Shoup/Montgomery modular product needs 3 of them.
HBM_PEAK_GBS = 8000.0
MODMUL_PEAK_T = 157.3 / 2 / 2 / 3
# Issue rates measured on the device at start-up (inspire_ctx_measure_rates):
# mmul_tops / smmul_tops (the transforms' Montgomery products), mad64_tops.
RATES: dict = {}


def tree_modmul(m: int = GAMMA):
    """Algorithmic modular multiplications per query of tree packing
    (respond_one_packing) over m columns: (leaves kernel, combine kernels).
    A tree node at level l costs the digit NTTs ((ell - 1) K: the last digit's
    transform is recovered from tau(s1.a^) with ell - 1 products per slot,
    exact since q <= z^ell), the key MAC (2 ell K d) and the INTT of the new a
    (K); nodes with an odd child also multiply odd.a^ and odd.b by y in the
    NTT domain (2 K d).  The leaves kernel adds the first-dimension MAC
    (2 ell K d per column) and each column's a INTT (K); the root adds b's
    INTT (K)."""
    L = D.bit_length() - 1
    node = ((ELL - 1) * K + K) * NTT + 2 * ELL * K * D + (ELL - 1) * K * D
    lm = (m - 1).bit_length()
    leaves = m * (2 * ELL * K * D + K * NTT) + m * (L - lm) * node
    combine = 0
    for lv in range(L - lm + 1, L + 1):
        half = 1 << (L - lv)
        nodes = min(m, half)
        odd = max(0, min(m - half, half))
        combine += nodes * node + odd * 2 * K * D
    return leaves, combine + K * NTT


def synthetic_queries(rng: np.random.Generator, n: int, nshards: int):
    moduli = np.array([Q0, Q1], dtype=np.uint64)[:, None]
    b = (rng.integers(0, 2**63, (n, ELL, K, D), dtype=np.uint64) % moduli).astype(np.uint32)
    y = (rng.integers(0, 2**63, (n, ELL, K, D), dtype=np.uint64) % moduli).astype(np.uint32)
    seeds = rng.integers(0, 2**32, (n, ELL, 8), dtype=np.uint64).astype(np.uint32)
    shards = rng.integers(0, nshards, n, dtype=np.uint64).astype(np.uint32)
    return shards, b, seeds, y


def measured_traffic(kernel: str, batch: int):
    """HBM bytes per launch of `kernel` at this batch size, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs summarised under profiles/ (lines
    `TRAFFIC <kernel> <batch> <fetch bytes> <write bytes>`, already scaled by
    the copy-kernel calibration), or None when no such run is recorded."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None
    for name in sorted(os.listdir(pdir)):
        if not name.endswith((".md", ".txt")):
            continue
        with open(os.path.join(pdir, name)) as f:
            for line in f:
                tok = line.split()
                if len(tok) == 5 and tok[0] == "TRAFFIC" and tok[1] == kernel and int(tok[2]) == batch:
                    return int(tok[3]) + int(tok[4])
    return None


def percentile(xs, p):
    return float(np.percentile(np.asarray(xs), p))


def host_info():
    """The host the CPU baseline ran on: CPU model, logical CPUs, the CPUs
    this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable}


def cpu_baseline(db_bytes: np.ndarray, qset, gpu_out: np.ndarray, budget_s: float, w_seed: bytes,
                 literal_queries: int = -1):
    """The oracle (C restatement of the reference respond) on this host's
    cores: InsPIRe^2 with the query-independent tables hoisted, one query per
    thread.  Also compares the sampled responses with the GPU's, bit for bit.
    Then the literal form (respond_inspiring as the reference runs it, per
    query: NTT contexts, mul_ntt external products, PackParams::new with the
    brute-force automorphism table of every odd t < 2d, and
    OfflinePackingKeys::generate) on `literal_queries` queries, one per
    thread (-1: one per thread; 0: skipped)."""
    sys.path.insert(0, ROOT)
    import oracle as O

    o = O.Oracle(O.baseline_params())
    shards, b, seeds, y = qset
    threads = O.lib.or_num_threads()

    def oracle_inputs(i):
        s = int(shards[i])
        chunk = db_bytes[s * D * 32:(s + 1) * D * 32]
        polys = o.encode_shard(chunk, 32, 0)
        rg = np.zeros((2 * ELL, 2, K, D), dtype=np.uint64)
        for r in range(ELL):
            rg[ELL + r, 0] = o.poly_from_seed(seeds[i, r].astype("<u4").tobytes())
            rg[ELL + r, 1] = b[i, r]
        return polys, rg, y[i].astype(np.uint64)

    o.pack(GAMMA, w_seed)  # tables built once (the hoisted variant)
    done, t_total, mismatches = 0, 0.0, 0
    n = len(shards)
    while done < n and (t_total < budget_s or done == 0):
        m = min(threads, n - done)
        ins = [oracle_inputs(done + j) for j in range(m)]
        t0 = time.perf_counter()
        rc, outs = o.respond_inspiring_many(w_seed, [x[0] for x in ins], [x[1] for x in ins], [x[2] for x in ins])
        t_total += time.perf_counter() - t0
        for j in range(m):
            if rc[j] != 0 or not np.array_equal(outs[j].astype(np.uint32), gpu_out[done + j]):
                mismatches += 1
        done += m
    out = {
        "value": done / t_total,
        "unit": "queries/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{done} InsPIRe^2 queries of the timed workload, {t_total:.2f} s wall on {threads} threads "
                  f"({t_total * threads:.1f} core-s); "
                  f"PackParams/OfflinePackingKeys hoisted; bit-exact vs GPU: {done - mismatches}/{done}",
        "bit_exact_sampled": f"{done - mismatches}/{done}",
        "host": host_info(),
        "literal": None,
    }
    nl = threads if literal_queries < 0 else min(literal_queries, n)
    if nl > 0:
        ins = [oracle_inputs(j) for j in range(nl)]
        t0 = time.perf_counter()
        rc, outs = o.respond_inspiring_literal_many(w_seed, [x[0] for x in ins], [x[1] for x in ins],
                                                    [x[2] for x in ins])
        t_lit = time.perf_counter() - t0
        ok = sum(1 for j in range(nl) if rc[j] == 0 and np.array_equal(outs[j].astype(np.uint32), gpu_out[j]))
        used = min(threads, nl)
        out["literal"] = {
            "value": nl / t_lit,
            "unit": "queries/s",
            "cores": used,
            "kind": "port",
            "sample": f"{nl} InsPIRe^2 queries of the timed workload, one per thread, {t_lit:.2f} s wall on "
                      f"{used} threads; per query: an NTT context per column, mul_ntt external products, "
                      f"PackParams::new with all {D} brute-force automorphism tables and 4d monomials, "
                      f"OfflinePackingKeys::generate; bit-exact vs GPU: {ok}/{nl}",
            "bit_exact_sampled": f"{ok}/{nl}",
        }
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000,
                    help="timed steps; the lanes start together, so a short region weighs their fill and drain "
                         "(20 steps measure about 15 %% below the 1000-step steady state)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--prime-ms", type=float, default=300.0,
                    help="untimed run of the workload on every lane before the warmup steps (clocks at load)")
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lanes", type=int, default=0,
                    help="lanes (inspire_ctx_create_lane) submitting batches from their own threads; 0 (default): "
                         "tuned, the count among --tune-lanes with the highest median rate over --tune-rounds "
                         "interleaved runs of --tune-steps untimed steps")
    ap.add_argument("--tune-lanes", type=str, default="1,2,3,4", help="lane counts the lane tuning tries")
    ap.add_argument("--tune-steps", type=int, default=0,
                    help="untimed steps per lane count and round when tuning lanes; 0 (default): as many as the "
                         "timed region has, at most 96 (a short timed region favours fewer lanes: filling and "
                         "draining the lanes' pipelines weighs more)")
    ap.add_argument("--tune-rounds", type=int, default=0,
                    help="rounds of the lane tuning (every candidate once per round, the median rate kept); "
                         "0 (default): enough for about 480 steps per candidate, 5 to 25")
    ap.add_argument("--tuning-file", type=str, default="",
                    help="where the tuning report (kernel variants and lanes) is written; with --reuse-tuning, "
                         "read from instead of measuring")
    ap.add_argument("--reuse-tuning", action="store_true")
    ap.add_argument("--chain-window-us", type=int, default=-1,
                    help="inspire_ctx_set_chain_window: chains of lanes submitted within this many microseconds "
                         "of each other run one after another; -1 (default): the library's default")
    ap.add_argument("--lanes-sweep", type=str, default="1,3",
                    help="other lane counts timed after the headline (reported as lanes_sweep)")
    ap.add_argument("--db-log2", type=int, default=26)
    ap.add_argument("--single-steps", type=int, default=32, help="steps of the one-query-per-launch run")
    ap.add_argument("--host-steps", type=int, default=3, help="steps of the host-buffer (PCIe-inclusive) run")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--tree-steps", type=int, default=10,
                    help="steps of the tree-packing (respond_seeded_packed) batch leg; 0 skips it")
    ap.add_argument("--extract-steps", type=int, default=20,
                    help="launches of the batched client-side extract leg (0 skips it)")
    ap.add_argument("--extract-batch", type=int, default=4096, help="responses per extract launch")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-literal-queries", type=int, default=-1,
                    help="queries of the literal CPU baseline (nothing hoisted), one per thread; -1 (default): "
                         "one per oracle thread, 0 skips it")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (a known-size kernel for "
                         "calibrating rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--serve", action="store_true",
                    help="measure the serving loop instead: open-loop Poisson arrivals over 64 databases")
    ap.add_argument("--serve-log2", type=int, default=28, help="total entries of the serving databases (log2)")
    ap.add_argument("--serve-dbs", type=int, default=64)
    ap.add_argument("--serve-rates", type=str, default="100,300,1000,3000,10000,30000,100000,200000,300000")
    ap.add_argument("--serve-lanes", type=int, default=2)
    ap.add_argument("--serve-max-batch", type=int, default=256)
    ap.add_argument("--serve-pool", type=int, default=256, help="distinct queries (with CPU answers) drawn from")
    ap.add_argument("--serve-leg-rates", type=str, default="10000,100000,400000",
                    help="offered rates of the serving leg of the default run (the last one above saturation); "
                         "empty skips the leg")
    ap.add_argument("--serve-leg-queries", type=str, default="10000,50000,50000",
                    help="queries per rate of the serving leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the respond path shards by database across GPUs with no collective, "
              f"and this build measures one GPU's share (asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    import inspire_mi355x as M

    log = lambda *a: print("[bench]", *a, file=sys.stderr, flush=True)
    w_seed = bytes(range(32))
    if args.serve:
        print(json.dumps(serve_main(args, M, log, w_seed)))
        return
    rng = np.random.default_rng(2026)
    n_entries = 1 << args.db_log2
    nshards = n_entries // D
    log(f"generating {n_entries} x 32 B database")
    db = np.frombuffer(rng.bytes(n_entries * 32), dtype=np.uint8)
    ctx = M.Context(M.InspireParams(), w_seed)
    if args.chain_window_us >= 0:
        ctx.set_chain_window(args.chain_window_us)
    # Kernel variants (matrix-core vs VALU forward NTT, packing-chain width):
    # timed on the device by the library's tuner, the faster kept.
    prior = None
    if args.reuse_tuning and args.tuning_file and os.path.exists(args.tuning_file):
        prior = json.load(open(args.tuning_file))
        ctx.load_tuning(args.tuning_file)
        tuning = dict(prior)
    else:
        tuning = ctx.autotune()
    log(f"tuning: ntt_fwd_2048 {tuning['ntt_fwd_2048']}, chain {tuning['chain_values_per_thread']} values/thread "
        f"{tuning.get('measured', {})}")
    # measured issue rates of the arithmetic itself (register-resident probes), the roof beside the
    # figure derived from the published FP32 rate
    RATES.update({k_: round(v, 3) for k_, v in ctx.measure_rates().items()})
    log(f"measured issue rates (T/s): {RATES}")
    t0 = time.perf_counter()
    ctx.load_entries(0, db, 32)
    log(f"database resident: {nshards} shards, {ctx.device_bytes / 2**30:.1f} GiB on device, "
        f"load {time.perf_counter() - t0:.1f} s")

    B = args.batch
    n_sets = 4  # distinct query batches cycled through the steps
    qsets = [synthetic_queries(rng, B, nshards) for _ in range(n_sets)]
    dev = [(shards, M.DeviceArray.from_numpy(b), M.DeviceArray.from_numpy(seeds), M.DeviceArray.from_numpy(y))
           for shards, b, seeds, y in qsets]
    dout = M.DeviceArray(B * 2 * K * D * 4)

    def step(i, count=B, c=ctx, out=dout):
        shards, db_, ds_, dy_ = dev[i % n_sets]
        st = c.respond_batch_device(0, M.TWO_PACKING, shards[:count], db_.ptr, out.ptr, seeds_ptr=ds_.ptr,
                                    y_body_ptr=dy_.ptr)
        if (st != 0).any():
            raise RuntimeError(f"status {st[st != 0][:4]}")

    if args.calibrate_copy_mib:
        nb = args.calibrate_copy_mib << 20
        src, dst = M.DeviceArray(nb), M.DeviceArray(nb)
        M.DeviceArray.hip().hipMemcpy(dst.ptr, src.ptr, nb, 3)
        M.DeviceArray.hip().hipDeviceSynchronize()
        src.free(), dst.free()
        log(f"calibration copy: {args.calibrate_copy_mib} MiB device to device")

    # Lanes: each has its own stream and workspaces over the resident state,
    # and submits batches from its own thread, so one batch's host work and
    # transfers overlap another's kernels.  A step is one batch on one lane.
    cand = sorted({int(x) for x in args.tune_lanes.split(",") if x.strip()} | ({args.lanes} if args.lanes else set()))
    lanes = [ctx] + [ctx.lane() for _ in range(max(cand + [1]) - 1)]
    louts = [dout] + [M.DeviceArray(B * 2 * K * D * 4) for _ in lanes[1:]]
    # The headline's regime from priming on: HIP events on each lane's stream
    # time the chain kernel of every batch (its roofline), and a recurring
    # batch replays the HIP graphs captured for it, the chain launched between
    # the graphs before and after it (its events are recorded by its own
    # dispatch).  Priming, lane tuning, warmup and the timed region all run
    # this way, so the graphs are captured before the timed region.
    def headline_regime(ln, on=True):
        ln.set_profiling_kernels(["pack_chain"])
        ln.set_profiling(on)
    for ln in lanes:
        headline_regime(ln)
    def run_lanes(lns, outs, steps):
        """steps batches taken in turn by the lanes' threads: (elapsed s, per-step latencies ms)"""
        lat = []
        counter = [0]
        lock = threading.Lock()
        errors = []
        go = threading.Barrier(len(lns) + 1)

        def lane_loop(j):
            go.wait()
            try:
                while True:
                    with lock:
                        i = counter[0]
                        counter[0] += 1
                    if i >= steps:
                        return
                    t0 = time.perf_counter()
                    step(i, c=lns[j], out=outs[j])
                    with lock:
                        lat.append((time.perf_counter() - t0) * 1e3)
            except Exception as e:  # surfaced after the timed region
                errors.append(e)

        threads = [threading.Thread(target=lane_loop, args=(j,)) for j in range(len(lns))]
        for t in threads:
            t.start()
        go.wait()
        t_start = time.perf_counter()
        for t in threads:
            t.join()
        el = time.perf_counter() - t_start
        if errors:
            raise errors[0]
        return el, lat

    # Priming (untimed): about --prime-ms of the workload on every lane, so that
    # the GPU's clocks are at their loaded state when the warmup steps run.
    t_prime = time.perf_counter()
    while (time.perf_counter() - t_prime) * 1e3 < args.prime_ms:
        run_lanes(lanes, louts, 2 * len(lanes))
    # Lanes: the headline's lane count is tuned (untimed) unless given.
    lane_rates = {}
    t_steps = t_rounds = 0
    if args.lanes:
        n_lanes = args.lanes
    elif prior and "lanes" in prior:
        n_lanes = int(prior["lanes"])
    else:
        # every candidate once per round, rounds interleaved, so a drift of the clocks or a
        # slow run moves all candidates alike; the median of each candidate's rates is kept
        t_steps = args.tune_steps if args.tune_steps > 0 else max(1, min(args.steps, 96))
        t_rounds = args.tune_rounds if args.tune_rounds > 0 else min(25, max(5, 480 // t_steps))
        runs = {nl: [] for nl in cand}
        for _ in range(t_rounds):
            for nl in cand:
                el, _ = run_lanes(lanes[:nl], louts[:nl], t_steps)
                runs[nl].append(B * t_steps / el)
        for nl in cand:
            lane_rates[nl] = round(float(np.median(runs[nl])), 1)
        # the fewest lanes within 3 % of the best rate: more lanes only add latency beyond that
        best = max(lane_rates.values())
        n_lanes = min(nl for nl, r in lane_rates.items() if r >= 0.97 * best)
    tuning["lanes"] = n_lanes
    if lane_rates:
        tuning.setdefault("measured", {})["lanes_qps"] = {str(k): v for k, v in lane_rates.items()}
        tuning["measured"]["lanes_tuning"] = f"median of {t_rounds} interleaved rounds of {t_steps} steps"
    log(f"lanes: {n_lanes} {lane_rates}")
    if args.tuning_file and not args.reuse_tuning:
        with open(args.tuning_file, "w") as f:
            json.dump(tuning, f)
    spare = lanes[n_lanes:]
    lanes, louts, spare_outs = lanes[:n_lanes], louts[:n_lanes], louts[n_lanes:]
    for j, ln in enumerate(lanes):
        for i in range(args.warmup):
            step(i + j, c=ln, out=louts[j])
    # Timed region: HIP events bracket only the kernel the roofline is computed
    # for (every bracket costs host time inside the step).
    kernels = ("pack_chain", "pack_front", "first_dim", "pack_step1", "pack_b0", "expand_seeds", "rows_ntt", "y_ntt")
    for ln in lanes:
        for name in kernels:
            ln.profile(name, reset=True)
    caps0 = [ln.graph_stats() for ln in lanes]
    elapsed, lat = run_lanes(lanes, louts, args.steps)
    caps1 = [ln.graph_stats() for ln in lanes]
    headline_graphs = {"captures": int(sum(b[0] - a[0] for a, b in zip(caps0, caps1))),
                       "replays": int(sum(b[1] - a[1] for a, b in zip(caps0, caps1)))}
    n_lanes = len(lanes)
    prof = {}
    for name in kernels:
        tot = [ln.profile(name, reset=True) for ln in lanes]
        prof[name] = (sum(x[0] for x in tot), sum(x[1] for x in tot))
    for ln in lanes:
        ln.set_profiling(False)
    qps = B * args.steps / elapsed
    log(f"batched: {qps:.0f} queries/s, {elapsed / args.steps * 1e3:.3f} ms/step over {n_lanes} lanes, "
        f"p50 {percentile(lat, 50):.3f} ms, graphs {headline_graphs}")
    # The same timed region without profiling events, with graph replay of the
    # repeated batches on and then off.
    graphs_ab = {}
    for on in (True, False):
        for j, ln in enumerate(lanes):
            ln.set_profiling(False)
            ln.set_graphs(on)
            for i in range(max(args.warmup, 2 * n_sets)):  # every query set seen twice: captured
                step(i + j, c=ln, out=louts[j])
        caps0 = [ln.graph_stats() for ln in lanes]
        el, lt = run_lanes(lanes, louts, args.steps)
        caps1 = [ln.graph_stats() for ln in lanes]
        graphs_ab["graphs_on" if on else "graphs_off"] = {
            "value": round(B * args.steps / el, 1), "ms_per_step": round(el / args.steps * 1e3, 3),
            "p50_ms": round(percentile(lt, 50), 3), "p99_ms": round(percentile(lt, 99), 3),
            "graph_replays": int(sum(b[1] - a[1] for a, b in zip(caps0, caps1)))}
        log(f"  no profiling, graphs {'on' if on else 'off'}: {graphs_ab['graphs_on' if on else 'graphs_off']}")
    # the headline region once more (same K steps) after the others, with the
    # front end's launches bracketed too: the spread between two measurements
    # of the same thing in one process, and the front end's launch time there
    for j, ln in enumerate(lanes):
        ln.set_graphs(True)
        ln.set_profiling_kernels(["pack_chain", "pack_front"])
        ln.set_profiling(True)
        for i in range(max(args.warmup, 2 * n_sets)):
            step(i + j, c=ln, out=louts[j])
        for name in kernels:
            ln.profile(name, reset=True)
    el, lt = run_lanes(lanes, louts, args.steps)
    prof_repeat = {}
    for name in kernels:
        tot = [ln.profile(name, reset=True) for ln in lanes]
        prof_repeat[name] = (sum(x[0] for x in tot), sum(x[1] for x in tot))
    for ln in lanes:
        ln.set_profiling(False)
    graphs_ab["profiled_repeat"] = {
        "value": round(B * args.steps / el, 1), "ms_per_step": round(el / args.steps * 1e3, 3),
        "p50_ms": round(percentile(lt, 50), 3), "p99_ms": round(percentile(lt, 99), 3),
        "pack_chain_avg_ms": round(prof_repeat["pack_chain"][0] / max(prof_repeat["pack_chain"][1], 1), 4),
        "pack_front_avg_ms": round(prof_repeat["pack_front"][0] / max(prof_repeat["pack_front"][1], 1), 4)}
    log(f"  profiled repeat (chain and front end bracketed): {graphs_ab['profiled_repeat']}")
    # the same batches from other numbers of lanes (more lanes: more throughput,
    # longer latency), in the headline's regime
    sweep = []
    lanes, louts = lanes + spare, louts + spare_outs
    for j, ln in enumerate(lanes):
        headline_regime(ln)
        for i in range(max(args.warmup, 2 * n_sets)):
            step(i + j, c=ln, out=louts[j])
    for nl in sorted({int(x) for x in args.lanes_sweep.split(",") if x.strip()} | set(cand)):
        if nl < 1 or nl == n_lanes:
            continue
        while len(lanes) < nl:
            lanes.append(ctx.lane())
            louts.append(M.DeviceArray(B * 2 * K * D * 4))
            headline_regime(lanes[-1])
            for i in range(max(args.warmup, 2 * n_sets)):
                step(i + len(lanes) - 1, c=lanes[-1], out=louts[-1])
        el, lt = run_lanes(lanes[:nl], louts[:nl], args.steps)
        sweep.append({"lanes": nl, "value": round(B * args.steps / el, 1), "ms_per_step": round(el / args.steps * 1e3, 3),
                      "p50_ms": round(percentile(lt, 50), 3), "p99_ms": round(percentile(lt, 99), 3)})
        log(f"  {nl} lanes: {sweep[-1]['value']:.0f} queries/s, p50 {sweep[-1]['p50_ms']} ms, "
            f"p99 {sweep[-1]['p99_ms']} ms")
    # untimed: every kernel of the step alone on the GPU (one lane, eager
    # launches whose events are recorded by the kernels' own dispatches), for
    # the breakdown
    for ln in lanes:
        ln.set_profiling(False)
    ctx.set_graphs(False)
    ctx.set_profiling_kernels(None)
    for name in kernels:
        ctx.profile(name, reset=True)
    ctx.set_profiling(True)
    for i in range(min(args.steps, 5)):
        step(i)
    breakdown = {name: ctx.profile(name, reset=True) for name in kernels}
    ctx.set_profiling(False)
    ctx.set_graphs(True)
    for k_, (ms, n) in breakdown.items():
        if n:
            log(f"  {k_:13s} {ms / n:8.3f} ms/launch x {n}")

    # the same path, one query per launch
    single, single_qps = [], None
    if args.single_steps:
        for i in range(args.warmup):
            step(i, 1)
        t0 = time.perf_counter()
        for i in range(args.single_steps):
            t1 = time.perf_counter()
            step(i, 1)
            single.append((time.perf_counter() - t1) * 1e3)
        single_qps = args.single_steps / (time.perf_counter() - t0)
        log(f"one query per launch: {single_qps:.0f} queries/s, p50 {percentile(single, 50):.2f} ms")

    # the same batches through the host-buffer API (queries in the reference's
    # u64 layout in host memory, responses copied back): PCIe-inclusive rate
    host_qps = None
    if args.host_steps:
        shards, b, seeds, y = qsets[0]
        hq = []
        for i in range(B):
            rg = np.zeros((2 * ELL, K, D), dtype=np.uint64)
            rg[ELL:] = b[i]
            sd = np.zeros((2 * ELL, 32), dtype=np.uint8)
            sd[ELL:] = seeds[i].astype("<u4").view(np.uint8).reshape(ELL, 32)
            hq.append(M.ClientQuery(shard_id=int(shards[i]), rgsw=rg, seeds=sd, has_packing_keys=True,
                                    y_body=y[i].astype(np.uint64)))
        ids, vs = [0] * B, [M.TWO_PACKING] * B
        ctx.respond_batch(ids, hq, vs)
        t0 = time.perf_counter()
        for _ in range(args.host_steps):
            res = ctx.respond_batch(ids, hq, vs)
        host_qps = B * args.host_steps / (time.perf_counter() - t0)
        assert all(not isinstance(r, M.PirError) for r in res)
        log(f"host-buffer API (PCIe-inclusive): {host_qps:.0f} queries/s")

    # The chain's roofline uses the timed region's own events (its launches
    # there may share the GPU with the other lane's kernels).  The front end
    # runs mostly beside another lane's chain in the timed region, so its HBM
    # roofline uses the untimed single-lane pass, where it runs alone.
    chain_ms, chain_n = prof["pack_chain"]
    chain_avg_s = chain_ms / max(chain_n, 1) / 1e3
    front = "pack_front" if breakdown["pack_front"][1] else "first_dim"
    fd_ms, fd_n = breakdown[front]
    fd_avg_s = fd_ms / max(fd_n, 1) / 1e3
    iso_chain_ms = breakdown["pack_chain"][0] / max(breakdown["pack_chain"][1], 1)
    chain_tmodmul = MODMUL_CHAIN * B / chain_avg_s / 1e12
    roofline = {
        "bound": "valu",
        "kernel": "k_pack_chain_crt",
        "achieved": round(chain_tmodmul, 3),
        "peak": round(MODMUL_PEAK_T, 2),
        "unit": "Tmodmul/s",
        "frac": round(chain_tmodmul / MODMUL_PEAK_T, 4),
        "traffic": measured_traffic("k_pack_chain_crt", B),
        "algorithmic_per_launch": f"{MODMUL_CHAIN} modmul/query x {B} queries",
        "avg_launch_ms": round(chain_avg_s * 1e3, 4),
        "measured": f"HIP events on each lane's stream over the timed region, recorded by the chain's own "
                    f"dispatch between the batch's graph launches ({n_lanes} lanes, sharing CUs with the other "
                    "lanes' kernels)",
        "timed_region_graphs": headline_graphs,
        "avg_launch_ms_alone": round(iso_chain_ms, 4),
        # the same algorithmic work over the launch time of the kernel alone on the GPU (one lane)
        "frac_alone": (round(MODMUL_CHAIN * B / (iso_chain_ms * 1e-3) / 1e12 / MODMUL_PEAK_T, 4)
                       if iso_chain_ms else None),
        "peak_measured": RATES.get("smmul_tops"),
        "frac_measured": round(chain_tmodmul / RATES["smmul_tops"], 4) if RATES.get("smmul_tops") else None,
        "peak_measured_note": "the signed Montgomery product's measured issue rate (the chain's digit transforms "
                              "and multiply-accumulates are signed)",
    }
    roofline_hbm = {
        "bound": "hbm",
        "kernel": "k_pack_front" if front == "pack_front" else "k_first_dim",
        "achieved": round(FRONT_BYTES * B / fd_avg_s / 1e9, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(FRONT_BYTES * B / fd_avg_s / 1e9 / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic("k_pack_front", B),
        "algorithmic_per_launch": f"{FRONT_BYTES} B/query x {B} queries",
        "avg_launch_ms": round(fd_avg_s * 1e3, 4),
        "measured": "HIP events, untimed single-lane pass (the kernel alone on the GPU)",
        # its launches in the profiled repeat of the timed region (beside other lanes' kernels)
        "avg_launch_ms_timed_region": round(prof_repeat[front][0] / max(prof_repeat[front][1], 1), 4),
        # the device's measured streaming-read rate (16-byte loads over 1 GiB), the practical roof
        "peak_measured": round(RATES["hbm_read_tbs"] * 1e3, 1) if RATES.get("hbm_read_tbs") else None,
        "frac_measured": (round(FRONT_BYTES * B / fd_avg_s / 1e9 / (RATES["hbm_read_tbs"] * 1e3), 4)
                          if RATES.get("hbm_read_tbs") else None),
    }
    line = {
        "metric": METRIC,
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "InsPIRe^2 batched respond: 256 seeded queries + y_body keys per step, "
                        "d=2048, CRT [268369921, 249561089], 2^26 x 32 B DB resident in HBM",
            "batch": B,
            "lanes": n_lanes,
            "db_entries": n_entries,
            "entry_bytes": 32,
            "ring_dim": D,
            "gadget": "2^20 x 3",
            "chain_window_us": ctx.chain_window(),
        },
        "p50_ms": round(percentile(lat, 50), 3),
        "p99_ms": round(percentile(lat, 99), 3),
        "single_query_qps": round(single_qps, 1) if single_qps else None,
        "single_query_p50_ms": round(percentile(single, 50), 3) if single else None,
        "speedup_vs_single_query_launch": round(qps / single_qps, 2) if single_qps else None,
        "host_buffer_api_qps": round(host_qps, 1) if host_qps else None,
        "lanes_sweep": sweep,
        "graphs_ab": graphs_ab,
        "tuning": tuning,
        "measured_issue_rates_tops": dict(RATES),
        "modmul_per_query": MODMUL_QUERY,
        "modmul_per_query_reference_algorithm": MODMUL_QUERY_REF_ALGORITHM,
        "modmul_note": "modmul_per_query counts the work this path does (database digits stored in the NTT "
                       "domain, step 1 on residue 0 only, the rows' a expanded and transformed for residue 0 "
                       "only, the last gadget digit's transform recovered); the reference's algorithm as written does "
                       f"{MODMUL_QUERY_REF_ALGORITHM} per query; the roofline uses the former",
        "roofline": roofline,
        "roofline_hbm_front": roofline_hbm,
        "kernel_ms_per_step": {k_: round(ms / n, 4) for k_, (ms, n) in breakdown.items() if n},
        "cpu_baseline": None,
    }
    if args.tree_steps:
        line["tree_packing"] = tree_leg(args, M, ctx, db, dev, qsets, log)
    if args.extract_steps:
        line["extract"] = extract_leg(args, M, ctx, log)
    if not args.no_cpu:
        shards, b, seeds, y = qsets[(args.steps - 1) % n_sets]
        step(args.steps - 1)  # the batch of this query set, responses in pout
        gout = dout.download(np.zeros((B, 2, K, D), dtype=np.uint32))
        log("cpu baseline (oracle) ...")
        line["cpu_baseline"] = cpu_baseline(db, (shards, b, seeds, y), gout, args.cpu_seconds, w_seed,
                                            args.cpu_literal_queries)
    for _, db_, ds_, dy_ in dev:
        db_.free(), ds_.free(), dy_.free()
    for o_ in louts:
        o_.free()
    for ln in lanes[1:]:
        ln.close()
    ctx.close()
    if args.serve_leg_rates:
        # the serving configuration: continuous batching over 64 uneven databases (2^28 x 32 B
        # resident), Poisson arrivals, the ^0/^1/^2 mix; every response checked against the oracle
        rates = [float(x) for x in args.serve_leg_rates.split(",") if x]
        counts = [int(x) for x in args.serve_leg_queries.split(",") if x]
        sv = serve_main(args, M, log, w_seed, rates=rates, counts=counts)
        line["serving"] = {k: sv[k] for k in ("value", "unit", "config", "rates", "bit_exact", "server",
                                              "cpu_baseline")}
        line["serving"]["metric"] = "served queries/s at saturation (highest achieved rate), p50/p99 per offered rate"
    print(json.dumps(line))


def tree_leg(args, M, ctx, db, dev, qsets, log):
    """Tree packing (respond_seeded_packed, src/pir/respond.rs:428-435) of the
    same seeded query batches over the same resident database, with
    synthetic CRS galois keys: queries/s, per-kernel times, the leaves
    kernel's modmul roofline, and a sample checked bit for bit against the
    oracle (which is also the CPU baseline of this leg)."""
    B = args.batch
    rng = np.random.default_rng(2027)
    moduli = np.array([Q0, Q1], dtype=np.uint64)[:, None]
    L = D.bit_length() - 1
    gk = rng.integers(0, 2**63, (L, ELL, 2, K, D), dtype=np.uint64) % moduli
    ctx.set_galois_keys(gk)
    words = K * D + K * 2 * D
    tout = M.DeviceArray(B * words * 4)

    def tstep(i):
        shards, db_, ds_, _ = dev[i % len(dev)]
        st = ctx.respond_batch_device(0, M.TWO_PACKING, shards, db_.ptr, tout.ptr, seeds_ptr=ds_.ptr,
                                      packing_mode=M.PACKING_TREE)
        if (st != 0).any():
            raise RuntimeError(f"status {st[st != 0][:4]}")

    for i in range(args.warmup):
        tstep(i)
    names = ("tree_leaves", "tree_combine", "expand_seeds")
    ctx.set_profiling_kernels(["tree_leaves", "tree_combine"])
    ctx.set_profiling(True)
    for nm in names:
        ctx.profile(nm, reset=True)
    lat = []
    t_start = time.perf_counter()
    for i in range(args.tree_steps):
        t0 = time.perf_counter()
        tstep(i)
        lat.append((time.perf_counter() - t0) * 1e3)
    elapsed = time.perf_counter() - t_start
    prof = {nm: ctx.profile(nm, reset=True) for nm in names}
    ctx.set_profiling(False)
    ctx.set_profiling_kernels(None)
    qps = B * args.tree_steps / elapsed
    mm_leaves, mm_combine = tree_modmul()
    lv_ms, lv_n = prof["tree_leaves"]
    cb_ms, cb_n = prof["tree_combine"]
    lv_avg = lv_ms / max(lv_n, 1) / 1e3
    achieved = mm_leaves * B / lv_avg / 1e12
    log(f"tree packing: {qps:.0f} queries/s, {elapsed / args.tree_steps * 1e3:.2f} ms/step, "
        f"leaves {lv_avg * 1e3:.3f} ms ({achieved:.2f} Tmodmul/s), combine {cb_ms / max(lv_n, 1):.3f} ms/step")
    out = {
        "workload": f"respond_seeded_packed (tree packing, 16 LWEs, 11 levels): {B} seeded queries per step, "
                    "same 2^26 x 32 B database, synthetic CRS galois keys",
        "value": round(qps, 1),
        "unit": "queries/s",
        "steps": args.tree_steps,
        "ms_per_step": round(elapsed / args.tree_steps * 1e3, 3),
        "p50_ms": round(percentile(lat, 50), 3),
        "p99_ms": round(percentile(lat, 99), 3),
        "modmul_per_query": mm_leaves + mm_combine + 2 * ELL * K * NTT,
        "kernel_ms_per_step": {"tree_leaves": round(lv_avg * 1e3, 4),
                               "tree_combine": round(cb_ms / max(lv_n, 1), 4)},
        "roofline": {
            "bound": "valu",
            "kernel": "k_tree_leaves",
            "achieved": round(achieved, 3),
            "peak": round(MODMUL_PEAK_T, 2),
            "unit": "Tmodmul/s",
            "frac": round(achieved / MODMUL_PEAK_T, 4),
            "peak_measured": RATES.get("mmul_tops"),
            "frac_measured": round(achieved / RATES["mmul_tops"], 4) if RATES.get("mmul_tops") else None,
            "traffic": measured_traffic("k_tree_leaves", B),
            "algorithmic_per_launch": f"{mm_leaves} modmul/query x {B} queries",
            "avg_launch_ms": round(lv_avg * 1e3, 4),
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        sys.path.insert(0, ROOT)
        import oracle as O

        o = O.Oracle(O.baseline_params())
        i_last = args.tree_steps - 1
        tstep(i_last)
        gout = tout.download(np.zeros((B, words), dtype=np.uint32))
        shards, b, seeds, _ = qsets[i_last % len(qsets)]
        n = min(B, O.lib.or_num_threads())
        polys, rgsws = [], []
        for i in range(n):
            s = int(shards[i])
            polys.append(o.encode_shard(db[s * D * 32:(s + 1) * D * 32], 32, 0))
            rg = np.zeros((2 * ELL, 2, K, D), dtype=np.uint64)
            for r in range(ELL):
                rg[ELL + r, 0] = o.poly_from_seed(seeds[i, r].astype("<u4").tobytes())
                rg[ELL + r, 1] = b[i, r]
            rgsws.append(rg)
        t0 = time.perf_counter()
        outs = o.respond_one_packing_many(polys, rgsws, gk)
        t_cpu = time.perf_counter() - t0
        ok = sum(np.array_equal(outs[i].astype(np.uint32), gout[i]) for i in range(n))
        out["cpu_baseline"] = {
            "value": n / t_cpu,
            "unit": "queries/s",
            "cores": O.lib.or_num_threads(),
            "kind": "port",
            "sample": f"{n} tree-packed queries of the timed workload, one per thread, {t_cpu:.1f} s wall; "
                      f"bit-exact vs GPU: {ok}/{n}",
            "bit_exact_sampled": f"{ok}/{n}",
        }
        log(f"tree cpu baseline: {n / t_cpu:.1f} queries/s, bit-exact {ok}/{n}")
    tout.free()
    return out


def extract_leg(args, M, ctx, log):
    """Batched client-side extract (extract_inspiring, src/pir/extract.rs:
    163-192) of InspiRING responses of the headline shape ([2][K][d] u32, 32 B
    entries: 16 coefficients) with one secret key per response, all resident
    in HBM: responses/s, the kernel's modmul and HBM rooflines, and a sample
    checked bit for bit against the oracle (also the CPU baseline)."""
    E = args.extract_batch
    es, nvals = 32, 16
    rng = np.random.default_rng(2028)
    moduli = np.array([Q0, Q1], dtype=np.uint64)[:, None]
    resp = (rng.integers(0, 2**63, (E, 2, K, D), dtype=np.uint64) % moduli).astype(np.uint32)
    # RLWE secret keys: small signed coefficients as residues
    e = rng.integers(-8, 9, (E, 1, D))
    sk = np.where(e >= 0, e, np.array([Q0, Q1], dtype=np.int64)[None, :, None] + e).astype(np.uint32)
    dresp, dsk, dent = M.DeviceArray.from_numpy(resp), M.DeviceArray.from_numpy(sk), M.DeviceArray(E * es)

    def run():
        ctx.extract_batch_device(E, M.EXTRACT_INSPIRING, es, dresp.ptr, dsk.ptr, dent.ptr)

    for _ in range(args.warmup):
        run()
    ctx.set_profiling_kernels(["extract"])
    ctx.set_profiling(True)
    ctx.profile("extract", reset=True)
    t0 = time.perf_counter()
    for _ in range(args.extract_steps):
        run()
    elapsed = time.perf_counter() - t0
    ms, n = ctx.profile("extract", reset=True)
    ctx.set_profiling(False)
    ctx.set_profiling_kernels(None)
    avg = ms / max(n, 1) / 1e3
    rps = E * args.extract_steps / elapsed
    mm = nvals * K * D  # one product per term of each coefficient's dot product
    nbytes = K * D * 4 + K * D * 4 + nvals * K * 4 + es  # a, the key, the b words read, the entry
    mm_t = mm * E / avg / 1e12
    gbs = nbytes * E / avg / 1e9
    log(f"extract: {rps:.0f} responses/s, kernel {avg * 1e3:.4f} ms per {E} ({mm_t:.2f} Tmodmul/s, {gbs:.0f} GB/s)")
    out = {
        "workload": f"extract_inspiring of {E} device-resident InspiRING responses per launch (d=2048, CRT, "
                    "32 B entries), one secret key each",
        "value": round(rps, 1),
        "unit": "responses/s",
        "steps": args.extract_steps,
        "avg_launch_ms": round(avg * 1e3, 4),
        "roofline": {
            "bound": "valu" if mm_t / (RATES.get("mad64_tops") or MODMUL_PEAK_T) >= gbs / HBM_PEAK_GBS else "hbm",
            "kernel": "k_extract",
            "achieved_modmul": round(mm_t, 3), "peak_modmul": round(MODMUL_PEAK_T, 2), "unit_modmul": "Tmodmul/s",
            "frac_modmul": round(mm_t / MODMUL_PEAK_T, 4),
            "peak_mad64_measured": RATES.get("mad64_tops"), "unit_mad64": "T multiply-add/s",
            "frac_mad64_measured": round(mm_t / RATES["mad64_tops"], 4) if RATES.get("mad64_tops") else None,
            "mad64_note": "each term of the kernel's dot products is one 64-bit multiply-add, not a modular "
                          "product: frac_mad64_measured is its roof; frac_modmul prices it as modmuls",
            "achieved_hbm": round(gbs, 1), "peak_hbm": HBM_PEAK_GBS, "unit_hbm": "GB/s",
            "frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
            "frac_hbm_measured": round(gbs / (RATES["hbm_read_tbs"] * 1e3), 4) if RATES.get("hbm_read_tbs") else None,
            "traffic": measured_traffic("k_extract", E),
            "algorithmic_per_launch": f"{mm} modmul and {nbytes} B per response x {E} responses",
        },
        "cpu_baseline": None,
    }
    got = dent.download(np.zeros((E, es), dtype=np.uint8))
    if not args.no_cpu:
        sys.path.insert(0, ROOT)
        import oracle as O

        o = O.Oracle(O.baseline_params())
        n_cpu, t_cpu, ok = 0, 0.0, 0
        t0 = time.perf_counter()
        while n_cpu < E and (time.perf_counter() - t0) < max(args.cpu_seconds / 4, 1.0):
            want = o.extract_inspiring(resp[n_cpu].astype(np.uint64), sk[n_cpu].astype(np.uint64), es)
            ok += want == got[n_cpu].tobytes()
            n_cpu += 1
        t_cpu = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": n_cpu / t_cpu,
            "unit": "responses/s",
            "cores": 1,
            "kind": "port",
            "sample": f"the first {n_cpu} responses of the launch, one thread, {t_cpu:.1f} s; "
                      f"bit-exact vs GPU: {ok}/{n_cpu}",
            "bit_exact_sampled": f"{ok}/{n_cpu}",
        }
        log(f"extract cpu baseline: {n_cpu / t_cpu:.1f} responses/s, bit-exact {ok}/{n_cpu}")
    dresp.free(), dsk.free(), dent.free()
    return out


SERVE_METRIC = ("serving loop: queries/s + p50/p99 respond latency, InsPIRe^0/^1/^2 at 10/30/60 %, "
                "open-loop Poisson arrivals, 2^28 x 32 B in 64 databases")


def serve_sizes(rng, total_log2: int, ndb: int):
    """Log-uniform sizes in [2^20, 2^26] entries, scaled to a total of
    2^total_log2, in whole shards of D entries."""
    s = 2.0 ** rng.uniform(20, 26, ndb)
    target = (1 << total_log2) // D
    shards = np.maximum(1, np.round(s * target / s.sum())).astype(np.int64)
    shards[np.argmax(shards)] += target - shards.sum()
    return shards * D


def serve_world(M, rng, total_log2: int, ndb: int, w_seed: bytes, log, n_clients: int = 16, pool_size: int = 256):
    """The serving configuration: ndb databases of log-uniform size totalling
    2^total_log2 x 32 B of uniform bytes, resident; n_clients clients with
    registered packing keys; a pool of synthetic queries of the real shape
    mixing InsPIRe^0/^1/^2 at 10/30/60 % over uniform databases and shards."""
    sizes = serve_sizes(rng, total_log2, ndb)
    total = int(sizes.sum())
    log(f"serving databases: {len(sizes)} of {sizes.min()}..{sizes.max()} entries, total {total}")
    raw = rng.integers(0, 2**64 - 1, total * 32 // 8, dtype=np.uint64, endpoint=True).view(np.uint8)
    offs = np.concatenate([[0], np.cumsum(sizes)]) * 32
    ctx = M.Context(M.InspireParams(), w_seed)
    t0 = time.perf_counter()
    for i in range(len(sizes)):
        ctx.load_entries(i, raw[offs[i]:offs[i + 1]], 32)
    log(f"resident: {ctx.device_bytes / 2**30:.1f} GiB, load {time.perf_counter() - t0:.1f} s")
    moduli = np.array([Q0, Q1], dtype=np.uint64)[:, None]
    ybodies = [(rng.integers(0, 2**63, (ELL, K, D), dtype=np.uint64) % moduli) for _ in range(n_clients)]
    for c, yb in enumerate(ybodies):
        ctx.register_client_keys(c, yb)
    pool = []
    for i in range(pool_size):
        u = rng.random()
        variant = 0 if u < 0.1 else (1 if u < 0.4 else 2)
        db = int(rng.integers(0, len(sizes)))
        shard = int(rng.integers(0, sizes[db] // D))
        client = int(rng.integers(0, n_clients))
        if variant == 1:  # full rows ell..2ell-1 (a, b)
            rows = (rng.integers(0, 2**63, (ELL, 2, K, D), dtype=np.uint64) % moduli).astype(np.uint32)
            seeds = None
        else:  # seeded
            rows = (rng.integers(0, 2**63, (ELL, K, D), dtype=np.uint64) % moduli).astype(np.uint32)
            seeds = rng.integers(0, 256, (ELL, 32), dtype=np.uint8)
        pool.append(dict(db=db, shard=shard, variant=variant, client=client, rows=rows, seeds=seeds))
    return dict(ctx=ctx, sizes=sizes, raw=raw, offs=offs, ybodies=ybodies, pool=pool, n_clients=n_clients)


def request_structs(M, pool):
    structs, keep = [], []
    for q in pool:
        st, k = M.Server.request_struct(M.Request(q["db"], q["shard"], q["variant"], rows=q["rows"],
                                                  seeds=q["seeds"], client_id=q["client"]), borrow=True)
        structs.append(st)
        keep.append(k)
    return structs, keep


def poisson_run(M, srv, pool, structs, answers, lam: float, n: int, rng):
    """n queries drawn from the pool, arriving open-loop at Poisson rate lam
    (arrival times fixed in advance, submitted when due); every response is
    compared with the pool query's answer, by waiter threads that drain the
    completions in chunks (inspire_server_wait_many).  -> (result dict,
    mismatches)."""
    import threading

    pick = rng.integers(0, len(pool), n)
    gaps = rng.exponential(1.0 / lam, n)
    t_start = time.monotonic_ns() + 20_000_000
    arrival = (t_start + np.cumsum(gaps) * 1e9).astype(np.int64)
    tickets = np.zeros(n, dtype=np.uint64)
    done = np.zeros(n, dtype=np.int64)
    bad = [0]
    submitted = [0]

    ans_ct = np.stack([answers[i][0] for i in range(len(pool))])  # [P, 2, K, D] u64
    chunk, n_waiters = 64, 4

    def waiter(wi):
        # drains completions chunk by chunk (chunk c on waiter c % n_waiters) and checks every response
        cts = np.empty((chunk,) + ans_ct.shape[1:], dtype=np.uint64)
        colbufs = [np.empty((16, 2, K, D), dtype=np.uint64) for _ in range(chunk)]
        for c0 in range(wi * chunk, n, n_waiters * chunk):
            c = min(chunk, n - c0)
            with progress:
                progress.wait_for(lambda: submitted[0] >= c0 + c)
            pk = pick[c0:c0 + c]
            zero = [i for i in range(c) if pool[pk[i]]["variant"] == 0]
            st, dn, _ = srv.wait_many(tickets[c0:c0 + c], cts[:c], {i: colbufs[i] for i in zero})
            done[c0:c0 + c] = dn.astype(np.int64)
            ok = (st == M.OK) & (cts[:c].reshape(c, -1) == ans_ct[pk].reshape(c, -1)).all(axis=1)
            for i in zero:
                ok[i] &= bool(np.array_equal(colbufs[i], answers[pk[i]][1]))
            with lock:
                bad[0] += int(c - ok.sum())

    lock = threading.Lock()
    progress = threading.Condition()  # the submitter's count of submitted queries
    ths = [threading.Thread(target=waiter, args=(wi,)) for wi in range(n_waiters)]
    for th in ths:
        th.start()
    i = 0
    while i < n:
        now = time.monotonic_ns()
        k = int(np.searchsorted(arrival, now, side="right"))
        if k > i:
            arr = (M.RequestStruct * (k - i))()
            for j in range(i, k):
                arr[j - i] = structs[pick[j]]
                arr[j - i].arrival_ns = int(arrival[j])
            tickets[i:k] = srv.submit_structs(arr, k - i)
            with progress:
                crossed = k // chunk != i // chunk or k == n  # a chunk became complete
                submitted[0] = k
                if crossed:
                    progress.notify_all()
            i = k
        else:
            dt = (arrival[i] - now) / 1e9
            if dt > 300e-6:
                time.sleep(dt - 200e-6)
    for th in ths:
        th.join()
    lat = (done - arrival) / 1e6
    span = (done.max() - arrival[0]) / 1e9
    return {
        "offered_qps": lam,
        "queries": n,
        "achieved_qps": round(n / span, 1),
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "bit_exact": f"{n - bad[0]}/{n}",
    }, bad[0]


def serve_main(args, M, log, w_seed: bytes, rates=None, counts=None):
    """Open-loop Poisson arrivals into the serving loop (inspire_server_*)
    over databases of uneven size: throughput and p50/p99 latency per offered
    rate.  Queries are drawn from a pool of synthetic queries of the real
    shape; the CPU leg computes the oracle's answer for each pool query (and
    times it), and every served response is compared with that answer."""
    rng = np.random.default_rng(2028)
    W = serve_world(M, rng, args.serve_log2, args.serve_dbs, w_seed, log, pool_size=args.serve_pool)
    ctx, sizes, pool = W["ctx"], W["sizes"], W["pool"]
    cpu = serve_cpu_answers(pool, W["raw"], W["offs"], W["ybodies"], w_seed, log)
    structs, keep = request_structs(M, pool)
    results = []
    if rates is None:
        rates = [float(x) for x in args.serve_rates.split(",") if x]
    mismatches_total, checked_total = 0, 0
    with M.Server(ctx, lanes=args.serve_lanes, max_batch=args.serve_max_batch) as srv:
        for j, lam in enumerate(rates):
            n = counts[j] if counts else int(min(50000, max(200, lam * 1.0)))
            res, bad = poisson_run(M, srv, pool, structs, cpu, lam, n, rng)
            mismatches_total += bad
            checked_total += n
            results.append(res)
            log(f"offered {lam:>8.0f} q/s: achieved {res['achieved_qps']:>9.0f} q/s, p50 {res['p50_ms']} ms, "
                f"p99 {res['p99_ms']} ms, bit-exact {res['bit_exact']}")
        stats = srv.stats()
    ctx.close()
    peak = max(results, key=lambda r: r["achieved_qps"])
    total = int(sizes.sum())
    return {
        "metric": SERVE_METRIC,
        "value": peak["achieved_qps"],
        "unit": "queries/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "serving loop: open-loop Poisson arrivals, InsPIRe^0/^1/^2 at 10/30/60 %, "
                        f"{len(sizes)} databases of log-uniform size totalling 2^{args.serve_log2} x 32 B, "
                        f"{W['n_clients']} clients with resident packing keys",
            "db_entries_min": int(sizes.min()),
            "db_entries_max": int(sizes.max()),
            "db_entries_total": total,
            "lanes": args.serve_lanes,
            "max_batch": args.serve_max_batch,
        },
        "rates": results,
        "bit_exact": f"{checked_total - mismatches_total}/{checked_total}",
        "server": stats,
        "cpu_baseline": cpu["baseline"],
    }


def serve_cpu_answers(pool, raw, offs, ybodies, w_seed: bytes, log):
    """The oracle (C restatement of the reference respond) on this host's
    cores for every pool query: its answers check the served responses, and
    its time is the CPU baseline of the mix."""
    sys.path.insert(0, ROOT)
    import oracle as O

    o = O.Oracle(O.baseline_params())
    o.pack(GAMMA, w_seed)
    threads = O.lib.or_num_threads()

    def full_rows(q):
        rg = np.zeros((2 * ELL, 2, K, D), dtype=np.uint64)
        for r in range(ELL):
            if q["variant"] == 1:
                rg[ELL + r] = q["rows"][r]
            else:
                rg[ELL + r, 0] = o.poly_from_seed(q["seeds"][r].tobytes())
                rg[ELL + r, 1] = q["rows"][r]
        return rg

    def shard_polys(q):
        base = offs[q["db"]] + q["shard"] * D * 32
        return o.encode_shard(raw[base:base + D * 32], 32, 0)

    answers = {}
    packed = [i for i, q in enumerate(pool) if q["variant"] != 0]
    zero = [i for i, q in enumerate(pool) if q["variant"] == 0]
    ins = [(shard_polys(pool[i]), full_rows(pool[i]), ybodies[pool[i]["client"]]) for i in packed]
    t0 = time.perf_counter()
    rc, outs = o.respond_inspiring_many(w_seed, [x[0] for x in ins], [x[1] for x in ins], [x[2] for x in ins])
    t_cpu = time.perf_counter() - t0
    for j, i in enumerate(packed):
        assert rc[j] == 0
        answers[i] = (outs[j], None)
    for i in zero:
        sp, rg = shard_polys(pool[i]), full_rows(pool[i])
        t0 = time.perf_counter()
        ct, cols = o.respond(sp, rg)
        t_cpu += time.perf_counter() - t0
        answers[i] = (ct, cols)
    log(f"cpu answers: {len(pool)} queries ({len(zero)} InsPIRe^0) in {t_cpu:.1f} s")
    answers["baseline"] = {
        "value": round(len(pool) / t_cpu, 1),
        "unit": "queries/s",
        "cores": threads,
        "kind": "port",
        "sample": f"the {len(pool)} queries of the serving pool (every served query is one of them): packed ones "
                  f"{threads} at a time, InsPIRe^0 one at a time; {t_cpu:.1f} s",
    }
    return answers


if __name__ == "__main__":
    main()
