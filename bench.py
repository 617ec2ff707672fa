"""Benchmark: plain single-trace lifted-PCS proving on one MI355X.

Workload (the single-GPU prove configuration): a Goldilocks trace of 2^22 rows
x 64 columns, uniform random from SmallRng(2025); log_blowup 2 (LDE of 2^24
rows), Poseidon2-Goldilocks-12 LMCS, DEEP at 2 extension points, FRI arity 4
down to degree 2^4, 100 queries.  One step = one full proof with the trace
already resident in HBM: commit_traces (coset LDE + LMCS tree), observe the
root, sample two points, open_with_channel (DEEP, FRI commit phase, query
openings); the proof (TranscriptData) is copied back to the host.

Prints one JSON line.  `roofline` describes the kernel family with the most
device time in the timed region, measured with HIP events on the library's
stream.  `cpu_baseline` times the oracle (the C restatement of the reference
CPU prover, OpenMP) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("p3-miden_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import numpy as np  # noqa: E402

LOG_N, WIDTH, LOG_BLOWUP = 22, 64, 2
PARAMS = (2, 2, 4, 0, 0, 100, 0)  # PcsParams(log_blowup, log_arity, log_final_degree, fold/deep pow, queries, query pow)
SEED = 2025

# MI355X published peaks
HBM_PEAK_GBS = 8000.0
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # 32-bit VALU lane-ops/s: 256 CUs x 64 lanes x 2.4 GHz


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic(n, w, b):
    """Algorithmic bytes / permutations per launch of each kernel family, for one proof."""
    L = n << b
    trace = 8 * n * w
    lde = 8 * L * w
    return {
        # LDE: read trace, write LDE (counted once over all passes of one trace)
        "lde": trace + lde,
        # leaf hashing: read every LDE element once, write one 32-byte digest per leaf
        "p2_leaf_bytes": lde + 32 * L,
        "p2_leaf_perms": L * -(-w // 8),
        # LDE + commit (the headline GB/s): trace in, LDE resident, all digest layers out
        "lde_commit": trace + lde + 32 * (2 * L - 1),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-log-rows", type=int, default=17)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: multi-GPU runs are not part of this build "
              "(independent proofs shard one per device; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import p3gpu as G

    n, w, b = 1 << LOG_N, WIDTH, LOG_BLOWUP
    log(f"generating trace 2^{LOG_N} x {w} from SmallRng({SEED})")
    host = G.fill_random(SEED, n * w).reshape(n, w)
    dev = G.Device(0)
    trace = dev.upload(host)
    del host
    cfg = G.LmcsConfig(dev, G.HASH_POSEIDON2)
    params = G.PcsParams(*PARAMS)

    def step():
        return G.prove(cfg, params, [trace])

    log(f"warmup x{args.warmup}")
    first = None
    for _ in range(args.warmup):
        first = step()
    dev.synchronize()
    dev.profile(True)
    dev.profile_reset()
    log(f"timing x{args.steps}")
    t0 = time.perf_counter()
    out = None
    for _ in range(args.steps):
        out = step()
    dev.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    if first is not None:
        assert np.array_equal(first[2], out[2]) and np.array_equal(first[3], out[3]), "proofs differ between runs"

    fams = ["lde_inverse_pass", "lde_forward_pass", "p2_leaf", "compress_layer", "compress_top", "deep_weights",
            "deep_colsum", "deep_quotient", "fri_fold"]
    per = {}
    for f in fams:
        tot, cnt = dev.profile_read(f)
        per[f] = (tot / args.steps, cnt // max(1, args.steps))
    dev.profile(False)
    dom = max(per, key=lambda k: per[k][0])
    alg = algorithmic(n, w, b)

    # the leaf-hash launch that hashes the 2^24-row LDE: one per proof for the trace
    # tree plus one per FRI round; its per-launch duration is taken from the trace
    # tree launch, measured separately below
    log("timing commit phase (LDE + LMCS) alone")
    dev.profile(True)
    dev.profile_reset()
    tc0 = time.perf_counter()
    reps = max(3, args.steps // 2)
    for _ in range(reps):
        t = G.commit_traces(cfg, [trace], b)
        t.free()
    dev.synchronize()
    tc1 = time.perf_counter()
    commit_ms = (tc1 - tc0) * 1e3 / reps
    leaf_ms, leaf_n = dev.profile_read("p2_leaf")
    leaf_ms /= max(1, leaf_n)
    lde_ms = (dev.profile_read("lde_inverse_pass")[0] + dev.profile_read("lde_forward_pass")[0]) / reps
    dev.profile(False)

    perms = alg["p2_leaf_perms"]
    leaf_bytes = alg["p2_leaf_bytes"]
    # VALU work of the leaf kernel: 32-bit lane-ops per Poseidon2 permutation, from the
    # instruction count of the compiled kernel (SQ_INSTS_VALU per wave / 64 leaves per wave
    # / permutations per leaf) recorded in profiles/p2_leaf_c2.md
    valu_ops_per_perm = None
    prof_file = os.path.join(ROOT, "profiles", "p2_leaf_c2.md")
    traffic = None
    if os.path.exists(prof_file):
        for line in open(prof_file):
            if line.startswith("valu_lane_ops_per_perm:"):
                valu_ops_per_perm = float(line.split(":")[1])
            if line.startswith("hbm_bytes_per_launch:"):
                traffic = float(line.split(":")[1])
    roof = {
        "kernel": "p2_leaf (Poseidon2-GL12 LMCS leaf hashing, 2^24 leaves x 8 permutations)",
        "bound": "valu",
        "achieved": None,
        "peak": round(VALU_PEAK_TOPS, 2),
        "unit": "TOP/s",
        "frac": None,
        "traffic": traffic,
        "launch_ms": round(leaf_ms, 3),
        "permutations_per_s": round(perms / (leaf_ms * 1e-3), 1) if leaf_ms else None,
        "hbm_view": {"achieved": round(leaf_bytes / (leaf_ms * 1e-3) / 1e9, 1) if leaf_ms else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s"},
    }
    if valu_ops_per_perm and leaf_ms:
        ach = perms * valu_ops_per_perm / (leaf_ms * 1e-3) / 1e12
        roof["achieved"] = round(ach, 2)
        roof["frac"] = round(ach / VALU_PEAK_TOPS, 3)

    cpu = None
    if not args.no_cpu_baseline:
        import oracle as O

        ls = args.cpu_sample_log_rows
        log(f"cpu baseline: oracle on 2^{ls} x {w} (same parameters), {O.get_threads()} threads")
        tr = [O.rand_matrix(SEED, 1 << ls, w)]
        c0 = time.perf_counter()
        O.pcs_trace_prove(O.PcsParams(*PARAMS), O.HASH_P2, tr)
        c1 = time.perf_counter()
        scale = (1 << LOG_N) // (1 << ls)
        cpu = {
            "value": round(1.0 / ((c1 - c0) * scale), 6),
            "unit": "proofs/s",
            "cores": O.get_threads(),
            "kind": "port",
            "sample": f"oracle proof of a 2^{ls} x {w} trace with the same parameters took {c1 - c0:.2f} s; "
                      f"scaled x{scale} (linear in rows) to the 2^{LOG_N}-row workload",
        }

    line = {
        "metric": "proofs/s batched vs plain on 1 MI355X, bit-exact; LDE+commit GB/s; p99 latency",
        "value": round(1000.0 / ms, 4),
        "unit": "proofs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (Goldilocks)",
        "data": "synthetic: uniform Goldilocks trace from SmallRng(2025)",
        "config": {"workload": "plain single-trace prove: 2^22 x 64 trace, log_blowup 2 (LDE 2^24 rows), "
                               "Poseidon2 LMCS, DEEP at 2 points, FRI arity 4 to degree 16, 100 queries",
                   "mode": "plain", "parallelism": "1 GPU"},
        "lde_commit_gbps": round(alg["lde_commit"] / (commit_ms * 1e-3) / 1e9, 1),
        "commit_ms": round(commit_ms, 3),
        "lde_ms": round(lde_ms, 3),
        "kernel_ms_per_proof": {k: round(v[0], 3) for k, v in per.items()},
        "dominant_kernel": dom,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    dev.close()


if __name__ == "__main__":
    main()
