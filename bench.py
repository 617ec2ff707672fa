"""Benchmark of the lifted-PCS prover core on one MI355X.

Headline (`value`): plain single-trace proving of a Goldilocks trace of 2^22
rows x 64 columns, uniform random from SmallRng(2025); log_blowup 2 (LDE of
2^24 rows), Poseidon2-Goldilocks-12 LMCS, DEEP at 2 extension points, FRI
arity 4 down to degree 2^4, 100 queries.  One step = one full proof with the
trace already resident in HBM: commit_traces (coset LDE + LMCS tree), observe
the root, sample two points, open_with_channel (DEEP, FRI commit phase, query
openings); the proof (TranscriptData) is copied back to the host.

Other sections of the same JSON line:
  batched  - 12 traces of mixed heights (2^12..2^22) and widths (8..256),
             log_blowup 1, one lifted proof: the batched commit plan (grouped
             NTT launches, hipGraph replay, autotuned shapes) against the
             per-trace launch path, with the two proofs compared word for word;
  keccak   - 2^16 x 2600 bit-structured trace, Keccak LMCS, arity 2, 100 queries;
  quotient_commit - commit_quotient of quotient evaluations resident in HBM
             (N = 2^16, 2^17 with D = 2, B = 2; N = 2^22 with D = 4, B = 4);
  serving  - Poisson arrivals of prove requests (1-8 traces of 2^14..2^22 rows,
             16..128 columns, log_blowup 1..3) into the prove-request server;
             saturation throughput, then p50/p99 latency at 0.5x, 0.8x and 1.0x
             of it, against the same requests proven one after another.

`roofline` describes the kernel family with the most device time in the
headline proof, timed with HIP events on the library's stream.
`cpu_baseline` times the oracle (the C restatement of the reference CPU prover,
OpenMP) on a bounded sample of the headline workload.
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("p3-miden_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import numpy as np  # noqa: E402

SEED = 2025

# MI355X published peaks
HBM_PEAK_GBS = 8000.0
# 32-bit VALU lane-ops/s at 2 wave64 instructions per SIMD per quad-cycle (the published 157.3 TFLOP/s
# FP32 vector rate / 2 flops per FMA) = 256 CUs x 128 lanes x 2.4 GHz; the measured peak replaces it
# when the calibrate section runs
VALU_PEAK_TOPS = 256 * 128 * 2.4e9 / 1e12
# VALU issue slots: 1,024 SIMDs (256 CUs x 4), one wave64 VALU instruction per SIMD per 4 cycles at the
# 2.4 GHz peak engine clock, in 10^9 quad-cycles/s
ISSUE_PEAK_G = 1024 * 2.4e9 / 4 / 1e9


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def lde_commit_bytes(shapes, log_blowup):
    """Algorithmic bytes of LDE + commit: read each trace, write its LDE, write all digest layers."""
    B = 1 << log_blowup
    L = max(h for h, _ in shapes) * B
    return sum(8 * h * w * (1 + B) for h, w in shapes) + 32 * (2 * L - 1)


def leaf_perms(shapes, log_blowup, rate):
    B = 1 << log_blowup
    return sum(h * B * -(-w // rate) for h, w in shapes)


def read_profile_constants():
    """Per-launch VALU instructions, issue occupancy and HBM traffic of the leaf kernel (rocprofv3 PMC runs)."""
    out = {}
    path = os.path.join(ROOT, "profiles", "valu_issue_c2_pmc.md")
    if os.path.exists(path):
        for line in open(path):
            for key in ("valu_insts_per_launch", "valu_issue_cycles_per_launch", "valu_issue_busy",
                        "valu_paired_share", "hbm_bytes_per_launch"):
                if line.startswith(key + ":"):
                    out[key] = float(line.split(":")[1])
    return out


def timed(fn, steps, dev):
    dev.synchronize()
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = fn()
    dev.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps, out


# ------------------------------------------------------------------ measured peaks
def bench_calibrate(dev):
    """Sustained VALU issue rates per instruction class and device copy bandwidth, measured on this
    GPU: the ceilings the kernels' roofline fractions are quoted against beside the published ones."""
    valu = [dev.calibrate_valu(k) for k in range(len(dev.VALU_KINDS))]
    copy = dev.calibrate_copy(4 << 30, 5)
    add = valu[0]
    return {
        "valu": {v["kind"]: {"ms": round(v["ms"], 3), "wave_insts_per_s": round(v["wave_insts_per_s"] / 1e12, 4),
                             "lane_ops_tops": round(v["lane_ops_per_s"] / 1e12, 2)} for v in valu},
        "valu_issue_peak_tops": round(add["lane_ops_per_s"] / 1e12, 2),
        "copy_kernel_gbs": round(copy["kernel_gbs"], 1),
        "copy_memcpy_gbs": round(copy["memcpy_gbs"], 1),
        "note": "valu: 8192 x 256 lanes, independent chains of one instruction (inline asm); wave_insts_per_s in "
                "10^12 wave64 instructions/s, lane_ops = 64 x that.  copy: 4 GiB device-to-device, read + write "
                "bytes per second",
    }


# ------------------------------------------------------------------ headline: plain single-trace proof
def bench_c2(G, W, dev, steps, warmup, peaks=None):
    n, w, b = 1 << W.C2_LOG_ROWS, W.C2_WIDTH, W.C2_PARAMS[0]
    log(f"[c2] generating trace 2^{W.C2_LOG_ROWS} x {w} from SmallRng({SEED})")
    host = G.fill_random(SEED, n * w).reshape(n, w)
    trace = dev.upload(host)
    del host
    cfg = G.LmcsConfig(dev, G.HASH_POSEIDON2)
    params = G.PcsParams(*W.C2_PARAMS)

    def step():
        return G.prove(cfg, params, [trace])

    log(f"[c2] warmup x{warmup}")
    first = None
    for _ in range(warmup):
        first = step()
    dev.profile(True)
    dev.profile_reset()
    log(f"[c2] timing x{steps}")
    ms, out = timed(step, steps, dev)
    if first is not None:
        assert np.array_equal(first[2], out[2]) and np.array_equal(first[3], out[3]), "proofs differ between runs"
    fams = ["lde_inverse_pass", "lde_forward_pass", "p2_leaf", "compress_layer", "compress_top", "deep_weights",
            "deep_colsum", "deep_quotient", "fri_fold"]
    per = {}
    for f in fams:
        tot, cnt = dev.profile_read(f)
        per[f] = (tot / steps, cnt // max(1, steps))
    dev.profile(False)
    dom = max(per, key=lambda k: per[k][0])

    log(f"[c2] the same proofs as a stream through the prove-request server (2 workers) x{steps}")
    srv = G.Server(0, workers=2, hbm_budget_bytes=100 << 30, max_batch=1)
    for r in [srv.submit([trace], params) for _ in range(2)]:
        srv.result(r)
    t0 = time.perf_counter()
    tickets = [srv.submit([trace], params) for _ in range(steps)]
    res = [srv.result(t) for t in tickets]
    stream_ms = (time.perf_counter() - t0) * 1e3 / steps
    stream_same = all(np.array_equal(r.fields, out[2]) and np.array_equal(r.commitments, out[3]) for r in res)
    srv.close()
    dev.trim()

    log("[c2] timing the commit phase (LDE + LMCS) alone")
    dev.profile(True)
    dev.profile_reset()
    reps = max(3, steps // 2)

    def commit():
        t = G.commit_traces(cfg, [trace], b)
        t.free()

    commit_ms, _ = timed(commit, reps, dev)
    leaf_ms, leaf_n = dev.profile_read("p2_leaf")
    leaf_ms /= max(1, leaf_n)
    lde_ms = (dev.profile_read("lde_inverse_pass")[0] + dev.profile_read("lde_forward_pass")[0]) / reps
    dev.profile(False)
    trace.free()
    dev.trim()

    shapes = [(n, w)]
    perms = leaf_perms(shapes, b, 8)
    lde_bytes = 8 * n * w * (1 + (1 << b))
    pc = read_profile_constants()
    leaf_bytes = 8 * (n << b) * w + 32 * (n << b)  # read every LDE word once, write the 2^24 leaf digests
    peak_tops = (peaks or {}).get("valu_issue_peak_tops") or VALU_PEAK_TOPS
    issue = pc.get("valu_issue_cycles_per_launch")
    insts = pc.get("valu_insts_per_launch")
    roof = {
        "kernel": "p2_leaf (Poseidon2-GL12 LMCS leaf hashing of the 2^24 x 64 LDE: 2^24 leaves x 8 permutations)",
        "bound": "valu",
        "achieved": round(issue / (leaf_ms * 1e-3) / 1e9, 1) if issue and leaf_ms else None,
        "peak": ISSUE_PEAK_G,
        "unit": "G VALU issue quad-cycles/s",
        "frac": round(issue / (leaf_ms * 1e-3) / 1e9 / ISSUE_PEAK_G, 3) if issue and leaf_ms else None,
        "traffic": pc.get("hbm_bytes_per_launch"),
        "launch_ms": round(leaf_ms, 3),
        "permutations_per_s": round(perms / (leaf_ms * 1e-3), 1) if leaf_ms else None,
        "valu_issue_busy": pc.get("valu_issue_busy"),
        "valu_paired_share": pc.get("valu_paired_share"),
        "valu_insts_per_permutation": round(insts * 64 / perms) if insts else None,
        "valu_lane_ops_tops": round(insts * 64 / (leaf_ms * 1e-3) / 1e12, 2) if insts and leaf_ms else None,
        "valu_lane_ops_peak_tops": round(peak_tops, 2),
        "valu_lane_ops_peak_source": "measured v_add_u32 issue rate in this run (bench.py calibrate)" if peaks
                                     else "published FP32 vector rate / 2 (calibrate section not run)",
        "hbm": {"algorithmic_bytes": leaf_bytes,
                "achieved": round(leaf_bytes / (leaf_ms * 1e-3) / 1e9, 1) if leaf_ms else None,
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(leaf_bytes / (leaf_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if leaf_ms else None,
                "measured_copy_gbs": (peaks or {}).get("copy_kernel_gbs")},
        "note": "the VALU issue port: quad-cycles per launch in which a SIMD issued a VALU instruction "
                "(SQ_ACTIVE_INST_VALU - SQ_ACTIVE_INST_VALU2, rocprofv3 PMC) / the launch time from HIP events "
                "in this run, against 1,024 SIMDs x 2.4 GHz peak engine clock / 4; valu_issue_busy = the same "
                "share of the SIMDs' busy quad-cycles at the clock they ran (PMC); lane-ops = SQ_INSTS_VALU x 64 "
                "/ launch time",
    }
    others = {
        "lde (inverse + forward passes)": {"ms": round(lde_ms, 3), "bytes": lde_bytes,
                                           "gbps": round(lde_bytes / (lde_ms * 1e-3) / 1e9, 1) if lde_ms else None,
                                           "peak": HBM_PEAK_GBS},
    }
    cs_ms = per["deep_colsum"][0]
    if cs_ms:
        cs_bytes = 8 * n * w  # each trace's first n rows of its LDE
        others["deep_colsum"] = {"ms": round(cs_ms, 3), "bytes": cs_bytes,
                                 "gbps": round(cs_bytes / (cs_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS}
    ff_ms = per["fri_fold"][0]
    if ff_ms:
        # every round reads its m extension elements and writes m / a of them
        lb, la = W.C2_PARAMS[0], W.C2_PARAMS[1]
        m, ff_bytes = n << lb, 0
        for _ in range(params.num_rounds((n << lb).bit_length() - 1)):
            ff_bytes += 16 * m + 16 * (m >> la)
            m >>= la
        others["fri_fold (all rounds)"] = {"ms": round(ff_ms, 3), "bytes": ff_bytes,
                                           "gbps": round(ff_bytes / (ff_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                                           "note": "rounds below ~2^16 elements are launch-latency bound"}
    dq_ms = per["deep_quotient"][0]
    if dq_ms:
        dq_bytes = 8 * (n << b) * w + 16 * (n << b)
        others["deep_quotient"] = {"ms": round(dq_ms, 3), "bytes": dq_bytes,
                                   "gbps": round(dq_bytes / (dq_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS}
    return {
        "ms": ms,
        "lde_commit_gbps": round(lde_commit_bytes(shapes, b) / (commit_ms * 1e-3) / 1e9, 1),
        "commit_ms": round(commit_ms, 3),
        "kernel_ms_per_proof": {k: round(v[0], 3) for k, v in per.items()},
        "served_stream": {"workers": 2, "proofs": steps, "proofs_per_s": round(1000.0 / stream_ms, 3),
                          "ms_per_proof": round(stream_ms, 3), "speedup_vs_plain": round(ms / stream_ms, 3),
                          "bit_exact_vs_plain": bool(stream_same),
                          "note": "the same proofs submitted together to p3g_server with two workers: one "
                                  "proof's latency-bound tree tops, small FRI rounds and host transcript work "
                                  "overlap the other's kernels"},
        "dominant_kernel": dom,
        "roofline": roof,
        "other_kernels_hbm": others,
    }


# ------------------------------------------------------------------ headline with the trace in host memory
def bench_c2_host(G, W, dev, steps, warmup):
    """The headline proof when the caller hands over a host trace (as the reference's
    RowMajorMatrix): the 2 GiB trace crosses PCIe on every proof.
      - pageable: one proof from ordinary host memory (the runtime stages the copy);
      - pinned, sequential: upload, then prove, one after another on one context;
      - pinned, served: the prove-request server with two workers (each uploads, then proves),
        and with the upload stage added, which copies the next requests' traces on a stream
        of its own while the current proofs compute.
    Every proof is compared with the resident-trace proof of the same values."""
    n, w = 1 << W.C2_LOG_ROWS, W.C2_WIDTH
    pin = G.PinnedArray((n, w))
    pin.array[:] = G.fill_random(SEED, n * w).reshape(n, w)
    pageable = np.array(pin.array)
    cfg = G.LmcsConfig(dev, G.HASH_POSEIDON2)
    params = G.PcsParams(*W.C2_PARAMS)
    resident = dev.upload(pin.array)
    ref = G.prove(cfg, params, [resident])
    resident.free()
    dev.trim()

    def same(a_fields, a_comms):
        return np.array_equal(a_fields, ref[2]) and np.array_equal(a_comms, ref[3])

    ok = True
    t0 = time.perf_counter()
    out = G.prove(cfg, params, [pageable])
    pageable_ms = (time.perf_counter() - t0) * 1e3
    ok = ok and same(out[2], out[3])
    del pageable
    for _ in range(warmup):
        G.prove(cfg, params, [pin.array])
    log(f"[c2-host] pinned host trace, sequential x{steps}")
    seq_ms, out = timed(lambda: G.prove(cfg, params, [pin.array]), steps, dev)
    ok = ok and same(out[2], out[3])
    dev.trim()
    def served(workers, upload_ahead, count=12):
        # a stream of `count` requests (about a second of work), after 3 untimed ones
        srv = G.Server(0, workers=workers, hbm_budget_bytes=100 << 30, max_batch=1, upload_ahead=upload_ahead)
        for r in [srv.submit([pin.array], params) for _ in range(3)]:
            srv.result(r)
        t0 = time.perf_counter()
        tickets = [srv.submit([pin.array], params) for _ in range(count)]
        res = [srv.result(t) for t in tickets]
        ms = (time.perf_counter() - t0) * 1e3 / count
        srv.close()
        return ms, all(same(r.fields, r.commitments) for r in res)

    log("[c2-host] pinned host trace, server with 2 workers x12")
    over_ms, ok2 = served(2, 0)
    log("[c2-host] pinned host trace, server with 2 workers and the upload stage (4 ahead) x12")
    ahead_ms, ok3 = served(2, 4)
    ok = ok and ok2 and ok3
    pin.free()
    return {
        "workload": "the headline proof (2^22 x 64, C2 parameters) with the 2 GiB trace in host memory",
        "pageable_one_proof_ms": round(pageable_ms, 1),
        "pinned_sequential_proofs_per_s": round(1000.0 / seq_ms, 3),
        "pinned_two_workers_proofs_per_s": round(1000.0 / over_ms, 3),
        "pinned_upload_ahead_proofs_per_s": round(1000.0 / ahead_ms, 3),
        "bit_exact_vs_resident": bool(ok),
    }


# ------------------------------------------------------------------ batched multi-trace proof
def bench_c3(G, W, dev, steps, warmup):
    shapes = W.c3_shapes(SEED)
    params = G.PcsParams(*W.c3_params())
    cfg = G.LmcsConfig(dev, G.HASH_POSEIDON2)
    log(f"[c3] planning {len(shapes)} traces {shapes} (autotune)")
    plan = G.CommitPlan(cfg, shapes, 1, graph=True, autotune=True)
    plan.fill_random(SEED)
    inputs = [G.DeviceMatrix(dev, plan.input_ptr(i), h, w, owned=False) for i, (h, w) in enumerate(shapes)]

    def plain():
        return G.prove(cfg, params, inputs)

    def batched():
        return plan.prove(params)

    a, bt = plain(), batched()
    bit_exact = (a[0].tolist() == bt[0].tolist() and np.array_equal(a[2], bt[2]) and np.array_equal(a[3], bt[3]))
    for _ in range(max(0, warmup - 1)):
        plain()
        batched()
    log(f"[c3] timing x{steps} plain, x{steps} batched")
    plain_ms, _ = timed(plain, steps, dev)
    batched_ms, _ = timed(batched, steps, dev)

    def commit_plain():
        G.commit_traces(cfg, inputs, 1).free()

    def commit_batched():
        plan.run().free()

    cp_ms, _ = timed(commit_plain, steps, dev)
    cb_ms, _ = timed(commit_batched, steps, dev)
    info = plan.info()
    plan.free()
    dev.trim()
    nbytes = lde_commit_bytes(shapes, 1)
    return {
        "workload": f"12 traces, heights 2^12..2^22 log-uniform, widths 8..256 (seed {SEED}), log_blowup 1, "
                    "one lifted proof (Poseidon2 LMCS, arity 2 to degree 8, 100 queries)",
        "shapes": shapes,
        "plain_proofs_per_s": round(1000.0 / plain_ms, 3),
        "batched_proofs_per_s": round(1000.0 / batched_ms, 3),
        "speedup": round(plain_ms / batched_ms, 3),
        "bit_exact_batched_vs_plain": bool(bit_exact),
        "plain_commit_ms": round(cp_ms, 3),
        "batched_commit_ms": round(cb_ms, 3),
        "batched_lde_commit_gbps": round(nbytes / (cb_ms * 1e-3) / 1e9, 1),
        "plan": info,
    }


# ------------------------------------------------------------------ Keccak wide trace
def bench_c4(G, W, dev, steps, warmup):
    log("[c4] generating bit-structured 2^16 x 2600 trace")
    host = W.c4_trace(SEED)
    trace = dev.upload(host)
    del host
    cfg = G.LmcsConfig(dev, G.HASH_KECCAK)
    params = G.PcsParams(*W.C4_PARAMS)

    def step():
        return G.prove(cfg, params, [trace])

    first = None
    for _ in range(max(1, warmup)):
        first = step()
    ms, out = timed(step, steps, dev)
    assert np.array_equal(first[2], out[2]), "proofs differ between runs"
    # the trace tree's leaf launch alone (the FRI round trees hash with the same kernel)
    dev.profile(True)
    dev.profile_reset()
    for _ in range(steps):
        G.commit_traces(cfg, [trace], 1).free()
    dev.synchronize()
    leaf_ms, leaf_n = dev.profile_read("keccak_leaf")
    dev.profile(False)
    trace.free()
    dev.trim()
    n, w = 1 << W.C4_LOG_ROWS, W.C4_WIDTH
    perms = leaf_perms([(n, w)], 1, 17)
    per_launch = leaf_ms / max(1, leaf_n)
    return {
        "workload": "2^16 x 2600 trace (90% 0/1 columns, 16-bit limbs), log_blowup 1, Keccak LMCS, arity 2 to "
                    "degree 8, 100 queries",
        "proofs_per_s": round(1000.0 / ms, 3),
        "ms_per_proof": round(ms, 3),
        "keccak_leaf_ms_trace_tree": round(per_launch, 3),
        "keccak_f_per_s": round(perms / (per_launch * 1e-3), 1) if per_launch else None,
    }


# ------------------------------------------------------------------ serving
def bench_c5(G, W, dev, n_requests, workers):
    budget = 250 * 10**9
    rng = np.random.default_rng(SEED)
    pool_h, pool_w = 1 << 22, 256  # 8.6 GB of device-resident synthetic trace data
    log("[c5] filling the device trace pool")
    pool = dev.alloc_matrix(pool_h, pool_w)
    pool.fill_random(SEED)

    def make_request():
        while True:
            shapes, lb = W.c5_request(rng)
            traces = []
            for h, w in shapes:
                off = int(rng.integers(0, pool_h * pool_w - h * w + 1)) & ~31
                traces.append(pool.view(off, h, w))
            params = (lb, 2, 4, 0, 0, 100, 0)
            if G.request_footprint(G.HASH_POSEIDON2, G.PcsParams(*params), traces) <= budget:
                return traces, params, shapes

    reqs = [make_request() for _ in range(n_requests)]

    log("[c5] plain: the requests one after another on one context")
    cfg = G.LmcsConfig(dev, G.HASH_POSEIDON2)
    # steady state for both paths: one untimed pass over the whole request set
    # first (tables and allocation caches), then the timed pass
    for tr, p, _ in reqs:
        G.prove(cfg, G.PcsParams(*p), tr)
    t0 = time.perf_counter()
    plain = [G.prove(cfg, G.PcsParams(*p), tr) for tr, p, _ in reqs]
    plain_s = time.perf_counter() - t0
    dev.trim()

    srv = G.Server(0, workers=workers, hbm_budget_bytes=budget)
    # warm every worker context and the shared block pool with the whole set
    for r in [srv.submit(tr, G.PcsParams(*p)) for tr, p, _ in reqs]:
        srv.result(r)
    st_warm = srv.stats()

    log(f"[c5] saturation: {n_requests} requests submitted at once")
    t0 = time.perf_counter()
    tickets = [srv.submit(tr, G.PcsParams(*p)) for tr, p, _ in reqs]
    served = [srv.result(t) for t in tickets]
    sat_s = time.perf_counter() - t0
    sat_rate = n_requests / sat_s
    same = all(np.array_equal(a[2], s_.fields) and np.array_equal(a[3], s_.commitments)
               and a[0].tolist() == s_.root.tolist() for a, s_ in zip(plain, served))

    rates = []
    for frac in (0.5, 0.8, 1.0):
        rate = frac * sat_rate
        gaps = rng.exponential(1.0 / rate, size=n_requests)
        log(f"[c5] Poisson arrivals at {rate:.2f} requests/s")
        tickets = []
        start = time.perf_counter()
        due = start
        for (tr, p, _), g in zip(reqs, gaps):
            due += g
            d = due - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets.append(srv.submit(tr, G.PcsParams(*p)))
        res = [srv.result(t) for t in tickets]
        lat = np.array([r.latency_ms for r in res])
        span = (max(r.done_ms for r in res) - min(r.submit_ms for r in res)) / 1e3
        same = same and all(np.array_equal(a[2], r.fields) for a, r in zip(plain, res))
        rates.append({"offered_per_s": round(rate, 3), "proofs_per_s": round(n_requests / span, 3),
                      "p50_ms": round(float(np.percentile(lat, 50)), 2),
                      "p99_ms": round(float(np.percentile(lat, 99)), 2)})
    st = srv.stats()
    srv.close()
    pool.free()
    dev.trim()
    return {
        "workload": f"{n_requests} requests: 1-8 traces, heights 2^14..2^22 log-uniform, widths 16..128, "
                    "log_blowup 1..3, Poseidon2 LMCS, arity 4 to degree 16, 100 queries; traces resident in HBM",
        "workers": workers,
        "hbm_budget_gb": budget / 1e9,
        "saturation_proofs_per_s": round(sat_rate, 3),
        "plain_sequential_proofs_per_s": round(n_requests / plain_s, 3),
        "speedup_vs_plain": round(plain_s / sat_s, 3),
        "bit_exact_served_vs_plain": bool(same),
        "poisson": rates,
        "peak_inflight_gb": round(st["peak_inflight_bytes"] / 1e9, 2),
        "peak_running": st["peak_running"],
        "allocator": {"hipMalloc_calls": st["n_malloc"], "hipMalloc_ms": round(st["malloc_ms"], 1),
                      "hipFree_calls": st["n_free"], "hipFree_ms": round(st["free_ms"], 1),
                      "in_warm_pass": {"hipMalloc_calls": st_warm["n_malloc"],
                                       "hipMalloc_ms": round(st_warm["malloc_ms"], 1)},
                      "in_timed_passes": {"hipMalloc_calls": st["n_malloc"] - st_warm["n_malloc"],
                                          "hipMalloc_ms": round(st["malloc_ms"] - st_warm["malloc_ms"], 1)}},
    }


# ------------------------------------------------------------------ many small proofs
BATCH_WINDOW_MS = 2.0


def bench_small(G, W, dev, n_proofs, workers, max_batch=128):
    """n_proofs independent proofs of 2^14 x 32 traces (the CPU-sized configuration's shape):
    one after another on one context (plain) against the prove-request server."""
    n, w = 1 << W.C1_LOG_ROWS, W.C1_WIDTH
    params = G.PcsParams(*W.C1_PARAMS)
    pool = dev.alloc_matrix(n * n_proofs, w)
    pool.fill_random(SEED)
    traces = [pool.view(i * n * w, n, w) for i in range(n_proofs)]
    cfg = G.LmcsConfig(dev, G.HASH_POSEIDON2)
    for t in traces[:2]:
        G.prove(cfg, params, [t])
    t0 = time.perf_counter()
    plain = [G.prove(cfg, params, [t]) for t in traces]
    plain_s = time.perf_counter() - t0
    # two workers with batches of up to 128: with four (batches of 64) the
    # concurrent lock-step batches interleave their synchronised phases on the
    # GPU, and the stream measured 2,690-3,080 against 3,240-3,800 proofs/s
    srv = G.Server(0, workers=workers, hbm_budget_bytes=64 * 10**9, max_batch=max_batch, batch_window_ms=BATCH_WINDOW_MS)
    # warm every worker context and the shared block pool with the whole set,
    # twice: the four workers interleave differently from pass to pass, and
    # the second pass fills the size classes the first one's mix did not use
    for _ in range(2):
        for r in [srv.submit([t], params) for t in traces]:
            srv.result(r)
    s0 = srv.stats()
    t0 = time.perf_counter()
    res = [srv.result(tk) for tk in [srv.submit([t], params) for t in traces]]
    served_s = time.perf_counter() - t0
    sst = srv.stats()
    same = all(np.array_equal(a[2], r.fields) and np.array_equal(a[3], r.commitments) for a, r in zip(plain, res))
    lat = np.array([r.latency_ms for r in res])
    # a sustained stream: the set four times back to back, so that one batch's host phases
    # (transcripts, query indices, hints) overlap the next batch's kernels
    t0 = time.perf_counter()
    stream = [srv.result(tk) for tk in [srv.submit([traces[i % n_proofs]], params) for i in range(4 * n_proofs)]]
    stream_s = time.perf_counter() - t0
    same = same and all(np.array_equal(plain[i % n_proofs][2], r.fields) and
                        np.array_equal(plain[i % n_proofs][3], r.commitments) for i, r in enumerate(stream))
    srv.close()
    # lock-step batches: every device phase one launch over the whole batch
    batched = {}
    for bsz in (16, 64, 256):
        if bsz > n_proofs:
            continue
        G.prove_many(cfg, params, [[t] for t in traces[:bsz]])  # warm (tables, allocation cache)
        t0 = time.perf_counter()
        outs = []
        for b0 in range(0, n_proofs, bsz):
            outs += G.prove_many(cfg, params, [[t] for t in traces[b0:b0 + bsz]])
        el = time.perf_counter() - t0
        same = same and all(np.array_equal(a[2], o[2]) and np.array_equal(a[3], o[3]) for a, o in zip(plain, outs))
        batched[str(bsz)] = {"proofs_per_s": round(n_proofs / el, 2), "ms_per_batch": round(el * 1e3 * bsz / n_proofs, 2),
                             "speedup_vs_plain": round(plain_s / el, 2)}
    pool.free()
    dev.trim()
    return {
        "workload": f"{n_proofs} independent proofs of 2^14 x 32 traces (log_blowup 1, arity 2 to degree 8, 64 queries)",
        "plain_proofs_per_s": round(n_proofs / plain_s, 2),
        "plain_ms_per_proof": round(plain_s * 1e3 / n_proofs, 3),
        "served_proofs_per_s": round(n_proofs / served_s, 2),
        "speedup": round(plain_s / served_s, 3),
        "workers": workers,
        "served_batches": {"batches": sst["batches"] - s0["batches"],
                           "batched_requests": sst["batched_requests"] - s0["batched_requests"]},
        "served_hipMalloc_calls": sst["n_malloc"] - s0["n_malloc"],
        "bit_exact_served_and_batched_vs_plain": bool(same),
        "served_p50_ms": round(float(np.percentile(lat, 50)), 2),
        "served_p99_ms": round(float(np.percentile(lat, 99)), 2),
        "sustained_stream": {"requests": 4 * n_proofs, "proofs_per_s": round(4 * n_proofs / stream_s, 2),
                             "speedup_vs_plain": round(4 * plain_s / stream_s, 3)},
        "batched_lockstep": batched,
    }


# ------------------------------------------------------------------ grinding (the reference binary's parameters)
def bench_grind(G, dev, steps, warmup):
    """PcsParams(3, 1, 0, 16, 0, 100, 0) of the reference's end-to-end Goldilocks binary
    (p3-miden-lifted-examples/src/bin/lifted_miden.rs:34-54) on its two trace shapes,
    2^18 x 51 and 2^19 x 20 (LDE 2^22 rows): 19 FRI commit rounds, each grinding 16 bits,
    with the witness search on the GPU inside the device Fiat-Shamir rounds."""
    shapes = [(1 << 18, 51), (1 << 19, 20)]
    traces = []
    for i, (h, w) in enumerate(shapes):
        m = dev.alloc_matrix(h, w)
        m.fill_random(SEED + i)
        traces.append(m)
    cfg = G.LmcsConfig(dev, G.HASH_POSEIDON2)
    params = G.PcsParams(3, 1, 0, 16, 0, 100, 0)

    def step():
        return G.prove(cfg, params, traces)

    first = None
    for _ in range(max(1, warmup)):
        first = step()
    dev.profile(True)
    dev.profile_reset()
    ms, out = timed(step, steps, dev)
    assert np.array_equal(first[2], out[2]), "proofs differ between runs"
    g_ms, g_n = dev.profile_read("grind")
    dev.profile(False)
    for m in traces:
        m.free()
    dev.trim()
    # the same 16-bit search serially on the host (this library's transcript, one thread)
    tx = G.ProverTranscript()
    tx.observe(list(range(5)))
    h0 = time.perf_counter()
    tx.grind(16)
    host_ms = (time.perf_counter() - h0) * 1e3
    rounds = g_n // max(1, steps)
    return {
        "workload": "PcsParams(3, 1, 0, 16, 0, 100, 0) (the reference binary's), traces 2^18 x 51 and 2^19 x 20 "
                    "(LDE 2^22 rows), Poseidon2 LMCS: 19 FRI rounds with 16-bit folding grinding",
        "proofs_per_s": round(1000.0 / ms, 3),
        "ms_per_proof": round(ms, 3),
        "grind_ms_per_proof": round(g_ms / max(1, steps), 3),
        "grind_searches_per_proof": rounds,
        "host_serial_grind_ms_16_bits": round(host_ms, 2),
        "host_serial_grind_ms_per_proof_est": round(host_ms * rounds, 1),
    }


# ------------------------------------------------------------------ quotient commitment
def bench_quotient(G, dev, steps, warmup):
    """commit_quotient (quotient.rs:133-238) on quotient evaluations resident in HBM:
    the reference bench's shapes (N = 2^16, 2^17; D = 2; B = 2) and the quotient of the
    headline trace (N = 2^22, D = 4, B = 4)."""
    cfg = G.LmcsConfig(dev, G.HASH_POSEIDON2)
    out = []
    for logn, d, logb in ((16, 2, 1), (17, 2, 1), (22, 4, 2)):
        n, B = 1 << logn, 1 << logb
        q = dev.alloc_matrix(n, 2 * d)
        q.fill_random(SEED + logn)

        def run():
            G.commit_quotient(cfg, q, d, logb).free()

        for _ in range(warmup):
            run()
        dev.synchronize()
        reps = max(3, steps)
        t0 = time.perf_counter()
        for _ in range(reps):
            run()
        dev.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        L = n * B
        algo = 16 * n * d + 8 * L * 2 * d + 32 * (2 * L - 1)
        perms = L * -(-2 * d // 8) + (L - 1)
        out.append({"shape": f"N=2^{logn}, D={d}, log_blowup={logb} (LDE {L} x {2 * d})", "ms": round(ms, 3),
                    "gbps": round(algo / (ms * 1e-3) / 1e9, 1),
                    "poseidon2_perms_per_s": round(perms / (ms * 1e-3), 1)})
        if logn == 22:
            # the accumulation steps ahead of the commit, on the same n*d elements:
            # each reads and writes 16 bytes per element (HBM-bound)
            helpers = {}
            for name, fn in (("divide_by_vanishing", lambda: G.divide_by_vanishing(q, n, d, 7)),
                             ("cyclic_extend_and_scale (n*d/2 -> n*d)",
                              lambda: G.cyclic_extend_and_scale(q, n * d // 2, n * d, (3, 5)))):
                fn()
                dev.synchronize()
                t0 = time.perf_counter()
                for _ in range(reps):
                    fn()
                dev.synchronize()
                hms = (time.perf_counter() - t0) * 1e3 / reps
                helpers[name] = {"ms": round(hms, 3), "gbps": round(32 * n * d / (hms * 1e-3) / 1e9, 1)}
            out[-1]["accumulation_helpers"] = helpers
        q.free()
    dev.trim()
    return {"workload": "quotient commitment of N*D extension elements resident in HBM, Poseidon2 LMCS (aligned)",
            "runs": out}


# ------------------------------------------------------------------ constraint evaluation
def air_bytes_per_point(ops):
    """HBM bytes one point must move at least: every distinct main column it reads (8 B; a
    next-row value is another point's current row), every distinct aux extension column
    (16 B), and the accumulator read and written (32 B)."""
    from p3gpu import air
    main = {int(o["a"]) for o in ops if o["op"] == air.MAIN}
    aux = {int(o["a"]) for o in ops if o["op"] == air.AUX}
    return 8 * len(main) + 16 * len(aux) + 32


def bench_constraints(G, dev, steps, warmup, cpu=True):
    """evaluate_constraints_into (constraints/mod.rs:63-179) on committed LDEs resident in HBM,
    one kernel per AIR compiled at run time:
      - DummyMidenAir at the reference binary's shape (lifted_miden.rs: 2^19 x 20, log_blowup 3,
        quotient degree 8: 2^22 points), with the oracle on the host cores beside it;
      - Fibonacci with public values (selectors) and the running-sum AIR (aux extension column,
        periodic column, extension inputs) at 2^22 points;
      - a Poseidon2-GL12 AIR (3,047 ops over 248 columns) at 2^19 points and a 3000-op random
        program (every input kind) at 2^20 points, bound by arithmetic."""
    from p3gpu import example_airs as A
    cfg = G.LmcsConfig(dev, G.HASH_POSEIDON2)
    rng = np.random.default_rng(SEED)
    out = []

    def case(name, prog, trace, logb, log_d, aux=None, cpu_check=False, **kw):
        n = trace.shape[0]
        t0 = time.perf_counter()
        G.air_compile(prog)
        compile_s = time.perf_counter() - t0
        tree = G.commit_traces(cfg, [trace], logb)
        ptr, h, w = tree.matrix(0)
        main = G.DeviceMatrix(dev, ptr, h, w, owned=False)
        atree, da = None, None
        if aux is not None:
            atree = G.commit_traces(cfg, [aux], logb)
            p2, h2, w2 = atree.matrix(0)
            da = G.DeviceMatrix(dev, p2, h2, w2, owned=False)
        M = n << log_d
        acc = dev.alloc_matrix(M, 2)
        G.cyclic_extend_and_scale(acc, 0, M, (0, 0))
        alpha = (5, 8)

        def run():
            G.eval_constraints(prog, main, n, log_d, 7, alpha, acc, aux=da, **kw)

        for _ in range(warmup):
            run()
        dev.synchronize()
        reps = max(5, steps)
        dev.profile(True)
        dev.profile_reset()
        t0 = time.perf_counter()
        for _ in range(reps):
            run()
        dev.synchronize()
        wall_ms = (time.perf_counter() - t0) * 1e3 / reps
        kms, cnt = dev.profile_read("eval_constraints")
        dev.profile(False)
        kms /= cnt
        ops = prog.ops
        bpp = air_bytes_per_point(ops)
        row = {"air": name, "trace": f"2^{n.bit_length() - 1} x {trace.shape[1]}", "log_blowup": logb,
               "quotient_degree": 1 << log_d, "points": M, "ops": int(len(ops)),
               "constraints": int(sum(1 for o in ops if o["op"] >= 13)), "compile_s": round(compile_s, 2),
               "kernel_ms": round(kms, 4), "wall_ms": round(wall_ms, 4),
               "points_per_s": round(M / (kms * 1e-3), 1), "algorithmic_bytes_per_point": bpp,
               "algorithmic_gbps": round(bpp * M / (kms * 1e-3) / 1e9, 1)}
        if cpu_check:
            # the oracle (the C restatement, OpenMP on the host cores) on the same inputs:
            # it must agree word for word, and its time is the CPU figure beside the kernel's
            import oracle as O
            lde = main.download()
            acc0 = np.zeros((M, 2), np.uint64)
            G.cyclic_extend_and_scale(acc, 0, M, (0, 0))
            run()
            got = acc.download().reshape(-1, 2)
            c0 = time.perf_counter()
            want = O.eval_constraints(ops, lde, n, log_d, 7, alpha, acc=acc0)
            row["oracle_cpu_s"] = round(time.perf_counter() - c0, 3)
            row["oracle_threads"] = O.get_threads()
            row["bit_exact_vs_oracle"] = bool(np.array_equal(got, want))
            del lde, want
        out.append(row)
        log(f"constraints: {name} {row['kernel_ms']} ms ({row['algorithmic_gbps']} GB/s), compile {compile_s:.2f} s")
        acc.free()
        tree.free()
        if atree is not None:
            atree.free()

    case("DummyMidenAir (miden.rs:38-64)", A.miden_program(), A.miden_trace(1 << 19, 20, SEED), 3, 3, cpu_check=cpu)
    t, pub = A.fibonacci_trace(1 << 21, 1, 2)
    case("Fibonacci with public values", A.fibonacci_program(), t, 1, 1, publics=pub)
    n = 1 << 21
    main = rng.integers(0, A.P, size=(n, 2), dtype=np.uint64)
    aux = rng.integers(0, A.P, size=(n, 2), dtype=np.uint64)
    per = rng.integers(0, A.P, size=(8, 1), dtype=np.uint64)
    case("running sum (aux, periodic, extension inputs)", A.running_program(), main, 1, 1, aux=aux, periodic=per,
         ext_inputs=[3, 4, 5, 6])
    # Poseidon2-GL12 AIR (one permutation per row; round constants drawn here, which does
    # not change the work): 3,047 ops, 236 degree-3 constraints over 248 columns
    n = 1 << 18
    p2 = A.poseidon2_program(rng.integers(0, A.P, size=118, dtype=np.uint64),
                             rng.integers(0, A.P, size=12, dtype=np.uint64))
    case("Poseidon2-GL12 AIR (one permutation per row, 248 columns)", p2,
         rng.integers(0, A.P, size=(n, A.P2_COLS), dtype=np.uint64), 1, 1, cpu_check=cpu)
    prog = A.random_program(SEED, 3000, 9, 2, 3, 2, 2)
    case("random program, 3000 ops", prog, rng.integers(0, A.P, size=(n, 9), dtype=np.uint64), 2, 2,
         aux=rng.integers(0, A.P, size=(n, 4), dtype=np.uint64),
         periodic=rng.integers(0, A.P, size=(16, 2), dtype=np.uint64), publics=[1, 2, 3], ext_inputs=[4, 5, 6, 7])
    dev.trim()
    hbm_peak = 8000.0
    m = out[0]
    return {"workload": "constraint evaluation on the quotient domain, committed LDE in HBM (bit-reversed rows), "
                        "accumulator of n*d extension elements",
            "runs": out,
            "roofline": {"kernel": "p3g_air_kernel (DummyMidenAir, 2^22 points)", "bound": "hbm",
                         "achieved": m["algorithmic_gbps"], "peak": hbm_peak, "unit": "GB/s",
                         "frac": round(m["algorithmic_gbps"] / hbm_peak, 3),
                         "bytes_per_point": m["algorithmic_bytes_per_point"]}}


def cpu_baseline(W, log_rows, full_c2=False):
    """The oracle (the C restatement of the reference CPU prover, OpenMP) on the host cores, timed as
    p3-miden-lifted-examples/src/stats.rs:185-190 does (one warm-up, then BENCH_ITERS runs):
      - C1 (2^14 x 32, the workload the reference runs on CPU) at full size: 1 + 5 runs on all
        threads and 1 + 3 runs on one thread, median and p99;
      - the headline workload (2^22 x 64, C2 parameters) on a 2^log_rows-row sample, scaled by
        the row ratio (the prover is linear in rows at fixed width), which is `value`;
      - with full_c2, one C2 proof at full size as well (about a minute of CPU work)."""
    import oracle as O

    def pct(v, q):
        v = sorted(v)
        return v[min(len(v) - 1, int(round(q * (len(v) - 1))))]

    threads = O.get_threads()
    w = W.C2_WIDTH
    out = {"unit": "proofs/s", "cores": threads, "kind": "port"}
    c1 = [O.rand_matrix(SEED, 1 << W.C1_LOG_ROWS, W.C1_WIDTH)]
    c1p = O.PcsParams(*W.C1_PARAMS)

    def c1_runs(n):
        ts = []
        for i in range(n + 1):
            t0 = time.perf_counter()
            O.pcs_trace_prove(c1p, O.HASH_P2, c1)
            if i:
                ts.append(time.perf_counter() - t0)
        return ts

    log(f"cpu baseline: C1 2^{W.C1_LOG_ROWS} x {W.C1_WIDTH}, {threads} threads (1 + 5 runs) and 1 thread (1 + 3 runs)")
    ta = c1_runs(5)
    O.set_threads(1)
    try:
        t1 = c1_runs(3)
    finally:
        O.set_threads(threads)
    out["c1_full_size"] = {
        "workload": f"2^{W.C1_LOG_ROWS} x {W.C1_WIDTH}, PcsParams{W.C1_PARAMS}, Poseidon2",
        "all_threads": {"threads": threads, "median_s": round(pct(ta, 0.5), 4), "p99_s": round(pct(ta, 0.99), 4),
                        "proofs_per_s": round(1.0 / pct(ta, 0.5), 3)},
        "one_thread": {"median_s": round(pct(t1, 0.5), 4), "p99_s": round(pct(t1, 0.99), 4),
                       "proofs_per_s": round(1.0 / pct(t1, 0.5), 3)},
    }
    log(f"cpu baseline: oracle on 2^{log_rows} x {w} (C2 parameters), {threads} threads")
    tr = [O.rand_matrix(SEED, 1 << log_rows, w)]
    c0 = time.perf_counter()
    O.pcs_trace_prove(O.PcsParams(*W.C2_PARAMS), O.HASH_P2, tr)
    c1t = time.perf_counter()
    del tr
    scale = (1 << W.C2_LOG_ROWS) // (1 << log_rows)
    out["value"] = round(1.0 / ((c1t - c0) * scale), 6)
    out["sample"] = (f"oracle proof of a 2^{log_rows} x {w} trace with the C2 parameters took {c1t - c0:.2f} s on "
                     f"{threads} threads; scaled x{scale} (linear in rows) to the 2^{W.C2_LOG_ROWS}-row workload")
    if full_c2:
        log("cpu baseline: one C2 proof at full size")
        tr = [O.rand_matrix(SEED, 1 << W.C2_LOG_ROWS, w)]
        f0 = time.perf_counter()
        O.pcs_trace_prove(O.PcsParams(*W.C2_PARAMS), O.HASH_P2, tr)
        out["c2_full_size_s"] = round(time.perf_counter() - f0, 2)
        del tr
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--sections", default="calibrate,c2,c2host,c3,c4,c5,small,quotient,grind,constraints",
                    help="comma list of calibrate (measured peaks), c2 (headline), c2host (headline from host "
                         "memory), c3, c4, c5, small, quotient, grind (the reference binary's parameters), "
                         "constraints (constraint evaluation)")
    ap.add_argument("--small-proofs", type=int, default=256)
    ap.add_argument("--serve-requests", type=int, default=40)
    ap.add_argument("--serve-workers", type=int, default=4, help="server workers for the C5 mix")
    ap.add_argument("--small-workers", type=int, default=2, help="server workers for the small-proof stream")
    ap.add_argument("--small-max-batch", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run one 4 GiB device-to-device copy (reference dispatch for HBM byte counters)")
    ap.add_argument("--cpu-sample-log-rows", type=int, default=19)
    ap.add_argument("--cpu-full-c2", action="store_true", help="also time one C2 proof at full size on the CPU")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: multi-GPU runs are not part of this build "
              "(independent proofs shard one per device)", file=sys.stderr)
        sys.exit(2)

    import p3gpu as G
    from p3gpu import workloads as W

    sections = set(args.sections.split(","))
    dev = G.Device(0)
    if args.copy_calibration:
        a, b = dev.alloc_matrix(1 << 22, 128), dev.alloc_matrix(1 << 22, 128)
        a.fill_random(1)
        b.copy_from(a)
        a.free()
        b.free()
    peaks = bench_calibrate(dev) if "calibrate" in sections else None
    # the headline section always runs in the default configuration; leaving it
    # out of --sections is for profiling one other workload on its own
    c2 = bench_c2(G, W, dev, args.steps, args.warmup, peaks) if "c2" in sections else None
    c2h = bench_c2_host(G, W, dev, max(3, args.steps // 2), 1) if "c2host" in sections else None
    c3 = bench_c3(G, W, dev, max(2, args.steps // 3), args.warmup) if "c3" in sections else None
    c4 = bench_c4(G, W, dev, max(2, args.steps // 2), args.warmup) if "c4" in sections else None
    c5 = bench_c5(G, W, dev, args.serve_requests, args.serve_workers) if "c5" in sections else None
    small = bench_small(G, W, dev, args.small_proofs, args.small_workers, args.small_max_batch) \
        if "small" in sections else None
    quot = bench_quotient(G, dev, max(3, args.steps // 2), 1) if "quotient" in sections else None
    grind = bench_grind(G, dev, max(2, args.steps // 2), 1) if "grind" in sections else None
    cons = bench_constraints(G, dev, args.steps, 1, cpu=not args.no_cpu_baseline) \
        if "constraints" in sections else None
    cpu = None if args.no_cpu_baseline else cpu_baseline(W, args.cpu_sample_log_rows, args.cpu_full_c2)

    ms = c2["ms"] if c2 else None
    c2 = c2 or {}
    line = {
        "metric": "proofs/s batched vs plain on 1 MI355X, bit-exact; LDE+commit GB/s; p99 latency",
        "value": round(1000.0 / ms, 4) if ms else None,
        "unit": "proofs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3) if ms else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (Goldilocks)",
        "data": "synthetic: uniform Goldilocks traces from SmallRng(2025) and the device counter generator",
        "config": {"workload": "plain single-trace prove: 2^22 x 64 trace, log_blowup 2 (LDE 2^24 rows), "
                               "Poseidon2 LMCS, DEEP at 2 points, FRI arity 4 to degree 16, 100 queries",
                   "mode": "plain", "parallelism": "1 GPU"},
        "lde_commit_gbps": c2.get("lde_commit_gbps"),
        "commit_ms": c2.get("commit_ms"),
        "kernel_ms_per_proof": c2.get("kernel_ms_per_proof"),
        "dominant_kernel": c2.get("dominant_kernel"),
        "roofline": c2.get("roofline"),
        "host_input": c2h,
        "other_kernels_hbm": c2.get("other_kernels_hbm"),
        "served_stream": c2.get("served_stream"),
        "batched_vs_plain": {
            "small_proofs_lockstep": (small or {}).get("batched_lockstep"),
            "small_proofs_served_speedup": (small or {}).get("speedup"),
            "c3_one_lifted_proof_speedup": (c3 or {}).get("speedup"),
            "c2_served_stream_speedup": (c2.get("served_stream") or {}).get("speedup_vs_plain"),
        },
        "batched": c3,
        "keccak": c4,
        "serving": c5,
        "small_proofs": small,
        "quotient_commit": quot,
        "grinding": grind,
        "constraints": cons,
        "cpu_baseline": cpu,
        "measured_peaks": peaks,
    }
    print(json.dumps(line), flush=True)
    dev.close()


if __name__ == "__main__":
    main()
