"""Full-size configurations, checked through properties that do not need the
oracle to redo the whole computation:

  - the oracle's verifier accepts the GPU proof of the 2^22 x 64 trace
    (it recomputes every opened Merkle path, the DEEP quotient and every FRI
    fold at all 100 queries) and rejects it after one word is changed;
  - a shift-1 LDE of a 2^22-row trace holds the trace itself in its first
    2^22 (bit-reversed) rows: the iNTT -> scale -> NTT round trip at full size;
  - the step-level DEEP and FRI calls, composed in the opening's order, write
    the same 2^22 x 64 proof as the one-call opening;
  - the LDE is linear (LDE(a) + LDE(b) == LDE(a + b) mod p);
  - the batched commit plan of the 12-trace mixed-height workload gives the
    same root, layers and opened rows as the per-trace path, and its proof
    verifies.
"""
import numpy as np
import pytest

import oracle as O
import p3gpu as G
from p3gpu import workloads as W

pytestmark = pytest.mark.gpu

P = (1 << 64) - (1 << 32) + 1


def _bitrev_perm(logn):
    i = np.arange(1 << logn, dtype=np.uint64)
    r = np.zeros_like(i)
    for b in range(logn):
        r |= ((i >> np.uint64(b)) & np.uint64(1)) << np.uint64(logn - 1 - b)
    return r


def test_c2_proof_accepted_by_oracle_verifier(device):
    n, w = 1 << W.C2_LOG_ROWS, W.C2_WIDTH
    tr = O.rand_matrix(2025, n, w)
    params = W.C2_PARAMS
    root, z, f, c = G.prove(G.LmcsConfig(device, G.HASH_POSEIDON2), G.PcsParams(*params), [tr])
    logL = W.C2_LOG_ROWS + params[0]
    op = O.PcsParams(*params)
    assert O.pcs_trace_verify(op, O.HASH_P2, root, [w], logL, f, c) == 0
    bad = f.copy()
    bad[len(bad) // 2] = (int(bad[len(bad) // 2]) + 1) % P
    assert O.pcs_trace_verify(op, O.HASH_P2, root, [w], logL, bad, c) != 0


def test_c2_composed_steps_equal_one_call_opening(device):
    # at the full 2^22 x 64 shape: the step-level calls, composed in the opening's order, write
    # the same transcript as the one-call opening (whose proofs the oracle verifier accepts)
    from compose import compose_open
    n, w = 1 << W.C2_LOG_ROWS, W.C2_WIDTH
    params = W.C2_PARAMS
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    tr = device.alloc_matrix(n, w)
    tr.fill_random(2026)
    t = G.commit_traces(cfg, [tr], params[0])
    tx = G.ProverTranscript()
    tx.observe_commitment(t.root())
    z = [tx.sample_algebra_element(), tx.sample_algebra_element()]
    one_call = tx.clone()
    compose_open(cfg, params, [t], [w], z, tx)
    G.open_with_channel(G.PcsParams(*params), cfg, int(t.height()).bit_length() - 1, z, [t], one_call)
    f1, c1 = tx.data()
    f2, c2 = one_call.data()
    assert f1.size > 0 and np.array_equal(f1, f2) and np.array_equal(c1, c2)


def test_lde_shift_one_contains_trace_full_size(device):
    logn, w = 22, 8
    tr = O.rand_matrix(7, 1 << logn, w)
    lde = G.Dft(device).coset_lde_batch(tr, 2, 1)
    assert lde.shape == (4 << logn, w)
    assert np.array_equal(lde[: 1 << logn], tr[_bitrev_perm(logn)])


def test_lde_linearity(device):
    logn, w = 20, 16
    a = O.rand_matrix(1, 1 << logn, w)
    b = O.rand_matrix(2, 1 << logn, w)
    s = a + b  # wraps mod 2^64 where a + b >= 2^64
    wrap = s < a
    s = np.where(wrap | (s >= np.uint64(P)), s - np.uint64(P), s)
    dft = G.Dft(device)
    la, lb, ls = (dft.coset_lde_batch(x, 1, 7) for x in (a, b, s))
    t = la + lb
    t = np.where((t < la) | (t >= np.uint64(P)), t - np.uint64(P), t)
    assert np.array_equal(t, ls)


def test_batched_plan_equals_per_trace_path_c3(device):
    shapes = W.c3_shapes()
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    plan = G.CommitPlan(cfg, shapes, 1, graph=True, autotune=True)
    plan.fill_random(2025)
    inputs = [G.DeviceMatrix(device, plan.input_ptr(i), h, w, owned=False) for i, (h, w) in enumerate(shapes)]
    plain = G.commit_traces(cfg, inputs, 1)
    batched = plan.run()
    assert batched.root().tolist() == plain.root().tolist(), plan.info()
    H = batched.height()
    assert np.array_equal(batched.layer(12), plain.layer(12))
    for idx in (0, H // 3, H - 1):
        assert batched.rows(idx).tolist() == plain.rows(idx).tolist()
    batched.free()
    plain.free()
    params = W.c3_params()
    root, z, f, c = plan.prove(G.PcsParams(*params))
    logL = H.bit_length() - 1
    assert O.pcs_trace_verify(O.PcsParams(*params), O.HASH_P2, root, [w for _, w in shapes], logL, f, c) == 0


def test_quotient_commit_contains_chunk_values_full_size(device):
    # quotient of the 2^22-row trace at constraint degree 4, log_blowup 2.  Chunk t's
    # input column holds q_t on g w_J^t H, and g w_J^t w_H^r = g w_K^((B/D) t + B r)
    # lies in gK: the committed matrix (rows bit-reversed) repeats every input value
    logn, d, logb = 22, 4, 2
    n, B = 1 << logn, 1 << logb
    q = device.alloc_matrix(n, 2 * d)
    q.fill_random(31)
    qh = q.download()
    t = G.commit_quotient(G.LmcsConfig(device, G.HASH_POSEIDON2), q, d, logb)
    ptr, h, w = t.matrix(0)
    assert (h, w) == (n * B, 2 * d)
    lde = G.DeviceMatrix(device, ptr, h, w, owned=False).download()
    rows = _bitrev_perm(logn + logb)  # natural index -> stored row is bitrev
    r = np.arange(n, dtype=np.int64)
    for ti in range(d):
        nat = (B // d) * ti + B * r
        got = lde[rows[nat]][:, 2 * ti:2 * ti + 2]
        assert np.array_equal(got, qh[:, 2 * ti:2 * ti + 2]), ti
    t.free()
    q.free()
