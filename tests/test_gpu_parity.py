"""GPU parity: the HIP library, called through the C ABI, against the CPU
oracle on the same seeded inputs.  Bit-exact (integer field arithmetic).

Sizes stay where the oracle finishes in seconds; the full-size configuration
is covered in test_gpu_full_size.py through size-independent properties.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
import p3gpu as G

pytestmark = pytest.mark.gpu


def _mats(seed, shapes):
    return [O.rand_matrix(seed + i, h, w) for i, (h, w) in enumerate(shapes)]


# ---------------------------------------------------------------- coset LDE
@pytest.mark.parametrize("logn,w,bits", [(0, 1, 1), (0, 5, 3), (1, 1, 1), (2, 3, 2), (3, 16, 1), (5, 17, 2),
                                         (8, 4, 1), (9, 33, 1), (10, 64, 2), (12, 8, 3), (14, 32, 1), (16, 3, 2),
                                         # two-pass transforms with 9..11-stage passes (8- and 4-column tiles)
                                         (17, 8, 1), (18, 12, 2), (20, 16, 1), (21, 9, 1), (22, 8, 2)])
def test_coset_lde_matches_oracle(device, logn, w, bits):
    tr = O.rand_matrix(1000 + logn, 1 << logn, w)
    shift = O.gl_pow(7, 1 + logn % 3)
    want = O.coset_lde_batch(tr, bits, shift)
    got = G.Dft(device).coset_lde_batch(tr, bits, shift)
    assert np.array_equal(got, want)


def _lde_sweep_cases(n=24, seed=77):
    """Seeded shapes across the register-tile passes: 5..16 stages (so every pass
    size 5..8 and both first-pass modes), widths around the 16-column tile edges
    (masked columns), blowups 2..8."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        logn = int(rng.integers(5, 17))
        w = int(rng.choice([8, 9, 11, 15, 16, 17, 23, 24, 31, 33, 40]))
        bits = int(rng.integers(1, 4))
        out.append((logn, w, bits))
    return out


@pytest.mark.parametrize("logn,w,bits", _lde_sweep_cases())
def test_coset_lde_sweep_matches_oracle(device, logn, w, bits):
    tr = O.rand_matrix(3000 + 31 * logn + w, 1 << logn, w)
    shift = O.gl_pow(7, 1 + (logn + w) % 5)
    assert np.array_equal(G.Dft(device).coset_lde_batch(tr, bits, shift), O.coset_lde_batch(tr, bits, shift))


@pytest.mark.parametrize("logn,w,bits", [(7, 16, 1), (12, 17, 2), (15, 8, 3)])
def test_coset_lde_edge_values_match_oracle(device, logn, w, bits):
    """Columns of p - 1, of 0 / p - 1 alternating, of 2^k and of p - 2^k: butterfly
    sums landing on p and above, differences borrowing from zero, shift twiddles
    of words with the top bits set."""
    n = 1 << logn
    P = (1 << 64) - (1 << 32) + 1
    r = np.arange(n, dtype=np.uint64)
    cols = [np.full(n, P - 1, np.uint64), np.where(r % 2 == 0, 0, P - 1).astype(np.uint64),
            (np.uint64(1) << (r % np.uint64(64))) % np.uint64(P),
            (np.uint64(P) - (np.uint64(1) << (r % np.uint64(63)))).astype(np.uint64),
            np.where(r % 3 == 0, P - 1, 1).astype(np.uint64)]
    tr = np.stack([cols[c % len(cols)] if c < 2 * len(cols) else O.rand_matrix(c, n, 1)[:, 0]
                   for c in range(w)], axis=1).astype(np.uint64)
    assert (tr < np.uint64(P)).all()
    shift = O.gl_pow(7, 2)
    assert np.array_equal(G.Dft(device).coset_lde_batch(tr, bits, shift), O.coset_lde_batch(tr, bits, shift))


def test_coset_lde_natural_order(device):
    tr = O.rand_matrix(5, 256, 7)
    want = O.coset_lde_batch(tr, 2, 7, bit_reverse=False)
    got = G.Dft(device).coset_lde_batch(tr, 2, 7, bit_reverse_rows=False)
    assert np.array_equal(got, want)


# ---------------------------------------------------------------- LMCS
SCEN_P2 = [[(1, 1)], [(1, 7)], [(2, 3), (4, 5), (8, 8)], [(1, 5), (1, 3), (2, 7), (4, 1), (8, 9)],
           [(2, 7), (4, 8), (8, 11)], [(4, 13), (8, 25)], [(1, 16), (2, 15), (4, 16), (8, 22)],
           [(4, 7), (4, 8), (8, 11), (8, 16)], [(8, 7)], [(1024, 64)], [(64, 3), (4096, 40), (4096, 1)]]
SCEN_KECCAK = [[(1, 1)], [(2, 16), (4, 17), (8, 18)], [(4, 34), (8, 35)], [(512, 100)], [(16, 2600)]]


@pytest.mark.parametrize("kind,scen", [(G.HASH_POSEIDON2, s) for s in SCEN_P2] + [(G.HASH_KECCAK, s) for s in SCEN_KECCAK])
def test_lmcs_tree_matches_oracle(device, kind, scen):
    mats = _mats(77, scen)
    want = O.Tree(kind, mats, alignment=O.hash_alignment(kind))
    cfg = G.LmcsConfig(device, kind)
    got = cfg.build_aligned_tree(mats)
    assert got.root().tolist() == want.root().tolist()
    H = got.height()
    for l in range(H.bit_length()):
        assert np.array_equal(got.layer(l), want.layer(l)), l
    for idx in {0, H - 1, H // 3}:
        assert got.rows(idx).tolist() == want.rows(idx).tolist()
    assert got.widths() == [O.aligned_len(w, O.hash_alignment(kind)) for _, w in scen]
    # prove_batch hints: identical transcript streams
    idx = [H - 1, 0, H // 2, H // 2, H // 5]
    gt = G.ProverTranscript(0 if kind == G.HASH_POSEIDON2 else 1)
    ot = O.Transcript(0 if kind == G.HASH_POSEIDON2 else 1)
    got.prove_batch(idx, gt)
    want.prove_batch(idx, ot)
    gf, gc = gt.data()
    of, oc = ot.data()
    assert np.array_equal(gf, of) and np.array_equal(gc, oc)


def test_lmcs_unaligned_and_errors(device):
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    mats = _mats(3, [(4, 3), (8, 5)])
    t = cfg.build_tree(mats)
    assert t.root().tolist() == O.Tree(O.HASH_P2, mats).root().tolist()
    assert t.widths() == [3, 5]
    with pytest.raises(G.P3GError) as e:
        cfg.build_tree(_mats(3, [(8, 3), (4, 5)]))
    assert e.value.code == 3  # unsorted
    with pytest.raises(G.P3GError) as e:
        cfg.build_tree(_mats(3, [(6, 3)]))
    assert e.value.code == 2  # not a power of two
    with pytest.raises(G.P3GError) as e:
        cfg.build_tree([])
    assert e.value.code == 6  # empty
    with pytest.raises(G.P3GError):
        t.rows(8)


def test_device_resident_inputs(device):
    mats = _mats(11, [(256, 24), (1024, 9)])
    dms = [device.upload(m) for m in mats]
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    t = cfg.build_aligned_tree(dms)
    assert t.root().tolist() == O.Tree(O.HASH_P2, mats, alignment=8).root().tolist()


# ---------------------------------------------------------------- commit_traces
@pytest.mark.parametrize("kind,shapes,lb", [
    (G.HASH_POSEIDON2, [(1 << 10, 8)], 1),
    (G.HASH_POSEIDON2, [(16, 3), (256, 10), (1024, 50), (1024, 16)], 2),
    (G.HASH_POSEIDON2, [(2, 1), (8, 20)], 3),
    (G.HASH_KECCAK, [(64, 40), (256, 17)], 1),
])
def test_commit_traces_matches_oracle(device, kind, shapes, lb):
    tr = _mats(31, shapes)
    want = O.commit_traces(kind, tr, lb)
    got = G.commit_traces(G.LmcsConfig(device, kind), tr, lb)
    assert got.root().tolist() == want.root().tolist()
    H = got.height()
    assert got.rows(H - 3).tolist() == want.rows(H - 3).tolist()


# ---------------------------------------------------------------- full PCS opening
CASES = [
    # (kind, shapes, PcsParams args)  — C1 first: 2^14 x 32, blowup 1, arity 2, final degree 8, 64 queries
    (G.HASH_POSEIDON2, [(1 << 14, 32)], (1, 1, 3, 0, 0, 64, 0)),
    (G.HASH_POSEIDON2, [(1 << 10, 16)], (2, 2, 4, 0, 0, 100, 0)),
    (G.HASH_POSEIDON2, [(1 << 9, 9)], (1, 3, 2, 0, 0, 30, 0)),
    (G.HASH_POSEIDON2, [(8, 2), (64, 11), (256, 7), (256, 30)], (1, 1, 1, 0, 0, 40, 0)),
    (G.HASH_POSEIDON2, [(1 << 8, 5)], (3, 2, 0, 1, 2, 20, 3)),
    (G.HASH_KECCAK, [(1 << 9, 40)], (1, 1, 2, 0, 0, 50, 0)),
    (G.HASH_KECCAK, [(32, 18), (128, 17)], (2, 2, 1, 0, 1, 25, 2)),
    # zero FRI rounds: the final polynomial is the whole DEEP polynomial (fri/tests.rs:294-340);
    # final degree equal to and above the trace degree
    (G.HASH_POSEIDON2, [(1 << 4, 6)], (1, 1, 4, 0, 0, 8, 0)),
    (G.HASH_POSEIDON2, [(1 << 4, 6)], (1, 1, 6, 0, 0, 8, 0)),
    (G.HASH_KECCAK, [(1 << 4, 6)], (2, 2, 5, 0, 0, 10, 0)),
]


@pytest.mark.parametrize("kind,shapes,params", CASES)
def test_prove_matches_oracle_bit_exact(device, kind, shapes, params):
    tr = _mats(2025, shapes)
    op = O.PcsParams(*params)
    root_o, z_o, f_o, c_o, tree_o = O.pcs_trace_prove(op, kind, tr)
    gp = G.PcsParams(*params)
    root_g, z_g, f_g, c_g = G.prove(G.LmcsConfig(device, kind), gp, tr)
    assert root_g.tolist() == root_o.tolist()
    assert z_g == tuple(z_o)
    assert f_g.size == f_o.size and c_g.shape == c_o.shape
    assert np.array_equal(f_g, f_o)
    assert np.array_equal(c_g, c_o)
    assert G.proof_bytes(f_g, c_g) == G.proof_bytes(f_o, c_o)
    logL = tree_o.height().bit_length() - 1
    assert O.pcs_trace_verify(op, kind, root_g, [w for _, w in shapes], logL, f_g, c_g) == 0


def test_open_with_channel_multi_tree_and_points(device):
    # two commitment groups (main + aux style) opened at three points
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    tr_a = _mats(5, [(64, 6), (128, 10)])
    tr_b = _mats(9, [(128, 3)])
    ga, gb = G.commit_traces(cfg, tr_a, 1), G.commit_traces(cfg, tr_b, 1)
    oa, ob = O.commit_traces(O.HASH_P2, tr_a, 1), O.commit_traces(O.HASH_P2, tr_b, 1)
    z = [(11, 12), (13, 14), (15, 16)]
    params = (1, 2, 1, 0, 0, 17, 0)
    gt, ot = G.ProverTranscript(), O.Transcript()
    G.open_with_channel(G.PcsParams(*params), cfg, 8, z, [ga, gb], gt)
    O.open_with_channel(O.PcsParams(*params), O.HASH_P2, [oa, ob], 8, z, ot)
    gf, gc = gt.data()
    of, oc = ot.data()
    assert np.array_equal(gf, of) and np.array_equal(gc, oc)
    v = O.Transcript()
    v.load(gf, gc)
    rc = O.verify(O.PcsParams(*params), O.HASH_P2, [ga.root(), gb.root()], [[8, 16], [8]], 8, z, v)
    assert rc == 0


def test_repeated_proofs_are_identical(device):
    # the allocation cache recycles buffers between proofs: results must not depend on it
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    tr = _mats(44, [(1 << 11, 12)])
    p = G.PcsParams(1, 2, 2, 0, 0, 30, 0)
    a = G.prove(cfg, p, tr)
    b = G.prove(cfg, p, tr)
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])


class _OracleChannel:
    """A caller-owned prover channel (the oracle's transcript) in ProverChannel vocabulary."""

    def __init__(self, tx):
        self.tx = tx

    def send_field_slice(self, v): self.tx.send_fields(v)
    def send_commitment(self, c): self.tx.send_commitment(c)
    def hint_field_slice(self, v): self.tx.hint_fields(v)
    def hint_commitment(self, c): self.tx.hint_commitment(c)
    def grind(self, bits): return self.tx.grind(bits)
    def sample(self): return self.tx.sample()
    def sample_bits(self, bits): return self.tx.sample_bits(bits)


@pytest.mark.parametrize("kind", [G.HASH_POSEIDON2, G.HASH_KECCAK])
def test_open_through_caller_channel_callbacks(device, kind):
    # the drop-in path: the caller's own channel is driven by the library's callbacks
    cfg = G.LmcsConfig(device, kind)
    ch_kind = 0 if kind == G.HASH_POSEIDON2 else 1
    tr = _mats(21, [(128, 9), (256, 20)])
    params = (1, 2, 2, 1, 2, 20, 3)
    gt = G.commit_traces(cfg, tr, 1)
    ot = O.commit_traces(kind, tr, 1)
    z = [(101, 202), (303, 404)]
    via_cb = O.Transcript(ch_kind)
    G.open_with_channel(G.PcsParams(*params), cfg, 9, z, [gt], _OracleChannel(via_cb))
    want = O.Transcript(ch_kind)
    O.open_with_channel(O.PcsParams(*params), kind, [ot], 9, z, want)
    a, b = via_cb.data(), want.data()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # a failing callback aborts with a transcript status
    class Broken(_OracleChannel):
        def sample(self):
            raise RuntimeError("channel closed")
    with pytest.raises(G.P3GError) as e:
        G.open_with_channel(G.PcsParams(*params), cfg, 9, z, [gt], Broken(O.Transcript(ch_kind)))
    assert e.value.code == 7


# ---------------------------------------------------------------- batched commit plans
PLAN_CASES = [
    (G.HASH_POSEIDON2, [(1 << 10, 8)], 1, True, False),
    (G.HASH_POSEIDON2, [(16, 3), (256, 10), (1024, 50), (1 << 13, 17)], 2, True, "per_trace"),
    (G.HASH_POSEIDON2, [(16, 3), (256, 10), (1024, 50), (1024, 16)], 2, True, True),
    (G.HASH_POSEIDON2, [(2, 1), (8, 20), (4096, 33)], 1, False, False),
    (G.HASH_POSEIDON2, [(1 << 12, 8), (1 << 13, 100), (1 << 14, 17)], 1, True, True),
    (G.HASH_KECCAK, [(64, 40), (256, 17), (256, 2)], 1, True, False),
    (G.HASH_POSEIDON2, [(1 << 9, 5), (1 << 17, 8), (1 << 18, 12)], 1, True, True),
    (G.HASH_POSEIDON2, [(1 << 17, 8), (1 << 18, 12)], 1, False, "per_trace"),
]


@pytest.mark.parametrize("kind,shapes,lb,graph,autotune", PLAN_CASES)
def test_commit_plan_matches_oracle(device, kind, shapes, lb, graph, autotune):
    cfg = G.LmcsConfig(device, kind)
    plan = G.CommitPlan(cfg, shapes, lb, graph=graph, autotune=autotune is True, per_trace=autotune == "per_trace")
    if autotune == "per_trace":
        assert plan.info()["grouped"] is False
    for rnd in range(2):  # the second run replays the same plan on new inputs
        tr = _mats(500 + 17 * rnd, shapes)
        for i, m in enumerate(tr):
            plan.set_input(i, m)
        got = plan.run()
        want = O.commit_traces(kind, tr, lb)
        assert got.root().tolist() == want.root().tolist(), (rnd, plan.info())
        H = got.height()
        for l in (0, 1, H.bit_length() - 2):
            assert np.array_equal(got.layer(l), want.layer(l)), (rnd, l)
        assert got.rows(H - 3).tolist() == want.rows(H - 3).tolist()
        got.free()


@pytest.mark.parametrize("kind,shapes,params", [
    (G.HASH_POSEIDON2, [(8, 2), (64, 11), (256, 7), (256, 30)], (1, 1, 1, 0, 0, 40, 0)),
    (G.HASH_POSEIDON2, [(1 << 10, 16)], (2, 2, 4, 0, 0, 100, 0)),
    (G.HASH_KECCAK, [(32, 18), (128, 17)], (2, 2, 1, 0, 1, 25, 2)),
])
def test_prove_planned_matches_oracle(device, kind, shapes, params):
    tr = _mats(2025, shapes)
    op = O.PcsParams(*params)
    root_o, z_o, f_o, c_o, _ = O.pcs_trace_prove(op, kind, tr)
    plan = G.CommitPlan(G.LmcsConfig(device, kind), shapes, params[0], graph=True)
    for i, m in enumerate(tr):
        plan.set_input(i, m)
    for _ in range(2):  # graph replays give the same proof
        root_g, z_g, f_g, c_g = plan.prove(G.PcsParams(*params))
        assert root_g.tolist() == root_o.tolist() and z_g == tuple(z_o)
        assert np.array_equal(f_g, f_o) and np.array_equal(c_g, c_o)


def test_commit_plan_rejects_bad_shapes(device):
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    for shapes, code in [([], 6), ([(8, 3), (4, 5)], 3), ([(6, 3)], 2), ([(8, 0)], 1)]:
        with pytest.raises(G.P3GError) as e:
            G.CommitPlan(cfg, shapes, 1)
        assert e.value.code == code, shapes
    plan = G.CommitPlan(cfg, [(8, 3)], 1)
    with pytest.raises(G.P3GError):
        plan.input_ptr(1)
    with pytest.raises(G.P3GError):
        plan.prove(G.PcsParams(2, 1, 0, 0, 0, 4, 0))  # log_blowup differs from the plan's


def _splitmix_fill(seed, n):
    golden, p, m = 0x9E3779B97F4A7C15, (1 << 64) - (1 << 32) + 1, (1 << 64) - 1
    out = []
    for i in range(n):
        z = (seed + (i + 1) * golden) & m
        while True:
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
            z ^= z >> 31
            if z < p:
                break
            z = (z + golden) & m
        out.append(z)
    return out


def test_fill_random_device(device):
    d = device.upload(np.zeros((1, 1000), np.uint64))
    G.fill_random_device(device, 99, 1000, d.ptr)
    got = d.download().ravel().tolist()
    assert got == _splitmix_fill(99, 1000)


# ---------------------------------------------------------------- committed regression digests
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "regression.json")) as _fh:
    _REG = json.load(_fh)


@pytest.mark.parametrize("name", sorted(_REG))
def test_gpu_proofs_match_golden_digests(device, name):
    import hashlib
    r = _REG[name]
    tr = [O.rand_matrix(r["seed"] + i, h, w) for i, (h, w) in enumerate(r["shapes"])]
    root, z, f, c = G.prove(G.LmcsConfig(device, r["kind"]), G.PcsParams(*r["params"]), tr)
    assert [int(x) for x in root] == r["root"]
    assert [list(p) for p in z] == r["eval_points"]
    b = f.astype("<u8").tobytes() + c.astype("<u8").ravel().tobytes()
    assert hashlib.sha256(b).hexdigest() == r["proof_sha256"]


# ---------------------------------------------------------------- batched lock-step proofs
MANY_CASES = [
    (G.HASH_POSEIDON2, [(1 << 10, 16)], (2, 2, 4, 0, 0, 100, 0), 5),
    (G.HASH_POSEIDON2, [(8, 2), (64, 11), (256, 7), (256, 30)], (1, 1, 1, 0, 0, 40, 0), 3),
    (G.HASH_POSEIDON2, [(1 << 14, 32)], (1, 1, 3, 0, 0, 64, 0), 4),
    (G.HASH_POSEIDON2, [(1 << 8, 5)], (3, 2, 0, 1, 2, 20, 3), 6),
    (G.HASH_POSEIDON2, [(1 << 9, 9)], (1, 3, 2, 0, 0, 30, 0), 1),
    (G.HASH_KECCAK, [(32, 18), (128, 17)], (2, 2, 1, 0, 1, 25, 2), 3),
    (G.HASH_KECCAK, [(1 << 9, 40)], (1, 1, 2, 0, 0, 50, 0), 2),
    (G.HASH_POSEIDON2, [(1 << 17, 8)], (1, 1, 3, 0, 0, 20, 0), 2),
    (G.HASH_POSEIDON2, [(1 << 4, 6)], (1, 1, 4, 0, 0, 8, 0), 3),  # zero FRI rounds
]


@pytest.mark.parametrize("kind,shapes,params,k", MANY_CASES)
def test_prove_many_matches_oracle(device, kind, shapes, params, k):
    jobs = [_mats(3000 + 10 * j, shapes) for j in range(k)]
    got = G.prove_many(G.LmcsConfig(device, kind), G.PcsParams(*params), jobs)
    assert len(got) == k
    for tr, (root, z, f, c) in zip(jobs, got):
        root_o, z_o, f_o, c_o, _ = O.pcs_trace_prove(O.PcsParams(*params), kind, tr)
        assert root.tolist() == root_o.tolist() and z == tuple(z_o)
        assert np.array_equal(f, f_o) and np.array_equal(c, c_o)


def test_prove_many_device_inputs_and_errors(device):
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    params = G.PcsParams(1, 2, 1, 0, 0, 16, 0)
    host = [_mats(70 + j, [(256, 12)]) for j in range(4)]
    jobs = [[device.upload(m) for m in h] for h in host]
    got = G.prove_many(cfg, params, jobs)
    for h, (root, z, f, c) in zip(host, got):
        r1, z1, f1, c1 = G.prove(cfg, params, h)
        assert root.tolist() == r1.tolist() and np.array_equal(f, f1) and np.array_equal(c, c1)
    with pytest.raises(G.P3GError) as e:
        G.prove_many(cfg, params, [_mats(1, [(256, 12)]), _mats(2, [(256, 13)])])
    assert e.value.code == 1  # shapes differ between jobs


# ---------------------------------------------------------------- hiding LMCS (caller-given salt)
@pytest.mark.parametrize("kind,ok,shapes,salt_w", [
    (G.HASH_POSEIDON2, O.HASH_P2, [(8, 4)], 4),
    (G.HASH_POSEIDON2, O.HASH_P2, [(4, 3), (8, 5), (16, 7)], 2),
    (G.HASH_POSEIDON2, O.HASH_P2, [(64, 3), (4096, 40), (4096, 1)], 8),
    (G.HASH_POSEIDON2, O.HASH_P2, [(1 << 15, 16)], 4),
    (G.HASH_KECCAK, O.HASH_KECCAK, [(4, 17), (8, 3)], 4),
    (G.HASH_KECCAK, O.HASH_KECCAK, [(512, 100)], 2),
])
def test_hiding_tree_matches_oracle(device, kind, ok, shapes, salt_w):
    # HidingLmcsConfig::build_aligned_tree with the same salt: root, every layer,
    # salts, and prove_batch hints (rows padded, then the salt) equal the oracle's
    mats = _mats(700 + salt_w, shapes)
    H = shapes[-1][0]
    salt = O.rand_matrix(800 + salt_w, H, salt_w)
    cfg = G.LmcsConfig(device, kind)
    t = cfg.build_aligned_tree(mats, salt=salt)
    want = O.Tree(ok, mats, alignment=O.hash_alignment(ok), salt=salt)
    assert t.root().tolist() == want.root().tolist()
    for lvl in range(0, H.bit_length(), 2):
        assert np.array_equal(t.layer(lvl), want.layer(lvl))
    for i in (0, H // 2, H - 1):
        assert t.salt(i).tolist() == salt[i].tolist()
    idx = [0, H - 1, H // 3, H // 3, 1 % H]
    tx_g, tx_o = G.ProverTranscript(G.CHALLENGER_DUPLEX_P2), O.Transcript(O.CH_DUPLEX_P2)
    t.prove_batch(idx, tx_g)
    want.prove_batch(idx, tx_o)
    fg, cg = tx_g.data()
    fo, co = tx_o.data()
    assert np.array_equal(fg, fo) and np.array_equal(cg, co)
    # the oracle verifier opens the GPU's hints
    v = O.Transcript(O.CH_DUPLEX_P2)
    v.load(fg, cg)
    rc, _ = O.open_batch(ok, t.root(), t.widths(), H.bit_length() - 1, idx, v, salt_width=salt_w)
    assert rc == 0
    t.free()


def test_hiding_tree_device_salt_and_errors(device):
    mats = [O.rand_matrix(900, 1 << 10, 8)]
    salt = O.rand_matrix(901, 1 << 10, 4)
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    dm, ds = device.upload(mats[0]), device.upload(salt)
    t = cfg.build_tree([dm], salt=ds)
    assert t.root().tolist() == O.Tree(O.HASH_P2, mats, salt=salt).root().tolist()
    t.free()
    assert cfg.build_tree(mats).salt(0).size == 0  # a tree without salt
    with pytest.raises(G.P3GError) as e:
        cfg.build_tree(mats, salt=O.rand_matrix(902, 1 << 9, 4))  # salt height below the tree's
    assert e.value.code == 1
    dm.free()
    ds.free()
