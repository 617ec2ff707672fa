"""Proof-of-work grinding with the witness search on the GPU.

The reference grinds with DuplexChallenger::grind (p3-miden-transcript/src/
prover.rs:140-144): the smallest witness i with observe(i); sample_bits(b) == 0.
It grinds before the DEEP challenges (deep/prover.rs:156), in every FRI commit
round between the root and beta (fri/prover.rs:149-156), and before the query
indices (lifted-fri/src/prover.rs:85).  Here the search runs on the device (and
inside the device Fiat-Shamir rounds for folding grinding); each proof must be
word-for-word the oracle's, whose grind is the serial search, and the oracle's
verifier must accept it.  The reference's end-to-end Goldilocks binary uses
PcsParams(3, 1, 0, 16, 0, 100, 0) (p3-miden-lifted-examples/src/bin/
lifted_miden.rs:34-54); that set runs here on smaller traces.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import p3gpu as G

pytestmark = pytest.mark.gpu

CASES = [
    # (shapes, params): log_blowup, log_arity, log_final_degree, folding, deep, queries, query pow
    ([(1 << 8, 8)], (1, 1, 2, 8, 6, 20, 10)),
    ([(1 << 10, 16)], (2, 2, 1, 12, 12, 30, 12)),
    ([(1 << 6, 5), (1 << 9, 24)], (1, 3, 0, 10, 0, 16, 4)),
    ([(1 << 9, 51), (1 << 10, 20)], (3, 1, 0, 16, 0, 100, 0)),  # the reference binary's parameters
]


def _traces(shapes, seed):
    return [O.rand_matrix(seed + i, h, w) for i, (h, w) in enumerate(shapes)]


@pytest.mark.parametrize("shapes,params", CASES)
def test_prove_with_grinding_matches_oracle(device, shapes, params):
    tr = _traces(shapes, 40)
    print(f"grinding proof: shapes {shapes} params {params}", flush=True)
    root_g, z_g, f_g, c_g = G.prove(G.LmcsConfig(device, G.HASH_POSEIDON2), G.PcsParams(*params), tr)
    root_o, z_o, f_o, c_o, _ = O.pcs_trace_prove(O.PcsParams(*params), O.HASH_P2, tr)
    assert root_g.tolist() == root_o.tolist()
    assert np.array_equal(f_g, f_o) and np.array_equal(c_g, c_o)
    logL = max(h for h, _ in shapes).bit_length() - 1 + params[0]
    op = O.PcsParams(*params)
    assert O.pcs_trace_verify(op, O.HASH_P2, root_g, [w for _, w in shapes], logL, f_g, c_g) == 0


@pytest.mark.parametrize("shapes,params", CASES[:3])
def test_prove_many_with_grinding_matches_oracle(device, shapes, params):
    jobs = [_traces(shapes, 70 + 5 * j) for j in range(5)]
    print(f"grinding lock-step batch: 5 jobs of {shapes} params {params}", flush=True)
    outs = G.prove_many(G.LmcsConfig(device, G.HASH_POSEIDON2), G.PcsParams(*params), jobs)
    for tr, (root_g, z_g, f_g, c_g) in zip(jobs, outs):
        root_o, z_o, f_o, c_o, _ = O.pcs_trace_prove(O.PcsParams(*params), O.HASH_P2, tr)
        assert root_g.tolist() == root_o.tolist()
        assert np.array_equal(f_g, f_o) and np.array_equal(c_g, c_o)


def test_grinding_witnesses_are_the_smallest(device):
    # every witness in the proof passes the check and no smaller candidate does (brute force, oracle side)
    shapes, params = [(1 << 8, 8)], (1, 1, 2, 8, 6, 20, 10)
    tr = _traces(shapes, 40)
    root_g, z_g, f_g, c_g = G.prove(G.LmcsConfig(device, G.HASH_POSEIDON2), G.PcsParams(*params), tr)
    tx = O.Transcript(O.CH_DUPLEX_P2)
    tx.observe_commitment(root_g)
    tx.sample_ef()
    tx.sample_ef()
    # DEEP grind follows the two points' aligned evaluations (2 * 2 * 8 words)
    evals = f_g[:32]
    tx.send_fields(evals)
    w = int(f_g[32])
    for cand in range(w):
        assert not tx.clone().check_witness(6, cand), cand
    assert tx.clone().check_witness(6, w)


def test_pow_bits_beyond_the_bound_are_rejected_before_committing(device):
    tr = O.rand_matrix(1, 1 << 6, 4)
    ch = G.ProverTranscript()
    m = G._Matrix(tr.ctypes.data_as(G._u64p), tr.shape[0], tr.shape[1])
    root = np.zeros(4, np.uint64)
    z = np.zeros(4, np.uint64)
    for params in ((1, 1, 0, 32, 0, 4, 0), (1, 1, 0, 0, 40, 4, 0), (1, 1, 0, 0, 0, 4, 63)):
        p = G.PcsParams.__new__(G.PcsParams)
        C.Structure.__init__(p, *params)
        rc = G.lib().p3g_prove(device.h, G.HASH_POSEIDON2, C.byref(p), C.byref(m), 1, 0, ch.h, G._p(root), G._p(z))
        assert rc == 1, params  # P3G_E_INVAL
    f, c = ch.data()
    assert f.size == 0 and c.size == 0  # the transcript was not touched
