"""The wide bit-structured Keccak workload at its real shape: 2^16 rows x 2600
columns (about 90% of the columns 0/1 bits, the rest 16-bit limbs), log_blowup 1,
Keccak-f LMCS for leaves and nodes, FRI arity 2 down to degree 2^3, 100 queries.

This is the only shape that selects the 2^17-leaf Keccak leaf launch, the DEEP
quotient's split over 163 column tiles (deep_quotient_partial_kernel plus its
finishing pass) and the two-pass 2600-column NTT, so it is proven here in full:

  - the proof is compared word for word with the oracle's proof of the same
    trace (the oracle proves it in tens of seconds on the host cores);
  - the oracle's verifier accepts it and rejects it after one tampered word;
  - a 2^10 x 2600 trace of the same structure is compared bit for bit as well,
    so a failure at full size can be told apart from one at the width.

Workload shape: p3-miden-lmcs/benches/utils.rs:46-53,92-113 (the Goldilocks
Keccak LMCS), with the trace generator of p3gpu.workloads.c4_trace.
"""
import numpy as np
import pytest

import oracle as O
import p3gpu as G
from p3gpu import workloads as W

pytestmark = pytest.mark.gpu

P = (1 << 64) - (1 << 32) + 1


def _prove_both(device, log_rows):
    tr = W.c4_trace(2025, log_rows)
    params = W.C4_PARAMS
    print(f"c4 proof: 2^{log_rows} x {W.C4_WIDTH}, params {params}", flush=True)
    root_g, z_g, f_g, c_g = G.prove(G.LmcsConfig(device, G.HASH_KECCAK), G.PcsParams(*params), [tr])
    root_o, z_o, f_o, c_o, _ = O.pcs_trace_prove(O.PcsParams(*params), O.HASH_KECCAK, [tr])
    return params, (root_g, z_g, f_g, c_g), (root_o, z_o, f_o, c_o)


@pytest.mark.parametrize("log_rows", [10, W.C4_LOG_ROWS])
def test_c4_proof_bit_exact_and_verified(device, log_rows):
    params, (root_g, z_g, f_g, c_g), (root_o, z_o, f_o, c_o) = _prove_both(device, log_rows)
    assert root_g.tolist() == root_o.tolist()
    assert [tuple(int(v) for v in p) for p in z_g] == [tuple(int(v) for v in p) for p in z_o]
    assert f_g.shape == f_o.shape and np.array_equal(f_g, f_o)
    assert c_g.shape == c_o.shape and np.array_equal(c_g, c_o)
    logL = log_rows + params[0]
    op = O.PcsParams(*params)
    assert O.pcs_trace_verify(op, O.HASH_KECCAK, root_g, [W.C4_WIDTH], logL, f_g, c_g) == 0
    bad = f_g.copy()
    k = len(bad) // 2
    bad[k] = (int(bad[k]) + 1) % P
    assert O.pcs_trace_verify(op, O.HASH_KECCAK, root_g, [W.C4_WIDTH], logL, bad, c_g) != 0
    badc = c_g.copy()
    badc[len(badc) // 2, 0] = (int(badc[len(badc) // 2, 0]) + 1) % P
    assert O.pcs_trace_verify(op, O.HASH_KECCAK, root_g, [W.C4_WIDTH], logL, f_g, badc) != 0
