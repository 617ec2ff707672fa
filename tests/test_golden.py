"""The oracle against the committed fixtures in tests/golden/: the reference's
own known-answer tests (reference_kats.json) and regression digests of oracle
proofs (regression.json, written by tests/golden/make_golden.py)."""
import ctypes as C
import hashlib
import json
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KATS = json.load(open(os.path.join(HERE, "reference_kats.json")))
REG = json.load(open(os.path.join(HERE, "regression.json")))
P = O.P


def test_stateful_sponge_mock_kat():
    k = KATS["stateful_sponge_mock"]
    O.mock_config(k["width"], k["rate"])
    st = O.sponge_absorb(O.HASH_MOCK, np.zeros(k["width"], np.uint64), k["input"])
    assert st[: len(k["output"])].tolist() == k["output"]


def test_compute_field_alignment_table():
    for fb, wb, rate, want in KATS["compute_field_alignment"]["cases"]:
        assert O.field_alignment(fb, wb, rate) == want, (fb, wb, rate)
    # the two configurations used here
    assert O.hash_alignment(O.HASH_KECCAK) == O.field_alignment(8, 8, 17) == 17
    assert O.hash_alignment(O.HASH_P2) == 8


def _ef(a, b=0):
    return (a % P, b % P)


def _ef_add(x, y):
    return ((x[0] + y[0]) % P, (x[1] + y[1]) % P)


def test_deep_negative_coefficients_equal_negated_horner():
    # coeff_i = -c^(W-1-i) over the aligned widths (the prover's DEEP reduction)
    # against Horner over the zero-padded rows (the verifier's reduction)
    for case in KATS["deep_neg_coeffs_horner"]["cases"]:
        c, a, rows = _ef(case["c"]), case["alignment"], case["rows"]
        padded = [v for r in rows for v in (r + [0] * ((-len(r)) % a))]
        W = len(padded)
        explicit = _ef(0)
        pw = (P - 1, 0)  # -1
        coeffs = [None] * W
        for i in range(W - 1, -1, -1):
            coeffs[i] = pw
            pw = tuple(int(x) for x in O.ef_mul(pw, c))
        for i, v in enumerate(padded):
            explicit = _ef_add(explicit, tuple(int(x) for x in O.ef_mul(coeffs[i], _ef(v))))
        horner = _ef(0)
        for v in padded:
            horner = _ef_add(tuple(int(x) for x in O.ef_mul(horner, c)), _ef(v))
        assert explicit == ((P - horner[0]) % P, (P - horner[1]) % P)


def test_fri_final_polynomial_kat():
    k = KATS["fri_final_polynomial"]
    lpd, lb, lfd = k["log_poly_degree"], k["log_blowup"], k["log_final_degree"]
    rng = np.random.default_rng(123)
    fdeg, pdeg = 1 << lfd, 1 << lpd
    stride = pdeg // fdeg
    g = [(int(rng.integers(0, P, dtype=np.uint64)), int(rng.integers(0, P, dtype=np.uint64))) for _ in range(fdeg)]
    L = pdeg << lb
    coeffs = np.zeros((L, 2), np.uint64)
    for i, (c0, c1) in enumerate(g):
        coeffs[i * stride] = (c0, c1)
    evals = O.bit_reverse_rows(O.naive_coset_dft(coeffs, 1))  # f on the subgroup of order L, bit-reversed
    params = O.PcsParams(lb, k["log_arity"], lfd, 0, 0, 1, 0)
    rounds, final = O.fri_prove(params, O.HASH_P2, evals, lpd + lb, O.Transcript())
    assert rounds == lpd - lfd
    assert [tuple(map(int, r)) for r in final] == list(reversed(g))


def _digest(f, c):
    b = np.asarray(f, np.uint64).astype("<u8").tobytes() + np.asarray(c, np.uint64).astype("<u8").tobytes()
    return hashlib.sha256(b).hexdigest()


@pytest.mark.parametrize("name", sorted(REG))
def test_oracle_proofs_match_regression_digests(name):
    r = REG[name]
    tr = [O.rand_matrix(r["seed"] + i, h, w) for i, (h, w) in enumerate(r["shapes"])]
    root, z, f, c, _ = O.pcs_trace_prove(O.PcsParams(*r["params"]), r["kind"], tr)
    assert [int(x) for x in root] == r["root"]
    assert [list(map(int, p)) for p in z] == r["eval_points"]
    assert (f.size, c.shape[0]) == (r["num_fields"], r["num_commitments"])
    assert _digest(f, c) == r["proof_sha256"]


def test_xoshiro_and_constants_from_fixture():
    k = KATS["xoshiro256pp_reference"]

    class Rng(C.Structure):
        _fields_ = [("s", C.c_uint64 * 4)]

    r = Rng()
    r.s[:] = k["state"]
    O.lib.orc_rng_next.restype = C.c_uint64
    O.lib.orc_rng_next.argtypes = [C.POINTER(Rng)]
    assert [int(O.lib.orc_rng_next(C.byref(r))) for _ in k["outputs"]] == k["outputs"]
    assert O.keccak256(b"").hex() == KATS["keccak256_empty"]["digest"]
    assert O.two_adic_generator(32) == KATS["goldilocks_two_adic_generator_32"]["value"]
