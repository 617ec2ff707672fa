"""Writes tests/golden/*.json.

reference_kats.json holds the known-answer values stated in the reference's
own tests (transcribed as data, each with its source file:line) and published
vectors of the primitives the reference relies on.  regression.json holds
digests of oracle outputs for fixed seeds (the C1 proof and smaller cases) so
later changes to the oracle or the GPU path that alter a single output word
are caught without recomputing anything expensive.

Run from the repository root: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402

KATS = {
    "stateful_sponge_mock": {
        "source": "p3-miden-stateful-hasher/src/field_sponge.rs:80-107",
        "width": 4, "rate": 2, "input": [1, 2, 3, 4, 5], "output": [327, 327],
    },
    "compute_field_alignment": {
        "source": "p3-miden-stateful-hasher/src/serializing_sponge.rs:276-318",
        "cases": [[4, 4, 8, 8], [4, 8, 4, 8], [8, 4, 8, 4], [8, 8, 4, 4], [32, 8, 2, 1],
                  [4, 8, 3, 6], [8, 4, 3, 3], [4, 4, 7, 7], [8, 8, 5, 5]],
        "columns": ["field_bytes", "word_bytes", "rate", "alignment"],
    },
    "deep_neg_coeffs_horner": {
        "source": "p3-miden-lifted-fri/src/deep/prover.rs:379-451",
        "cases": [
            {"c": 2, "alignment": 3, "rows": [[1, 2], [3, 4, 5]]},
            {"c": 7, "alignment": 4, "rows": [[10, 20, 30], [1, 2, 3, 4, 5], [100, 200]]},
        ],
    },
    "fri_final_polynomial": {
        "source": "p3-miden-lifted-fri/src/fri/tests.rs:343-401",
        "log_poly_degree": 6, "log_blowup": 2, "log_final_degree": 3, "log_arity": 1,
        "statement": "f(x) = g(x^8) folded 3 times leaves g; final poly sent in descending degree order",
    },
    "xoshiro256pp_reference": {
        "source": "published xoshiro256++ reference output (state 1,2,3,4)",
        "state": [1, 2, 3, 4],
        "outputs": [41943041, 58720359, 3588806011781223, 3591011842654386, 9228616714210784205,
                    9973669472204895162, 14011001112246962877, 12406186145184390807,
                    15849039046786891736, 10450023813501588000],
    },
    "keccak256_empty": {
        "source": "Keccak-256 (original 0x01 padding) of the empty string",
        "digest": "c5d2460186f7233c927e7db2dcc703c0e500b653ca82273b7bfad8045d85a470",
    },
    "goldilocks_two_adic_generator_32": {
        "source": "7^((p-1)/2^32)", "value": 1753635133440165772,
    },
}


def proof_digest(fields, comms):
    b = np.asarray(fields, np.uint64).astype("<u8").tobytes() + np.asarray(comms, np.uint64).astype("<u8").tobytes()
    return hashlib.sha256(b).hexdigest()


REGRESSION_CASES = [
    # name, hash kind, shapes, PcsParams args, seed
    ("c1_poseidon2_2^14x32", O.HASH_P2, [(1 << 14, 32)], (1, 1, 3, 0, 0, 64, 0), 2025),
    ("multi_trace_poseidon2", O.HASH_P2, [(8, 2), (64, 11), (256, 7), (256, 30)], (1, 1, 1, 0, 0, 40, 0), 2025),
    ("keccak_grinding", O.HASH_KECCAK, [(32, 18), (128, 17)], (2, 2, 1, 0, 1, 25, 2), 2025),
    ("keccak_wide_2^4x2600", O.HASH_KECCAK, [(16, 2600)], (1, 1, 2, 0, 0, 20, 0), 77),
]


def regression():
    out = {}
    for name, kind, shapes, params, seed in REGRESSION_CASES:
        tr = [O.rand_matrix(seed + i, h, w) for i, (h, w) in enumerate(shapes)]
        root, z, f, c, _ = O.pcs_trace_prove(O.PcsParams(*params), kind, tr)
        out[name] = {"kind": kind, "shapes": shapes, "params": list(params), "seed": seed,
                     "root": [int(x) for x in root], "eval_points": [list(map(int, p)) for p in z],
                     "num_fields": int(f.size), "num_commitments": int(c.shape[0]),
                     "proof_sha256": proof_digest(f, c)}
    return out


if __name__ == "__main__":
    with open(os.path.join(HERE, "reference_kats.json"), "w") as fh:
        json.dump(KATS, fh, indent=1)
    with open(os.path.join(HERE, "regression.json"), "w") as fh:
        json.dump(regression(), fh, indent=1)
    print("wrote", os.listdir(HERE))
