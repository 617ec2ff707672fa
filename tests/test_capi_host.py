"""CPU-side checks of the C-ABI library (no GPU needed): it loads, exports every
symbol include/p3gpu.h declares, and its host-side pieces (transcript /
challenger, sponge hash, compression, seeded generator, parameter rules)
agree with the oracle."""
import ctypes as C
import os
import random
import re

import numpy as np
import pytest

import oracle as O
import p3gpu as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "p3gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(p3g_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(G.LIB_PATH)
    names = _declared_symbols()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_fill_random_matches_oracle_generator():
    assert G.fill_random(2025, 1000).tolist() == O.rand_matrix(2025, 1, 1000).ravel().tolist()


@pytest.mark.parametrize("kind", [G.HASH_POSEIDON2, G.HASH_KECCAK])
def test_lmcs_hash_and_compress_match_oracle(kind):
    rng = random.Random(kind)
    # LmcsConfig::hash / compress are host-side: no device needed
    cfg = G.LmcsConfig.__new__(G.LmcsConfig)
    cfg.hash_kind = kind
    for widths in [[1], [7], [8], [9, 3], [17, 17, 1], [0, 5], [40]]:
        rows = [[rng.randrange(O.P) for _ in range(w)] for w in widths]
        flat = [x for r in rows for x in r]
        want = O.Tree(kind, [np.array([r], dtype=np.uint64) for r in rows if r] or [np.zeros((1, 0), np.uint64)])
        st = np.zeros(25, np.uint64)
        for r in rows:
            st = O.sponge_absorb(kind, st, r)
        assert cfg.hash(rows).tolist() == st[:4].tolist()
        del want, flat
    l = [rng.randrange(O.P) for _ in range(4)]
    r = [rng.randrange(O.P) for _ in range(4)]
    assert cfg.compress(l, r).tolist() == O.compress(kind, l, r).tolist()


@pytest.mark.parametrize("kind", [G.CHALLENGER_DUPLEX_P2, G.CHALLENGER_KECCAK])
def test_transcript_matches_oracle(kind):
    rng = random.Random(7 + kind)
    g = G.ProverTranscript(kind)
    o = O.Transcript(kind)
    for step in range(300):
        op = rng.randrange(7)
        if op == 0:
            v = [rng.randrange(O.P) for _ in range(rng.randrange(1, 12))]
            g.send_field_slice(v); o.send_fields(v)
        elif op == 1:
            c = [rng.randrange(O.P) for _ in range(4)]
            g.send_commitment(c); o.send_commitment(c)
        elif op == 2:
            assert g.sample() == o.sample()
        elif op == 3:
            assert g.sample_algebra_element() == o.sample_ef()
        elif op == 4:
            b = rng.randrange(1, 25)
            assert g.sample_bits(b) == o.sample_bits(b)
        elif op == 5:
            v = [rng.randrange(O.P) for _ in range(rng.randrange(1, 5))]
            g.hint_field_slice(v); o.hint_fields(v)
        else:
            b = rng.randrange(0, 7)
            assert g.grind(b) == o.grind(b)
    gf, gc = g.data()
    of, oc = o.data()
    assert np.array_equal(gf, of) and np.array_equal(gc, oc)


def test_transcript_rejects_non_canonical_fields():
    g = G.ProverTranscript()
    with pytest.raises(G.P3GError):
        g.send_field_slice([O.P])


def test_pcs_params_validation_mirrors_reference():
    # params.rs:170-204
    with pytest.raises(ValueError):
        G.PcsParams(1, 4, 3)
    with pytest.raises(ValueError):
        G.PcsParams(0, 1, 3)
    with pytest.raises(ValueError):
        G.PcsParams(1, 1, 64)
    with pytest.raises(ValueError):
        G.PcsParams(1, 1, 3, num_queries=0)
    p = G.PcsParams(2, 2, 4, 0, 0, 100, 0)
    assert p.num_rounds(24) == 9 and p.final_poly_degree(24) == 16


def test_proof_bytes_layout():
    f = np.array([1, 2], np.uint64)
    c = np.array([[3, 4, 5, 6]], np.uint64)
    b = G.proof_bytes(f, c)
    assert len(b) == 48 and b[:8] == (1).to_bytes(8, "little") and b[16:24] == (3).to_bytes(8, "little")
