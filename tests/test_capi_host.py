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


def test_rust_shim_declarations_match_the_header():
    # the Rust shim's extern block (p3-miden_amd/rust/src/ffi.rs) declares only functions of
    # include/p3gpu.h, with the header's parameter counts, and the library exports them all
    text = open(os.path.join(ROOT, "p3-miden_amd", "rust", "src", "ffi.rs")).read()
    block = text[text.index('unsafe extern "C" {'):]
    block = block[:block.index("\n}\n")]
    rust = dict(re.findall(r"pub fn (p3g_[a-z0-9_]+)\(([^)]*)\)", block, flags=re.S))
    assert len(rust) > 15
    header = open(os.path.join(ROOT, "include", "p3gpu.h")).read()
    header = re.sub(r"/\*.*?\*/", "", header, flags=re.S)
    lib = C.CDLL(G.LIB_PATH)
    for name, args in rust.items():
        m = re.search(r"\b" + name + r"\s*\(([^)]*)\)", header, flags=re.S)
        assert m, name
        nh = 0 if m.group(1).strip() in ("", "void") else m.group(1).count(",") + 1
        nr = 0 if not args.strip() else args.rstrip().rstrip(",").count(",") + 1
        assert nh == nr, (name, nh, nr)
        assert hasattr(lib, name), name


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


# ---------------------------------------------------------------- proof wire format
def test_proof_encoding_known_bytes():
    # header "P3GT", version 1, nf, nc; then canonical words, all little-endian
    f = np.array([1, O.P - 1], np.uint64)
    c = np.array([[2, 3, 4, 5]], np.uint64)
    b = G.encode_proof(f, c)
    want = (b"P3GT" + (1).to_bytes(4, "little") + (2).to_bytes(8, "little") + (1).to_bytes(8, "little")
            + b"".join(int(x).to_bytes(8, "little") for x in [1, O.P - 1, 2, 3, 4, 5]))
    assert b == want and len(b) == 24 + 8 * 2 + 32
    df, dc = G.decode_proof(b)
    assert df.tolist() == f.tolist() and dc.tolist() == c.tolist()
    assert G.encode_proof([], np.zeros((0, 4), np.uint64)) == b"P3GT" + (1).to_bytes(4, "little") + bytes(16)


def test_proof_encoding_round_trips_an_oracle_proof():
    tr = [O.rand_matrix(5, 1 << 6, 9)]
    _, _, f, c, _ = O.pcs_trace_prove(O.PcsParams(1, 1, 2, 0, 0, 8, 0), O.HASH_P2, tr)
    b = G.encode_proof(f, c)
    assert len(b) == 24 + 8 * f.size + 32 * c.shape[0]
    assert b[24:] == G.proof_bytes(f, c)  # the header-free word stream follows the header
    df, dc = G.decode_proof(b)
    assert np.array_equal(df, f) and np.array_equal(dc, c)


def test_transcript_encode_matches_encode_proof():
    tx = G.ProverTranscript(G.CHALLENGER_DUPLEX_P2)
    tx.send_field_slice([7, 8, 9])
    tx.send_commitment([1, 2, 3, 4])
    tx.hint_field_slice([10])
    f, c = tx.data()
    assert tx.encode() == G.encode_proof(f, c)


def test_proof_decoding_rejects_malformed_input():
    good = G.encode_proof([1, 2, 3], [[4, 5, 6, 7]])
    bad = [
        good[:-1],                                        # truncated
        good + b"\0",                                     # trailing byte
        b"P3GX" + good[4:],                               # magic
        good[:4] + (2).to_bytes(4, "little") + good[8:],  # version
        good[:8] + (4).to_bytes(8, "little") + good[16:], # field count does not match the length
        good[:24] + O.P.to_bytes(8, "little") + good[32:],  # non-canonical word
        b"",
    ]
    for b in bad:
        with pytest.raises(G.P3GError) as e:
            G.decode_proof(b)
        assert e.value.code == 1
    with pytest.raises(G.P3GError) as e:
        G.encode_proof([O.P], np.zeros((0, 4), np.uint64))
    assert e.value.code == 1


def _bitrev(i, bits):
    return int(format(i, f"0{bits}b")[::-1], 2) if bits else 0


@pytest.mark.parametrize("logn", [0, 1, 2, 3, 5, 7])
def test_fri_final_poly_recovers_coefficients(logn):
    # closed form: evaluate p(x) = sum_k c_k x^k at w_n^bitrev(i) (the last FRI domain in
    # bit-reversed order); the call must return the c_k in descending degree
    P = 0xFFFFFFFF00000001
    n = 1 << logn
    rng = random.Random(40 + logn)
    c = [(rng.randrange(P), rng.randrange(P)) for _ in range(n)]
    w = O.two_adic_generator(logn)
    ev = np.zeros((n, 2), np.uint64)
    for i in range(n):
        x = O.gl_pow(w, _bitrev(i, logn))
        xp, a0, a1 = 1, 0, 0
        for c0, c1 in c:
            a0, a1 = (a0 + c0 * xp) % P, (a1 + c1 * xp) % P
            xp = xp * x % P
        ev[i] = (a0, a1)
    got = G.fri_final_poly(ev)
    assert np.array_equal(got, np.array(c[::-1], np.uint64).reshape(n, 2))


def test_fri_final_poly_matches_oracle_idft():
    # the oracle's inverse DFT on the natural-order values, reversed
    n, lf = 16, 4
    ev = O.rand_matrix(71, n, 2)
    nat = np.zeros_like(ev)
    for i in range(n):
        nat[_bitrev(i, lf)] = ev[i]
    want = O.idft_ext(nat)[::-1]
    assert np.array_equal(G.fri_final_poly(ev), want)


def test_fri_final_poly_rejects_bad_input():
    P = 0xFFFFFFFF00000001
    with pytest.raises(G.P3GError):
        G.fri_final_poly(np.zeros((3, 2), np.uint64))
    with pytest.raises(G.P3GError):
        G.fri_final_poly(np.array([[P, 0], [0, 0]], np.uint64))


def test_binding_fails_loudly_without_the_library(monkeypatch):
    # no CPU fallback: every entry point needs libp3gpu.so
    monkeypatch.setattr(G, "_lib", None)
    with pytest.raises(ImportError):
        G.fill_random(1, 4)
    with pytest.raises(ImportError):
        G.fri_final_poly(np.zeros((2, 2), np.uint64))


def test_product_code_never_references_the_oracle():
    # the oracle is test infrastructure: nothing under the package may import, link or load it
    pkg = os.path.join(ROOT, "p3-miden_amd")
    hits = []
    for d, _, files in os.walk(pkg):
        if "_build" in d:
            continue
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h")) or f == "Makefile":
                text = open(os.path.join(d, f), errors="replace").read()
                if re.search(r"\boracle\b|liboracle|orc_", text):
                    hits.append(f)
    assert not hits, hits
