"""ctypes wrapper over liboracle.so, the CPU restatement of the reference's
lifted-PCS path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of bench.py, always as the checker.  The product
(the HIP library and its host driver) never imports this module.

Each wrapper names the reference function it restates; the C sources carry
the file:line citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

P = 0xFFFFFFFF00000001
GENERATOR = 7
HASH_P2 = 0
HASH_KECCAK = 1
HASH_MOCK = 2
CH_DUPLEX_P2 = 0
CH_KECCAK = 1

_u64p = C.POINTER(C.c_uint64)
_szp = C.POINTER(C.c_size_t)


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    lib = C.CDLL(_LIB_PATH)
    vp = C.c_void_p
    sig = {
        "orc_rand_fill": (None, [C.c_uint64, C.c_size_t, _u64p]),
        "orc_p2_init_constants": (None, [C.c_uint64]),
        "orc_p2_constants": (None, [_u64p]),
        "orc_p2_diag": (None, [_u64p]),
        "orc_p2_m4": (None, [_u64p]),
        "orc_p2_external_linear": (None, [_u64p]),
        "orc_p2_internal_linear": (None, [_u64p]),
        "orc_p2_permute": (None, [_u64p]),
        "orc_keccakf": (None, [_u64p]),
        "orc_keccak256": (None, [C.c_char_p, C.c_size_t, C.c_char_p]),
        "orc_gl_mul": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "orc_gl_add": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "orc_gl_sub": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "orc_gl_inv": (C.c_uint64, [C.c_uint64]),
        "orc_gl_pow": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "orc_two_adic_generator": (C.c_uint64, [C.c_uint]),
        "orc_ef_mul": (None, [_u64p, _u64p, _u64p]),
        "orc_ef_inv": (None, [_u64p, _u64p]),
        "orc_ntt": (None, [_u64p, C.c_size_t, C.c_int]),
        "orc_eval_constraints": (C.c_int, [C.c_void_p, C.c_size_t, _u64p, C.c_size_t, C.c_size_t, _u64p,
                                           C.c_size_t, C.c_size_t, _u64p, C.c_size_t, C.c_size_t, _u64p,
                                           C.c_size_t, _u64p, C.c_size_t, C.c_size_t, C.c_uint, C.c_uint64,
                                           _u64p, _u64p]),
        "orc_coset_lde_batch": (None, [_u64p, C.c_size_t, C.c_size_t, C.c_uint, C.c_uint64, _u64p]),
        "orc_bit_reverse_rows": (None, [_u64p, C.c_size_t, C.c_size_t]),
        "orc_naive_coset_dft": (None, [_u64p, C.c_size_t, C.c_size_t, C.c_uint64, _u64p]),
        "orc_hash_alignment": (C.c_size_t, [C.c_int]),
        "orc_field_alignment": (C.c_size_t, [C.c_size_t, C.c_size_t, C.c_size_t]),
        "orc_fri_prove": (C.c_size_t, [C.c_void_p, C.c_int, _u64p, C.c_uint, C.c_void_p, _u64p]),
        "orc_mock_config": (None, [C.c_size_t, C.c_size_t]),
        "orc_sponge_absorb": (None, [C.c_int, _u64p, _u64p, C.c_size_t]),
        "orc_compress": (None, [C.c_int, _u64p, _u64p, _u64p]),
        "orc_hash_rows": (None, [C.c_int, _u64p, C.c_size_t, _u64p]),
        "orc_lmcs_build": (vp, [C.c_int, C.POINTER(_u64p), _szp, _szp, C.c_size_t, C.c_size_t]),
        "orc_lmcs_build_salted": (vp, [C.c_int, C.POINTER(_u64p), _szp, _szp, C.c_size_t, C.c_size_t, _u64p, C.c_size_t]),
        "orc_tree_salt": (C.c_size_t, [vp, C.c_size_t, _u64p]),
        "orc_lmcs_open_batch_salted": (C.c_int, [C.c_int, _u64p, _szp, C.c_size_t, C.c_uint, _szp, C.c_size_t, C.c_size_t,
                                                 vp, _u64p]),
        "orc_tree_free": (None, [vp]),
        "orc_tree_root": (None, [vp, _u64p]),
        "orc_tree_height": (C.c_size_t, [vp]),
        "orc_tree_layer": (None, [vp, C.c_uint, _u64p]),
        "orc_tree_rows": (C.c_size_t, [vp, C.c_size_t, _u64p]),
        "orc_tree_prove_batch": (None, [vp, _szp, C.c_size_t, vp]),
        "orc_lmcs_open_batch": (C.c_int, [C.c_int, _u64p, _szp, C.c_size_t, C.c_uint, _szp, C.c_size_t, vp, _u64p]),
        "orc_tx_new": (vp, [C.c_int]),
        "orc_tx_clone": (vp, [vp]),
        "orc_tx_free": (None, [vp]),
        "orc_tx_observe": (None, [vp, C.c_uint64]),
        "orc_tx_observe_commitment": (None, [vp, _u64p]),
        "orc_tx_sample": (C.c_uint64, [vp]),
        "orc_tx_sample_ef": (None, [vp, _u64p]),
        "orc_tx_sample_bits": (C.c_size_t, [vp, C.c_uint]),
        "orc_tx_send_fields": (None, [vp, _u64p, C.c_size_t]),
        "orc_tx_send_commitment": (None, [vp, _u64p]),
        "orc_tx_hint_fields": (None, [vp, _u64p, C.c_size_t]),
        "orc_tx_hint_commitment": (None, [vp, _u64p]),
        "orc_tx_grind": (C.c_uint64, [vp, C.c_uint]),
        "orc_tx_nfields": (C.c_size_t, [vp]),
        "orc_tx_ncomms": (C.c_size_t, [vp]),
        "orc_tx_get": (None, [vp, _u64p, _u64p]),
        "orc_tx_load": (None, [vp, _u64p, C.c_size_t, _u64p, C.c_size_t]),
        "orc_tx_is_empty": (C.c_int, [vp]),
        "orc_tx_check_witness": (C.c_int, [vp, C.c_uint, C.c_uint64]),
        "orc_fri_num_rounds": (C.c_size_t, [vp, C.c_uint]),
        "orc_fri_final_degree": (C.c_size_t, [vp, C.c_uint]),
        "orc_bit_reversed_coset_points": (None, [C.c_uint, _u64p]),
        "orc_point_quotients": (None, [_u64p, C.c_size_t, _u64p, C.c_size_t, _u64p]),
        "orc_batch_eval_lifted": (None, [C.POINTER(vp), C.c_size_t, _u64p, C.c_size_t, C.c_uint, _u64p]),
        "orc_fri_fold_row": (None, [C.c_uint, _u64p, C.c_uint64, _u64p, _u64p]),
        "orc_idft_ext": (None, [_u64p, C.c_size_t]),
        "orc_open_with_channel": (None, [vp, C.c_int, C.POINTER(vp), C.c_size_t, C.c_uint, _u64p, C.c_size_t, vp, _u64p]),
        "orc_verify": (C.c_int, [vp, C.c_int, _u64p, _szp, _szp, C.c_size_t, C.c_uint, _u64p, C.c_size_t, vp, _u64p]),
        "orc_commit_traces": (vp, [C.c_int, C.POINTER(_u64p), _szp, _szp, C.c_size_t, C.c_uint]),
        "orc_quotient_lde": (C.c_int, [_u64p, C.c_size_t, C.c_size_t, C.c_uint, _u64p]),
        "orc_commit_quotient": (vp, [C.c_int, _u64p, C.c_size_t, C.c_size_t, C.c_uint]),
        "orc_cyclic_extend_scale": (C.c_int, [_u64p, C.c_size_t, _u64p, C.c_size_t, _u64p]),
        "orc_divide_by_vanishing": (C.c_int, [_u64p, C.c_size_t, C.c_size_t, C.c_uint64]),
        "orc_set_threads": (None, [C.c_int]),
        "orc_get_threads": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def _p(a: np.ndarray):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u64p)


def _sz(vals) -> C.Array:
    arr = (C.c_size_t * max(1, len(vals)))(*[int(v) for v in vals])
    return arr


def u64(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64))


# ---------------------------------------------------------------- generator
def rand_matrix(seed: int, height: int, width: int) -> np.ndarray:
    """RowMajorMatrix::rand(SmallRng::seed_from_u64(seed), h, w)."""
    out = np.empty(height * width, dtype=np.uint64)
    lib.orc_rand_fill(seed, out.size, _p(out))
    return out.reshape(height, width)


# ---------------------------------------------------------------- field
def gl_mul(a, b): return int(lib.orc_gl_mul(a, b))
def gl_add(a, b): return int(lib.orc_gl_add(a, b))
def gl_sub(a, b): return int(lib.orc_gl_sub(a, b))
def gl_inv(a): return int(lib.orc_gl_inv(a))
def gl_pow(a, e): return int(lib.orc_gl_pow(a, e))
def two_adic_generator(bits): return int(lib.orc_two_adic_generator(bits))


def ef_mul(a, b):
    x, y, o = u64(a), u64(b), np.zeros(2, np.uint64)
    lib.orc_ef_mul(_p(x), _p(y), _p(o))
    return (int(o[0]), int(o[1]))


def ef_inv(a):
    x, o = u64(a), np.zeros(2, np.uint64)
    lib.orc_ef_inv(_p(x), _p(o))
    return (int(o[0]), int(o[1]))


# ---------------------------------------------------------------- DFT
def ntt(v, inverse=False) -> np.ndarray:
    a = u64(v).copy()
    lib.orc_ntt(_p(a), a.size, 1 if inverse else 0)
    return a


def coset_lde_batch(trace: np.ndarray, added_bits: int, shift: int, bit_reverse: bool = True) -> np.ndarray:
    """coset_lde_batch(mat, added_bits, shift) then bit_reverse_rows (commit.rs:217-221)."""
    t = u64(trace)
    n, w = t.shape
    out = np.empty((n << added_bits) * w, dtype=np.uint64)
    lib.orc_coset_lde_batch(_p(t), n, w, added_bits, shift, _p(out))
    if bit_reverse:
        lib.orc_bit_reverse_rows(_p(out), n << added_bits, w)
    return out.reshape(n << added_bits, w)


def bit_reverse_rows(m: np.ndarray) -> np.ndarray:
    a = u64(m).copy()
    lib.orc_bit_reverse_rows(_p(a), a.shape[0], a.shape[1] if a.ndim > 1 else 1)
    return a


def naive_coset_dft(coeffs: np.ndarray, shift: int) -> np.ndarray:
    c = u64(coeffs)
    n, w = c.shape
    out = np.empty(n * w, dtype=np.uint64)
    lib.orc_naive_coset_dft(_p(c), n, w, shift, _p(out))
    return out.reshape(n, w)


# ---------------------------------------------------------------- permutations / sponges
def p2_permute(state) -> np.ndarray:
    s = u64(state).copy()
    assert s.size == 12
    lib.orc_p2_permute(_p(s))
    return s


def p2_constants() -> np.ndarray:
    out = np.zeros(118, dtype=np.uint64)
    lib.orc_p2_constants(_p(out))
    return out


def p2_diag() -> np.ndarray:
    out = np.zeros(12, dtype=np.uint64)
    lib.orc_p2_diag(_p(out))
    return out


def keccakf(state) -> np.ndarray:
    s = u64(state).copy()
    assert s.size == 25
    lib.orc_keccakf(_p(s))
    return s


def keccak256(data: bytes) -> bytes:
    out = C.create_string_buffer(32)
    lib.orc_keccak256(data, len(data), out)
    return out.raw


def mock_config(width: int, rate: int) -> None:
    """Configure the mock permutation (width, rate) used by HASH_MOCK."""
    lib.orc_mock_config(width, rate)


def hash_alignment(kind: int) -> int:
    return int(lib.orc_hash_alignment(kind))


def field_alignment(field_bytes: int, word_bytes: int, rate: int) -> int:
    """compute_field_alignment (serializing_sponge.rs:166-203)."""
    return int(lib.orc_field_alignment(field_bytes, word_bytes, rate))


def sponge_absorb(kind: int, state: np.ndarray, values) -> np.ndarray:
    s = u64(state).copy()
    v = u64(values)
    lib.orc_sponge_absorb(kind, _p(s), _p(v) if v.size else None, v.size)
    return s


def compress(kind: int, left, right) -> np.ndarray:
    a, b, o = u64(left), u64(right), np.zeros(4, np.uint64)
    lib.orc_compress(kind, _p(a), _p(b), _p(o))
    return o


def hash_rows(kind: int, rows) -> np.ndarray:
    flat = u64(np.concatenate([u64(r).ravel() for r in rows])) if rows else np.zeros(0, np.uint64)
    o = np.zeros(4, np.uint64)
    lib.orc_hash_rows(kind, _p(flat) if flat.size else None, flat.size, _p(o))
    return o


# ---------------------------------------------------------------- transcript
class Transcript:
    """ProverTranscript / VerifierTranscript over the configured challenger."""

    def __init__(self, kind: int = CH_DUPLEX_P2, _handle=None):
        self.kind = kind
        self.h = _handle if _handle is not None else lib.orc_tx_new(kind)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib.orc_tx_free(h)
            self.h = None

    def clone(self) -> "Transcript":
        return Transcript(self.kind, lib.orc_tx_clone(self.h))

    def observe(self, v: int): lib.orc_tx_observe(self.h, int(v))

    def observe_commitment(self, c):
        a = u64(c)
        lib.orc_tx_observe_commitment(self.h, _p(a))

    def sample(self) -> int: return int(lib.orc_tx_sample(self.h))

    def sample_ef(self):
        o = np.zeros(2, np.uint64)
        lib.orc_tx_sample_ef(self.h, _p(o))
        return (int(o[0]), int(o[1]))

    def sample_bits(self, bits: int) -> int: return int(lib.orc_tx_sample_bits(self.h, bits))

    def send_fields(self, v):
        a = u64(v)
        lib.orc_tx_send_fields(self.h, _p(a), a.size)

    def send_commitment(self, c):
        a = u64(c)
        lib.orc_tx_send_commitment(self.h, _p(a))

    def hint_fields(self, v):
        a = u64(v)
        lib.orc_tx_hint_fields(self.h, _p(a), a.size)

    def hint_commitment(self, c):
        a = u64(c)
        lib.orc_tx_hint_commitment(self.h, _p(a))

    def grind(self, bits: int) -> int: return int(lib.orc_tx_grind(self.h, bits))

    def check_witness(self, bits: int, w: int) -> bool: return bool(lib.orc_tx_check_witness(self.h, bits, w))

    def data(self):
        nf, nc = lib.orc_tx_nfields(self.h), lib.orc_tx_ncomms(self.h)
        f = np.zeros(max(nf, 1), np.uint64)
        c = np.zeros(max(4 * nc, 1), np.uint64)
        lib.orc_tx_get(self.h, _p(f), _p(c))
        return f[:nf].copy(), c[: 4 * nc].reshape(nc, 4).copy()

    def load(self, fields, comms):
        f = u64(fields).ravel()
        c = u64(comms).ravel()
        lib.orc_tx_load(self.h, _p(f) if f.size else None, f.size, _p(c) if c.size else None, c.size // 4)

    def is_empty(self) -> bool: return bool(lib.orc_tx_is_empty(self.h))


# ---------------------------------------------------------------- LMCS
class Tree:
    """LiftedMerkleTree built by LmcsConfig::build_tree / build_aligned_tree."""

    def __init__(self, kind: int, matrices, alignment: int = 1, _handle=None, salt=None):
        """salt: a height x SALT matrix for a hiding tree (HidingLmcsConfig, hiding_lmcs.rs:104-140)."""
        self.kind = kind
        if _handle is not None:
            self.h = _handle
            return
        mats = [u64(m) for m in matrices]
        ptrs = (_u64p * len(mats))(*[_p(m) for m in mats])
        hs = _sz([m.shape[0] for m in mats])
        ws = _sz([m.shape[1] for m in mats])
        if salt is None:
            self.h = lib.orc_lmcs_build(kind, ptrs, hs, ws, len(mats), alignment)
        else:
            sl = u64(salt)
            if sl.ndim != 2 or not mats or sl.shape[0] != mats[-1].shape[0] or sl.shape[1] == 0:
                raise ValueError("salt must be a tree-height x SALT matrix")
            self._salt_keep = sl
            self.h = lib.orc_lmcs_build_salted(kind, ptrs, hs, ws, len(mats), alignment, _p(sl), sl.shape[1])
        if not self.h:
            raise ValueError("invalid LMCS input (empty, non power-of-two or unsorted heights)")

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib.orc_tree_free(h)
            self.h = None

    def root(self) -> np.ndarray:
        o = np.zeros(4, np.uint64)
        lib.orc_tree_root(self.h, _p(o))
        return o

    def height(self) -> int: return int(lib.orc_tree_height(self.h))

    def salt(self, index: int) -> np.ndarray:
        n = lib.orc_tree_salt(self.h, index, None)
        o = np.zeros(max(n, 1), np.uint64)
        if n:
            lib.orc_tree_salt(self.h, index, _p(o))
        return o[:n]

    def layer(self, l: int) -> np.ndarray:
        n = self.height() >> l
        o = np.zeros(4 * n, np.uint64)
        lib.orc_tree_layer(self.h, l, _p(o))
        return o.reshape(n, 4)

    def rows(self, index: int) -> np.ndarray:
        n = lib.orc_tree_rows(self.h, index, None)
        o = np.zeros(max(n, 1), np.uint64)
        lib.orc_tree_rows(self.h, index, _p(o))
        return o[:n]

    def prove_batch(self, indices, tx: Transcript):
        idx = _sz(list(indices))
        lib.orc_tree_prove_batch(self.h, idx, len(list(indices)), tx.h)


def open_batch(kind: int, root, widths, log_max_height: int, indices, tx: Transcript, salt_width: int = 0):
    """LmcsConfig::open_batch (lmcs.rs:167-250); returns (code, rows) with code 0 ok / 1 invalid / 2 root
    mismatch / 3 transcript.  salt_width > 0: a hiding tree's openings (rows, then the salt)."""
    r = u64(root)
    uniq = sorted(set(int(i) for i in indices))
    tot = int(sum(widths))
    rows = np.zeros(max(1, len(uniq) * tot), np.uint64)
    idx = list(indices)
    rc = lib.orc_lmcs_open_batch_salted(kind, _p(r), _sz(widths), len(widths), log_max_height, _sz(idx), len(idx),
                                        salt_width, tx.h, _p(rows))
    return int(rc), rows[: len(uniq) * tot].reshape(len(uniq), tot) if uniq else rows[:0]


def commit_traces(kind: int, traces, log_blowup: int) -> Tree:
    ts = [u64(t) for t in traces]
    ptrs = (_u64p * len(ts))(*[_p(t) for t in ts])
    h = lib.orc_commit_traces(kind, ptrs, _sz([t.shape[0] for t in ts]), _sz([t.shape[1] for t in ts]), len(ts), log_blowup)
    if not h:
        raise ValueError("invalid traces")
    return Tree(kind, None, _handle=h)


def quotient_lde(q_evals: np.ndarray, d: int, log_blowup: int) -> np.ndarray:
    """The bit-reversed (N*B) x 2D matrix commit_quotient commits (quotient.rs:133-235).

    q_evals: N*D extension elements as an (N*D, 2) array (c0, c1), natural order on gJ."""
    q = u64(q_evals).reshape(-1, 2)
    if d <= 0 or q.shape[0] % d:
        raise ValueError("quotient: q_evals length must be divisible by D")
    n = q.shape[0] // d
    out = np.zeros(((n << log_blowup), 2 * d), np.uint64)
    if lib.orc_quotient_lde(_p(q), n, d, log_blowup, _p(out)) != 0:
        raise ValueError("quotient: N and D must be powers of two with D <= blowup")
    return out


def commit_quotient(kind: int, q_evals: np.ndarray, d: int, log_blowup: int) -> Tree:
    """commit_quotient (p3-miden-lifted-stark/src/prover/quotient.rs:133-238)."""
    q = u64(q_evals).reshape(-1, 2)
    if d <= 0 or q.shape[0] % d:
        raise ValueError("quotient: q_evals length must be divisible by D")
    h = lib.orc_commit_quotient(kind, _p(q), q.shape[0] // d, d, log_blowup)
    if not h:
        raise ValueError("quotient: N and D must be powers of two with D <= blowup")
    return Tree(kind, None, _handle=h)


def cyclic_extend_scale(acc: np.ndarray, beta, target: int) -> np.ndarray:
    """cyclic_extend_and_scale (quotient.rs:40-55): acc is (m, 2) EF; returns (target, 2)."""
    a = u64(acc).reshape(-1, 2)
    b = u64(np.asarray(beta, dtype=np.uint64).ravel())
    out = np.zeros((target, 2), np.uint64)
    if lib.orc_cyclic_extend_scale(_p(a) if a.size else None, a.shape[0], _p(b), target, _p(out)) != 0:
        raise ValueError("cyclic_extend_scale: sizes must be powers of two with m <= target")
    return out


def divide_by_vanishing(num: np.ndarray, d: int, shift: int) -> np.ndarray:
    """divide_by_vanishing_in_place (quotient.rs:72-101) on an (n*d, 2) EF array; returns a copy."""
    a = u64(num).reshape(-1, 2).copy()
    if d <= 0 or a.shape[0] % d:
        raise ValueError("divide_by_vanishing: length must be a multiple of D")
    if lib.orc_divide_by_vanishing(_p(a), a.shape[0] // d, d, shift) != 0:
        raise ValueError("divide_by_vanishing: bad shape or the vanishing polynomial is zero on the coset")
    return a


# ---------------------------------------------------------------- constraint evaluation
AIR_OP_DTYPE = np.dtype([("op", "<u4"), ("a", "<u4"), ("b", "<u4"), ("reserved", "<u4"), ("imm", "<u8")])


def eval_constraints(ops, main, n: int, log_d: int, shift: int, alpha, acc=None, aux=None, periodic=None,
                     publics=(), ext_inputs=()) -> np.ndarray:
    """evaluate_constraints_into (p3-miden-lifted-stark/src/prover/constraints/mod.rs:63-179) for a
    straight-line constraint program (structured array of AIR_OP_DTYPE).  main / aux: committed LDEs
    with bit-reversed rows (aux columns as c0, c1 pairs); periodic: natural-order table on gJ.
    Returns acc + the folded constraints, an (n*d, 2) array."""
    o = np.ascontiguousarray(np.asarray(ops, dtype=AIR_OP_DTYPE))
    m = u64(main)
    M = n << log_d
    out = np.zeros((M, 2), np.uint64) if acc is None else u64(acc).reshape(M, 2).copy()
    x = u64(aux) if aux is not None else None
    pr = u64(periodic) if periodic is not None else None
    pub = u64(np.atleast_1d(np.asarray(publics, dtype=np.uint64)))
    ext = u64(np.asarray(ext_inputs, dtype=np.uint64).reshape(-1))
    al = u64(alpha)
    rc = lib.orc_eval_constraints(o.ctypes.data_as(C.c_void_p), o.size, _p(m), m.shape[0], m.shape[1],
                                  _p(x) if x is not None else None, x.shape[0] if x is not None else 0,
                                  x.shape[1] if x is not None else 0,
                                  _p(pr) if pr is not None else None, pr.shape[0] if pr is not None else 0,
                                  pr.shape[1] if pr is not None else 0, _p(pub), len(publics), _p(ext),
                                  ext.size // 2, n, log_d, shift, _p(al), _p(out))
    if rc != 0:
        raise ValueError("eval_constraints: malformed program or shape")
    return out


# ---------------------------------------------------------------- PCS
class PcsParams(C.Structure):
    """PcsParams::new(log_blowup, log_folding_arity, log_final_degree, folding_pow_bits,
    deep_pow_bits, num_queries, query_pow_bits) (p3-miden-lifted-fri/src/params.rs:54-88)."""

    _fields_ = [(n, C.c_uint32) for n in ("log_blowup", "log_arity", "log_final_degree", "folding_pow_bits",
                                           "deep_pow_bits", "num_queries", "query_pow_bits")]

    def __init__(self, log_blowup, log_arity, log_final_degree, folding_pow_bits, deep_pow_bits, num_queries, query_pow_bits):
        if log_arity not in (1, 2, 3):
            raise ValueError("invalid folding arity")
        if log_blowup == 0:
            raise ValueError("log_blowup must be > 0")
        if log_final_degree + log_blowup > 64:
            raise ValueError("final domain too large")
        if num_queries == 0:
            raise ValueError("num_queries must be > 0")
        super().__init__(log_blowup, log_arity, log_final_degree, folding_pow_bits, deep_pow_bits, num_queries, query_pow_bits)

    def num_rounds(self, log_domain: int) -> int: return int(lib.orc_fri_num_rounds(C.byref(self), log_domain))

    def final_degree(self, log_domain: int) -> int: return int(lib.orc_fri_final_degree(C.byref(self), log_domain))


def bit_reversed_coset_points(log_n: int) -> np.ndarray:
    o = np.zeros(1 << log_n, np.uint64)
    lib.orc_bit_reversed_coset_points(log_n, _p(o))
    return o


def point_quotients(z_points, xs) -> np.ndarray:
    z = u64(z_points).ravel()
    x = u64(xs)
    o = np.zeros(x.size * (z.size // 2) * 2, np.uint64)
    lib.orc_point_quotients(_p(z), z.size // 2, _p(x), x.size, _p(o))
    return o.reshape(x.size, z.size // 2, 2)


def batch_eval_lifted_widths(trees, total_width: int, z_points, log_blowup: int) -> np.ndarray:
    """PointQuotients::batch_eval_lifted: returns [total_width, npts, 2]."""
    z = u64(z_points).ravel()
    n = z.size // 2
    hs = (C.c_void_p * len(trees))(*[t.h for t in trees])
    o = np.zeros(max(1, total_width * n * 2), np.uint64)
    lib.orc_batch_eval_lifted(hs, len(trees), _p(z), n, log_blowup, _p(o))
    return o[: total_width * n * 2].reshape(total_width, n, 2)


def idft_ext(values) -> np.ndarray:
    """Inverse DFT of n extension elements ([n][2], natural order in and out)."""
    v = u64(values).reshape(-1).copy()
    lib.orc_idft_ext(_p(v), v.size // 2)
    return v.reshape(-1, 2)


def fri_fold_row(log_arity: int, row, s_inv: int, beta) -> tuple:
    r, b, o = u64(row).ravel(), u64(beta), np.zeros(2, np.uint64)
    lib.orc_fri_fold_row(log_arity, _p(r), s_inv, _p(b), _p(o))
    return (int(o[0]), int(o[1]))


def open_with_channel(params: PcsParams, kind: int, trees, log_lde_height: int, z_points, tx: Transcript,
                      want_deep: bool = False):
    z = u64(z_points).ravel()
    hs = (C.c_void_p * len(trees))(*[t.h for t in trees])
    deep = np.zeros(2 << log_lde_height, np.uint64) if want_deep else None
    lib.orc_open_with_channel(C.byref(params), kind, hs, len(trees), log_lde_height, _p(z), z.size // 2, tx.h,
                              _p(deep) if want_deep else None)
    return deep.reshape(-1, 2) if want_deep else None


def verify(params: PcsParams, kind: int, roots, widths_per_tree, log_lde_height: int, z_points, tx: Transcript) -> int:
    """verify (aligned widths).  0 = accept; also requires the transcript to be fully consumed."""
    r = u64(np.concatenate([u64(x).ravel() for x in roots]))
    flat = [int(w) for ws in widths_per_tree for w in ws]
    z = u64(z_points).ravel()
    rc = lib.orc_verify(C.byref(params), kind, _p(r), _sz(flat), _sz([len(ws) for ws in widths_per_tree]),
                        len(widths_per_tree), log_lde_height, _p(z), z.size // 2, tx.h, None)
    if rc == 0 and not tx.is_empty():
        return 99
    return int(rc)


def aligned_len(n: int, a: int) -> int:
    return n if a <= 1 else (n + a - 1) // a * a


def pcs_trace_prove(params: PcsParams, kind: int, traces, seed_unused=None):
    """The pcs_trace harness flow (p3-miden-lifted-fri/benches/pcs_trace.rs:60-100) with commit_traces
    semantics: commit, observe the root, sample two EF points, open.  Returns (root, z, fields, comms, tree)."""
    ch = CH_DUPLEX_P2 if kind == HASH_P2 else CH_KECCAK
    tree = commit_traces(kind, traces, params.log_blowup)
    root = tree.root()
    tx = Transcript(ch)
    tx.observe_commitment(root)
    z0 = tx.sample_ef()
    z1 = tx.sample_ef()
    logL = tree.height().bit_length() - 1
    open_with_channel(params, kind, [tree], logL, [z0, z1], tx)
    fields, comms = tx.data()
    return root, (z0, z1), fields, comms, tree


def pcs_trace_verify(params: PcsParams, kind: int, root, widths, log_lde_height: int, fields, comms) -> int:
    ch = CH_DUPLEX_P2 if kind == HASH_P2 else CH_KECCAK
    tx = Transcript(ch)
    tx.observe_commitment(root)
    z0 = tx.sample_ef()
    z1 = tx.sample_ef()
    tx.load(fields, comms)
    al = hash_alignment(kind)
    return verify(params, kind, [root], [[aligned_len(w, al) for w in widths]], log_lde_height, [z0, z1], tx)


def set_threads(n: int) -> None:
    lib.orc_set_threads(n)


def get_threads() -> int:
    return int(lib.orc_get_threads())


def fri_prove(params: PcsParams, kind: int, evals, log_lde_height: int, tx: Transcript):
    """FriPolys::new over caller evaluations (L extension elements, bit-reversed order).
    Returns (rounds, final polynomial as (deg, 2) words in descending degree order)."""
    ev = u64(evals).reshape(-1, 2)
    assert ev.shape[0] == 1 << log_lde_height
    fdeg = params.final_degree(log_lde_height)
    out = np.zeros((max(1, fdeg), 2), np.uint64)
    n = lib.orc_fri_prove(C.byref(params), kind, _p(ev), log_lde_height, tx.h, _p(out))
    return int(n), out[:fdeg]
