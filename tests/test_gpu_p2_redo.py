"""The hashing kernels' Poseidon2 permutations (one lane per permutation, and
the lane-cooperative form of small trees) take a shortcut in their modular
products that is wrong with probability about 2^-32 per product (a borrow it
does not fold), flag it, and then redo the whole wave's permutation exactly.
Random inputs never reach the redo, so these tests build leaf rows that do and
check the trees against the oracle.

A product a*b is computed from a = a0 + 2^32 a1, b = b0 + 2^32 b1 as
  P = a0 b0, M = a1 b0 + P1, N + cn 2^64 = a0 b1 + M, R = a1 b1 + N1,
  lo = P0 + 2^32 N0, T + c 2^64 = lo + (2^32 - 1) R0,
and it borrows when T < R1 + cn.  For x = 3 * 2^48: x * x = 9 * 2^96, so lo = 0,
R0 = 0, R1 = 9 and T = 0 < 9.  x * x is the first product of an S-box, so a
leaf whose first external layer (plus the first round constant) puts x in
state word 0 takes the redo in its first permutation.
"""
import numpy as np
import pytest

import oracle as O
import p3gpu as G

P = 0xFFFFFFFF00000001
M32 = (1 << 32) - 1
M64 = (1 << 64) - 1
X_BORROW = 3 << 48


def _try_product_borrows(a: int, b: int) -> bool:
    a0, a1, b0, b1 = a & M32, a >> 32, b & M32, b >> 32
    p = a0 * b0
    m = a1 * b0 + (p >> 32)
    nf = a0 * b1 + m
    cn, n = nf >> 64, nf & M64
    r = a1 * b1 + (n >> 32)
    lo = (p & M32) | ((n & M32) << 32)
    t = (lo + M32 * (r & M32)) & M64
    return t < (r >> 32) + cn


def _ext_word0(s) -> int:
    """Word 0 of circ(2 M4, M4, M4) s, M4 = [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]]."""
    k = [4, 6, 2, 2, 2, 3, 1, 1, 2, 3, 1, 1]
    return sum(ki * int(si) for ki, si in zip(k, s)) % P


def _borrow_row(rng) -> np.ndarray:
    """A width-8 leaf row whose sponge input (row, 0, 0, 0, 0) has first external
    layer + rc_init[0][0] equal to X_BORROW in word 0."""
    rc00 = int(O.p2_constants()[0])
    s = [0] + [int(v) for v in rng.integers(0, P, 7, dtype=np.uint64)] + [0] * 4
    rest = _ext_word0(s)
    s[0] = (X_BORROW - rc00 - rest) * pow(4, P - 2, P) % P
    assert (_ext_word0(s) + rc00) % P == X_BORROW
    return np.array(s[:8], dtype=np.uint64)


def test_borrow_construction():
    assert _try_product_borrows(X_BORROW, X_BORROW)
    rng = np.random.default_rng(3)
    vals = [int(v) for v in rng.integers(0, P, 2000, dtype=np.uint64)]
    assert not any(_try_product_borrows(a, b) for a, b in zip(vals[::2], vals[1::2]))
    assert _borrow_row(np.random.default_rng(5)).shape == (8,)


@pytest.mark.gpu
@pytest.mark.parametrize("log_h,rows", [(15, [3, (1 << 14) + 5, (1 << 15) - 1]), (16, [0, 12345]),
                                        # trees of up to 2^14 leaves hash with the lane-cooperative permutation
                                        (10, [1, 700]), (14, [(1 << 14) - 2])])
def test_tree_with_redo_rows_matches_oracle(device, log_h, rows):
    h = 1 << log_h
    m = O.rand_matrix(404 + log_h, h, 8)
    rng = np.random.default_rng(log_h)
    for r in rows:
        m[r] = _borrow_row(rng)
    want = O.Tree(O.HASH_P2, [m], alignment=O.hash_alignment(O.HASH_P2))
    got = G.LmcsConfig(device, G.HASH_POSEIDON2).build_aligned_tree([m])
    assert got.root().tolist() == want.root().tolist()
    for layer in range(log_h + 1):
        assert np.array_equal(got.layer(layer), want.layer(layer)), layer


@pytest.mark.gpu
def test_lifted_tree_with_redo_rows_matches_oracle(device):
    """Two heights: the redo row in the shorter matrix's first permutation, and a
    longer matrix whose leaves inherit that state (a second group)."""
    small = O.rand_matrix(91, 1 << 15, 8)
    big = O.rand_matrix(92, 1 << 16, 16)
    small[7] = _borrow_row(np.random.default_rng(9))
    want = O.Tree(O.HASH_P2, [small, big], alignment=O.hash_alignment(O.HASH_P2))
    got = G.LmcsConfig(device, G.HASH_POSEIDON2).build_aligned_tree([small, big])
    assert got.root().tolist() == want.root().tolist()
    for layer in range(17):
        assert np.array_equal(got.layer(layer), want.layer(layer)), layer


def _trace_with_lde_row(seed: int, n: int, log_blowup: int, j: int, row) -> np.ndarray:
    """An n x 8 trace whose coset LDE (shift 7, bit-reversed rows, as commit_traces
    commits a single trace) has `row` at LDE row j: trace row 0 is moved along the
    LDE of the unit vector e_0, which is nonzero at every coset point."""
    t = O.rand_matrix(seed, n, 8)
    base = O.coset_lde_batch(t, log_blowup, 7)
    e0 = np.zeros((n, 1), dtype=np.uint64)
    e0[0, 0] = 1
    unit = int(O.coset_lde_batch(e0, log_blowup, 7)[j, 0])
    inv = pow(unit, P - 2, P)
    for c in range(8):
        delta = (int(row[c]) - int(base[j, c])) * inv % P
        t[0, c] = np.uint64((int(t[0, c]) + delta) % P)
    assert np.array_equal(O.coset_lde_batch(t, log_blowup, 7)[j], row)
    return t


@pytest.mark.gpu
def test_lockstep_proofs_with_redo_rows_match_oracle(device):
    """Two lock-step jobs of 2^13 x 8 (LDE 2^14 rows each, so the batch hashes its
    leaves with the job-indexed one-lane kernel): leaf rows of both LDEs take the
    redo; the proofs equal the oracle's."""
    n, log_blowup = 1 << 13, 1
    rng = np.random.default_rng(21)
    jobs = [[_trace_with_lde_row(500, n, log_blowup, 77, _borrow_row(rng))],
            [_trace_with_lde_row(501, n, log_blowup, (1 << 14) - 1, _borrow_row(rng))]]
    # the committed matrix of a single trace is its LDE with shift 7
    lde = O.coset_lde_batch(jobs[0][0], log_blowup, 7)
    assert (O.commit_traces(O.HASH_P2, jobs[0], log_blowup).root().tolist()
            == O.Tree(O.HASH_P2, [lde], alignment=O.hash_alignment(O.HASH_P2)).root().tolist())
    params = (log_blowup, 1, 3, 0, 0, 32, 0)
    got = G.prove_many(G.LmcsConfig(device, G.HASH_POSEIDON2), G.PcsParams(*params), jobs)
    for tr, (root, z, f, c) in zip(jobs, got):
        root_o, z_o, f_o, c_o, _ = O.pcs_trace_prove(O.PcsParams(*params), O.HASH_P2, tr)
        assert root.tolist() == root_o.tolist() and z == tuple(z_o)
        assert np.array_equal(f, f_o) and np.array_equal(c, c_o)
