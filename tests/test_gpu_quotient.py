"""GPU parity of the quotient commitment (commit_quotient,
p3-miden-lifted-stark/src/prover/quotient.rs:133-238) against the oracle.

The oracle follows the reference's steps (iDFT over H, (omega_J^-k)^t scaling,
zero padding, plain DFT over K, bit-reversed rows, aligned tree); the HIP path
computes the same matrix as one coset LDE with a per-column coefficient twist.
Bit-exact: committed matrix, root, every digest layer and the openings.
"""
import numpy as np
import pytest

import oracle as O
import p3gpu as G

pytestmark = pytest.mark.gpu

KINDS = [(G.HASH_POSEIDON2, O.HASH_P2), (G.HASH_KECCAK, O.HASH_KECCAK)]


def _q(seed, n, d):
    return O.rand_matrix(seed, n * d, 2)  # n*d extension elements (c0, c1)


def _lde_of(tree):
    ptr, h, w = tree.matrix(0)
    m = G.DeviceMatrix(tree.dev, ptr, h, w, owned=False)
    return m.download()


# (log N, D, log B): the reference bench's D = 2, B = 2 (N = 2^16, 2^17), and
# around it D = 1..8 with B >= D, one- and multi-pass transforms, N = 1
@pytest.mark.parametrize("logn,d,logb", [(0, 1, 1), (0, 2, 1), (1, 2, 1), (3, 4, 2), (6, 2, 1), (8, 8, 3),
                                         (10, 2, 2), (11, 4, 3), (12, 2, 1), (14, 1, 2), (16, 2, 1), (17, 2, 1)])
def test_commit_quotient_matches_oracle(device, logn, d, logb):
    n = 1 << logn
    q = _q(4000 + logn * 16 + d, n, d)
    want_lde = O.quotient_lde(q, d, logb)
    for gk, ok in KINDS:
        if ok == O.HASH_KECCAK and logn > 12:
            continue  # the Keccak tree is covered at the smaller sizes
        t = G.commit_quotient(G.LmcsConfig(device, gk), q, d, logb)
        want = O.commit_quotient(ok, q, d, logb)
        assert np.array_equal(_lde_of(t), want_lde), (gk, logn, d, logb)
        assert t.root().tolist() == want.root().tolist(), (gk, logn, d, logb)
        assert t.height() == n << logb
        a = O.hash_alignment(ok)  # widths are reported aligned, as for trace trees
        assert t.widths() == [-(-2 * d // a) * a] and t.alignment() == a
        for lvl in range(0, logn + logb + 1, 3):
            assert np.array_equal(t.layer(lvl), want.layer(lvl))
        t.free()


def test_commit_quotient_openings_match_oracle(device):
    n, d, logb = 1 << 9, 4, 2
    q = _q(77, n, d)
    t = G.commit_quotient(G.LmcsConfig(device, G.HASH_POSEIDON2), q, d, logb)
    want = O.commit_quotient(O.HASH_P2, q, d, logb)
    idx = [0, 5, 5, 1023, 2047, 777, 1]
    tx_g = G.ProverTranscript(G.CHALLENGER_DUPLEX_P2)
    tx_o = O.Transcript(O.CH_DUPLEX_P2)
    t.prove_batch(idx, tx_g)
    want.prove_batch(idx, tx_o)
    f_g, c_g = tx_g.data()
    f_o, c_o = tx_o.data()
    assert np.array_equal(f_g, f_o) and np.array_equal(c_g, c_o)
    for i in (0, 3, 2047):
        assert np.array_equal(t.rows(i), want.rows(i))
    t.free()


def test_commit_quotient_device_input(device):
    # the quotient accumulator already resident in HBM (N x 2D words)
    n, d, logb = 1 << 12, 2, 1
    q = _q(91, n, d)
    dm = device.upload(q.reshape(n, 2 * d))
    t = G.commit_quotient(G.LmcsConfig(device, G.HASH_POSEIDON2), dm, d, logb)
    assert t.root().tolist() == O.commit_quotient(O.HASH_P2, q, d, logb).root().tolist()
    t.free()
    assert np.array_equal(dm.download().reshape(-1, 2), q)  # the input is left as it was
    dm.free()


def test_commit_quotient_rejects_bad_shapes(device):
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    with pytest.raises(G.P3GError) as e:
        G.commit_quotient(cfg, _q(1, 16, 4), 4, 1)  # B = 2 < D = 4
    assert e.value.code == 1  # bad argument
    with pytest.raises(G.P3GError) as e:
        G.commit_quotient(cfg, O.rand_matrix(2, 12 * 2, 2), 2, 1)  # N = 12
    assert e.value.code == 2  # not a power of two
    with pytest.raises(G.P3GError) as e:
        G.commit_quotient(cfg, O.rand_matrix(3, 16 * 3, 2), 3, 2)  # D = 3
    assert e.value.code == 2  # not a power of two


# ---------------------------------------------------------------- auxiliary traces
def test_commit_aux_traces_is_commit_traces_of_the_flattened_traces(device):
    # prove_multi flattens each EF aux trace to base, c0 then c1 per element
    # (mod.rs:323-337), and commits the flattened traces with commit_traces
    shapes = [(1 << 6, 3), (1 << 9, 1), (1 << 9, 5)]
    aux = [O.rand_matrix(300 + i, h, 2 * w).reshape(h, w, 2) for i, (h, w) in enumerate(shapes)]
    for gk, ok in KINDS:
        t = G.commit_aux_traces(G.LmcsConfig(device, gk), aux, 2)
        flat = []
        for a in aux:
            f = np.empty((a.shape[0], 2 * a.shape[1]), np.uint64)
            f[:, 0::2], f[:, 1::2] = a[:, :, 0], a[:, :, 1]  # element j -> base columns 2j (c0), 2j + 1 (c1)
            flat.append(f)
        want = O.commit_traces(ok, flat, 2)
        assert t.root().tolist() == want.root().tolist()
        assert t.height() == (1 << 9) << 2
        t.free()


# ---------------------------------------------------------------- accumulation helpers (quotient.rs:28-101)
def _dev_ef(device, a, rows=None):
    """A device buffer of `rows` extension elements (2 words each) holding a at its start."""
    a = np.asarray(a, np.uint64).reshape(-1, 2)
    rows = rows or a.shape[0]
    buf = np.zeros((rows, 2), np.uint64)
    buf[: a.shape[0]] = a
    return device.upload(buf)


@pytest.mark.parametrize("m,target", [(0, 8), (1, 1), (1, 64), (4, 4), (4, 1 << 12), (1 << 10, 1 << 14)])
def test_cyclic_extend_and_scale_matches_oracle(device, m, target):
    acc = O.rand_matrix(60 + m, m, 2) if m else np.zeros((0, 2), np.uint64)
    beta = (123456789, O.P - 5)
    buf = _dev_ef(device, acc, target)
    G.cyclic_extend_and_scale(buf, m, target, beta)
    assert np.array_equal(buf.download().reshape(-1, 2), O.cyclic_extend_scale(acc, beta, target))
    buf.free()


@pytest.mark.parametrize("logn,d,shift", [(0, 1, 7), (4, 2, 7), (10, 4, 49), (10, 8, 7), (16, 2, 7), (12, 1, 2401)])
def test_divide_by_vanishing_matches_oracle(device, logn, d, shift):
    n = 1 << logn
    num = O.rand_matrix(70 + logn, n * d, 2)
    buf = _dev_ef(device, num)
    G.divide_by_vanishing(buf, n, d, shift)
    assert np.array_equal(buf.download().reshape(-1, 2), O.divide_by_vanishing(num, d, shift))
    buf.free()


def test_quotient_pipeline_matches_oracle(device):
    # prove_multi's quotient steps (prover/mod.rs:355-434) on given constraint
    # evaluations: accumulate two instances in ascending height with Horner folding in
    # beta, divide by Z_H on the max quotient coset (shift GENERATOR), commit the chunks
    d, logb = 4, 2
    n_small, n_big = 1 << 6, 1 << 9
    beta = (987654321, 42)
    e1 = O.rand_matrix(81, n_small * d, 2)
    e2 = O.rand_matrix(82, n_big * d, 2)
    def addmod(a, b):
        return ((a.astype(object) + b.astype(object)) % O.P).astype(np.uint64)

    # oracle: empty accumulator -> zeros, + e1; extend(acc * beta), + e2; divide; commit
    acc = addmod(O.cyclic_extend_scale(np.zeros((0, 2), np.uint64), beta, n_small * d), e1)
    acc = addmod(O.cyclic_extend_scale(acc, beta, n_big * d), e2)
    acc = O.divide_by_vanishing(acc, d, 7)
    want = O.commit_quotient(O.HASH_P2, acc, d, logb)
    # GPU: the same steps on one device buffer
    buf = _dev_ef(device, np.zeros((0, 2), np.uint64), n_big * d)
    G.cyclic_extend_and_scale(buf, 0, n_small * d, beta)  # empty accumulator: zeros
    host = buf.download().reshape(-1, 2)
    host[: n_small * d] = addmod(host[: n_small * d], e1)  # instance 1's constraint evaluations
    buf.free()
    buf = _dev_ef(device, host[: n_small * d], n_big * d)
    G.cyclic_extend_and_scale(buf, n_small * d, n_big * d, beta)
    host = buf.download().reshape(-1, 2)
    host = addmod(host, e2)  # instance 2's constraint evaluations
    buf.free()
    buf = device.upload(host.reshape(n_big, 2 * d))
    G.divide_by_vanishing(buf, n_big, d, 7)
    t = G.commit_quotient(G.LmcsConfig(device, G.HASH_POSEIDON2), buf, d, logb)
    assert t.root().tolist() == want.root().tolist()
    t.free()
    buf.free()


def test_quotient_helpers_reject_bad_arguments(device):
    buf = _dev_ef(device, np.zeros((16, 2), np.uint64))
    with pytest.raises(G.P3GError) as e:
        G.divide_by_vanishing(buf, 8, 2, 1)  # shift 1: the coset meets H, Z_H = 0 there
    assert e.value.code == 1
    with pytest.raises(G.P3GError) as e:
        G.divide_by_vanishing(buf, 6, 2, 7)
    assert e.value.code == 2
    with pytest.raises(G.P3GError) as e:
        G.cyclic_extend_and_scale(buf, 16, 8, (1, 0))  # target below the accumulator
    assert e.value.code == 1
    with pytest.raises(G.P3GError) as e:
        G.cyclic_extend_and_scale(buf, 3, 16, (1, 0))
    assert e.value.code == 2
    buf.free()
