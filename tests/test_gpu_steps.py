"""GPU parity of the step-level DEEP and FRI calls (batch_eval_lifted,
deep_quotient, fri_commit, fri_fold, fri_final_poly) through the C ABI: each against the
oracle's own function on the same inputs, and the steps composed in the order
of open_with_channel (p3-miden-lifted-fri/src/prover.rs:35-116) against the
oracle's whole opening, transcript word for word.  Bit-exact.
"""
import numpy as np
import pytest

import oracle as O
import p3gpu as G
from compose import aligned_per_point as _aligned_per_point, bitrev as _bitrev, compose_open

pytestmark = pytest.mark.gpu

P = 0xFFFFFFFF00000001


GROUPS = [
    # (hash, groups of trace shapes, log_blowup, points)
    (G.HASH_POSEIDON2, [[(64, 6), (128, 10)], [(128, 3)]], 1, [(11, 12), (13, 14), (15, 16)]),
    (G.HASH_KECCAK, [[(32, 5), (32, 1), (256, 9)]], 2, [(3, 5), (7, 9)]),
    (G.HASH_POSEIDON2, [[(512, 17)]], 3, [(21, 22), (23, 24), (25, 26), (27, 28)]),
    (G.HASH_POSEIDON2, [[(8, 2), (16, 4)]], 1, [(101, 0)]),  # base-field point off the coset
]


def _commit(device, kind, groups, lb, seed):
    cfg = G.LmcsConfig(device, kind)
    okind = O.HASH_P2 if kind == G.HASH_POSEIDON2 else O.HASH_KECCAK
    gts, ots, widths = [], [], []
    for g, shapes in enumerate(groups):
        tr = [O.rand_matrix(seed + 10 * g + i, h, w) for i, (h, w) in enumerate(shapes)]
        gts.append(G.commit_traces(cfg, tr, lb))
        ots.append(O.commit_traces(okind, tr, lb))
        widths += [w for _, w in shapes]
    return cfg, okind, gts, ots, widths


@pytest.mark.parametrize("case", range(len(GROUPS)))
def test_batch_eval_lifted_matches_oracle(device, case):
    kind, groups, lb, z = GROUPS[case]
    _, _, gts, ots, widths = _commit(device, kind, groups, lb, 300 + case)
    got = G.batch_eval_lifted(gts, z, lb)
    want = O.batch_eval_lifted_widths(ots, sum(widths), z, lb)
    assert got.shape == (sum(widths), len(z), 2)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("case", range(len(GROUPS)))
def test_deep_quotient_matches_oracle_opening(device, case):
    kind, groups, lb, z = GROUPS[case]
    _, okind, gts, ots, widths = _commit(device, kind, groups, lb, 400 + case)
    logL = int(gts[0].height()).bit_length() - 1
    params = O.PcsParams(lb, 1, 0, 0, 0, 3, 0)
    otx = O.Transcript(O.CH_DUPLEX_P2 if kind == G.HASH_POSEIDON2 else O.CH_KECCAK)
    ctx = otx.clone()
    want = O.open_with_channel(params, okind, ots, logL, z, otx, want_deep=True)
    # the challenges the oracle's opening drew: aligned evals sent per point, then alpha and beta
    ev = G.batch_eval_lifted(gts, z, lb)
    for flat in _aligned_per_point(ev, widths, gts[0].alignment()):
        ctx.send_fields(flat)
    ctx.grind(0)
    alpha, beta = ctx.sample_ef(), ctx.sample_ef()
    q = G.deep_quotient(gts, z, lb, ev, alpha, beta)
    assert np.array_equal(q.download(), want)


@pytest.mark.parametrize("logn,la", [(1, 1), (2, 2), (3, 3), (6, 1), (9, 2), (12, 3), (13, 1), (11, 2)])
def test_fri_fold_matches_oracle(device, logn, la):
    n, a = 1 << logn, 1 << la
    ev = O.rand_matrix(900 + logn * 4 + la, n, 2)
    beta = (int(O.rand_matrix(77 + la, 1, 1)[0, 0]), 12345)
    d = device.upload(ev)
    got = G.fri_fold(d, la, beta).download()
    ginv = O.gl_inv(O.two_adic_generator(logn))
    rows = n // a
    lr = logn - la
    want = np.array([O.fri_fold_row(la, ev[k * a:(k + 1) * a], O.gl_pow(ginv, _bitrev(k, lr)), beta)
                     for k in range(rows)], np.uint64)
    assert np.array_equal(got, want.reshape(rows, 2))


def test_fri_fold_of_a_low_degree_polynomial(device):
    # folding p(x) = sum_j x^j p_j(x^a) by beta gives sum_j beta^j p_j(y) on the squared domain:
    # with p of degree < a (one coefficient per p_j), every folded value equals p evaluated "at beta"
    # in the coefficient sense, i.e. sum_j c_j beta^j, whatever the point.
    la, logn = 2, 8
    n, a = 1 << logn, 1 << la
    c = [5, 9, 13, 17]  # base-field coefficients of p, degree < a
    w = O.two_adic_generator(logn)
    ev = np.zeros((n, 2), np.uint64)
    for i in range(n):
        x = O.gl_pow(w, _bitrev(i, logn))
        ev[i, 0] = sum(cj * O.gl_pow(x, j) for j, cj in enumerate(c)) % P
    beta = (3, 0)
    got = G.fri_fold(device.upload(ev), la, beta).download()
    want = sum(cj * 3 ** j for j, cj in enumerate(c)) % P
    assert np.all(got[:, 0] == want) and np.all(got[:, 1] == 0)


@pytest.mark.parametrize("kind", [G.HASH_POSEIDON2, G.HASH_KECCAK])
@pytest.mark.parametrize("logn,la", [(1, 1), (7, 1), (10, 2), (9, 3)])
def test_fri_commit_matches_oracle_tree(device, kind, logn, la):
    # the round tree is the unaligned tree of the (n / a) x 2a base matrix
    n, a = 1 << logn, 1 << la
    ev = O.rand_matrix(950 + logn, n, 2)
    cfg = G.LmcsConfig(device, kind)
    t = G.fri_commit(cfg, device.upload(ev), la)
    okind = O.HASH_P2 if kind == G.HASH_POSEIDON2 else O.HASH_KECCAK
    ot = O.Tree(okind, [ev.reshape(n // a, 2 * a)])
    assert np.array_equal(t.root(), ot.root())
    assert t.height() == n // a and t.widths() == [2 * a]
    assert np.array_equal(t.rows(n // a - 1), ot.rows(n // a - 1))
    view = G.DeviceMatrix(device, t._evals.ptr, n // a, 2 * a, owned=False)
    assert np.array_equal(cfg.build_tree([view]).root(), ot.root())
    d = device.upload(ev)
    for bad in (0, 4, logn + 1):
        with pytest.raises(G.P3GError):
            G.fri_commit(cfg, d, bad)


def test_fri_fold_rejects_bad_shapes(device):
    d = device.upload(np.zeros((8, 2), np.uint64))
    for la in (0, 4):
        with pytest.raises(G.P3GError):
            G.fri_fold(d, la, (1, 0))
    with pytest.raises(G.P3GError):
        G.fri_fold(device.upload(np.zeros((2, 2), np.uint64)), 2, (1, 0))  # domain below the arity
    with pytest.raises(G.P3GError):
        G.fri_fold(device.upload(np.zeros((12, 2), np.uint64)), 1, (1, 0))  # not a power of two
    with pytest.raises(G.P3GError):
        G.fri_fold(d, 1, (P, 0))  # non-canonical challenge


def test_deep_steps_reject_bad_inputs(device):
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    t = G.commit_traces(cfg, [O.rand_matrix(1, 16, 3)], 1)
    lb = 1
    g = 7  # the LDE coset is 7 * H_32: 7 itself lies on it
    with pytest.raises(G.P3GError):
        G.batch_eval_lifted([t], [(g, 0)], lb)
    with pytest.raises(G.P3GError):
        G.batch_eval_lifted([t], [(1, 2)] * 5, lb)  # at most four points
    with pytest.raises(G.P3GError):
        G.batch_eval_lifted([t], [(1, 2)], 6)  # blowup above the LDE height
    ev = G.batch_eval_lifted([t], [(1, 2)], lb)
    with pytest.raises(G.P3GError):
        G.deep_quotient([t], [(1, 2)], lb, ev[:-1], (1, 0), (2, 0))  # count does not match the trees
    t2 = G.commit_traces(cfg, [O.rand_matrix(2, 8, 3)], 1)
    with pytest.raises(G.P3GError):
        G.batch_eval_lifted([t, t2], [(1, 2)], lb)  # mixed LDE heights


@pytest.mark.parametrize("kind,params", [
    (G.HASH_POSEIDON2, (1, 1, 0, 0, 0, 9, 0)),
    (G.HASH_POSEIDON2, (2, 2, 1, 0, 2, 7, 1)),
    (G.HASH_KECCAK, (1, 3, 2, 1, 0, 5, 0)),
    (G.HASH_POSEIDON2, (3, 1, 6, 0, 0, 4, 0)),  # no FRI round: final polynomial only
])
def test_composed_steps_reproduce_open_with_channel(device, kind, params):
    groups = [[(32, 4), (64, 7)], [(64, 2)]]
    lb, la = params[0], params[1]
    cfg, okind, gts, ots, widths = _commit(device, kind, groups, lb, 500 + la)
    logL = int(gts[0].height()).bit_length() - 1
    z = [(31, 32), (33, 34)]
    ch = G.CHALLENGER_DUPLEX_P2 if kind == G.HASH_POSEIDON2 else G.CHALLENGER_KECCAK
    tx = G.ProverTranscript(ch)
    compose_open(cfg, params, gts, widths, z, tx)
    gf, gc = tx.data()
    otx = O.Transcript(O.CH_DUPLEX_P2 if kind == G.HASH_POSEIDON2 else O.CH_KECCAK)
    O.open_with_channel(O.PcsParams(*params), okind, ots, logL, z, otx)
    of, oc = otx.data()
    assert np.array_equal(gf, of) and np.array_equal(gc, oc)
    # and the library's own one-call opening writes the same words
    t2 = G.ProverTranscript(ch)
    G.open_with_channel(G.PcsParams(*params), cfg, logL, z, gts, t2)
    f2, c2 = t2.data()
    assert np.array_equal(f2, of) and np.array_equal(c2, oc)


def test_deep_size_query_validates_like_the_call(device):
    # asking for the evaluation count checks the trees and the point count as the call does
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    t = G.commit_traces(cfg, [O.rand_matrix(3, 16, 5)], 1)
    n = np.zeros(1, np.uint64)
    z = np.array([1, 2] * 5, np.uint64)
    hs = (G.C.c_void_p * 1)(t.h)
    assert G.lib().p3g_deep_batch_eval(device.h, 1, G._p(z), 5, hs, 1, None, 0, G._p(n)) == 1  # INVAL
    assert G.lib().p3g_deep_batch_eval(device.h, 1, G._p(z), 2, hs, 0, None, 0, G._p(n)) == 6  # EMPTY
    assert G.lib().p3g_deep_batch_eval(device.h, 1, G._p(z), 2, hs, 1, None, 0, G._p(n)) == 0
    assert int(n[0]) == 5 * 2 * 2


def test_deep_size_query_rejects_what_the_call_rejects(device):
    # the size query runs the call's input checks too: a point on the LDE coset, a blowup above
    # the LDE height, mixed heights; and a bad point count is refused before the points are read
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    t = G.commit_traces(cfg, [O.rand_matrix(4, 16, 3)], 1)
    t2 = G.commit_traces(cfg, [O.rand_matrix(5, 8, 3)], 1)
    n = np.zeros(1, np.uint64)
    hs = (G.C.c_void_p * 2)(t.h, t2.h)
    on_coset = np.array([7, 0], np.uint64)
    ok = np.array([1, 2], np.uint64)
    lib = G.lib()
    assert lib.p3g_deep_batch_eval(device.h, 1, G._p(on_coset), 1, hs, 1, None, 0, G._p(n)) == 1
    assert lib.p3g_deep_batch_eval(device.h, 6, G._p(ok), 1, hs, 1, None, 0, G._p(n)) == 1
    assert lib.p3g_deep_batch_eval(device.h, 1, G._p(ok), 1, hs, 2, None, 0, G._p(n)) == 1
    # n_points 0: rejected on the count before the point buffer is read
    assert lib.p3g_deep_batch_eval(device.h, 1, G._p(ok), 0, hs, 1, None, 0, G._p(n)) == 1
    assert lib.p3g_deep_batch_eval(device.h, 1, G._p(ok), 1, hs, 1, None, 0, G._p(n)) == 0
