"""Constraint evaluation on the GPU (p3g_eval_constraints) against the oracle's
restatement of evaluate_constraints_into (p3-miden-lifted-stark/src/prover/
constraints/mod.rs:63-179), word for word:

  - three AIRs (Fibonacci with public values, the reference's DummyMidenAir, a
    running sum over an extension aux column with periodic and permutation
    inputs), several trace heights and quotient degrees, a nonzero accumulator;
  - reading main and aux in place from committed trees (the bit-reversed LDEs
    of commit_traces), then the rest of the quotient pipeline of prove_multi
    (prover/mod.rs:370-432): divide_by_vanishing and commit_quotient, whose root
    must equal the oracle's;
  - the DummyMidenAir at the reference binary's shape (2^19 x 20, log_blowup 3,
    quotient degree 8: 2^22 points);
  - malformed programs and shapes rejected before anything is launched.
"""
import numpy as np
import pytest

import oracle as O
import p3gpu as G
import airs

pytestmark = pytest.mark.gpu

P = airs.P
G7 = 7


def _run_both(device, prog, main, n, logb, log_d, alpha, aux=None, periodic=None, publics=(), ext=(), seed=0):
    M = n << log_d
    rng = np.random.default_rng(seed)
    acc0 = rng.integers(0, P, size=(M, 2), dtype=np.uint64)
    main_lde = O.coset_lde_batch(main, logb, G7)
    aux_lde = O.coset_lde_batch(aux, logb, G7) if aux is not None else None
    want = O.eval_constraints(prog.ops, main_lde, n, log_d, G7, alpha, acc=acc0, aux=aux_lde, periodic=periodic,
                              publics=publics, ext_inputs=ext)
    dm = device.upload(main_lde)
    da = device.upload(aux_lde) if aux_lde is not None else None
    acc = device.upload(acc0)
    print(f"eval_constraints: n={n} log_blowup={logb} log_d={log_d} ops={len(prog.ops)}", flush=True)
    G.eval_constraints(prog, dm, n, log_d, G7, alpha, acc, aux=da, periodic=periodic, publics=publics,
                       ext_inputs=ext)
    got = acc.download()
    for m in (dm, da, acc):
        if m is not None:
            m.free()
    return got, want


@pytest.mark.parametrize("n,logb,log_d", [(16, 1, 1), (256, 2, 1), (1 << 12, 3, 2)])
def test_fibonacci_matches_oracle(device, n, logb, log_d):
    trace, pub = airs.fibonacci_trace(n, 3, 4)
    got, want = _run_both(device, airs.fibonacci_program(), trace, n, logb, log_d, (11, 12), publics=pub, seed=n)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n,width,logb,log_d", [(64, 9, 3, 3), (1 << 10, 20, 3, 3), (256, 51, 4, 4)])
def test_miden_dummy_matches_oracle(device, n, width, logb, log_d):
    trace = airs.miden_trace(n, width, width)
    got, want = _run_both(device, airs.miden_program(), trace, n, logb, log_d, (7, 9), seed=width)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n,logb", [(16, 1), (1 << 11, 2)])
def test_running_sum_matches_oracle(device, n, logb):
    r = (99, 100)
    main, aux, final = airs.running_trace(n, r, 5)
    per = np.array(airs.PERIODIC, dtype=np.uint64).reshape(4, 1)
    table = O.coset_lde_batch(per, 1, pow(G7, n // 4, P), bit_reverse=False)
    got, want = _run_both(device, airs.running_program(), main, n, logb, 1, (13, 14), aux=aux, periodic=table,
                          ext=[r[0], r[1], final[0], final[1]], seed=3)
    assert np.array_equal(got, want)


def test_quotient_pipeline_from_committed_trees(device):
    # prove_multi's quotient steps on the GPU: commit the trace, evaluate the constraints on
    # the committed LDE in place, divide by Z_H, commit the quotient; the root must equal the
    # oracle's for the same steps
    n, logb, log_d = 1 << 10, 2, 1
    trace, pub = airs.fibonacci_trace(n, 1, 2)
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    tree = G.commit_traces(cfg, [trace], logb)
    ptr, h, w = tree.matrix(0)
    main = G.DeviceMatrix(device, ptr, h, w, owned=False)
    acc = device.alloc_matrix(n, 2 << log_d)  # n*d extension elements, read by commit_quotient as n x 2d
    G.cyclic_extend_and_scale(acc, 0, n << log_d, (0, 0))  # an empty accumulator becomes zeros
    alpha = (17, 19)
    G.eval_constraints(airs.fibonacci_program(), main, n, log_d, G7, alpha, acc, publics=pub)
    G.divide_by_vanishing(acc, n, 1 << log_d, G7)
    qt = G.commit_quotient(cfg, acc, 1 << log_d, logb)
    num = O.eval_constraints(airs.fibonacci_program().ops, O.coset_lde_batch(trace, logb, G7), n, log_d, G7, alpha,
                             publics=pub)
    q = O.divide_by_vanishing(num, 1 << log_d, G7)
    assert np.array_equal(acc.download().reshape(-1, 2), q)
    assert qt.root().tolist() == O.commit_quotient(O.HASH_P2, q, 1 << log_d, logb).root().tolist()
    qt.free()
    acc.free()
    tree.free()


def test_miden_dummy_at_the_reference_binary_shape(device):
    # lifted_miden.rs: trace 2 is 2^19 x 20 with log_blowup 3 and quotient degree 8 (2^22 points)
    n, width, logb, log_d = 1 << 19, 20, 3, 3
    trace = airs.miden_trace(n, width, 42)
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    tree = G.commit_traces(cfg, [trace], logb)
    ptr, h, w = tree.matrix(0)
    main = G.DeviceMatrix(device, ptr, h, w, owned=False)
    acc = device.alloc_matrix(n << log_d, 2)
    G.cyclic_extend_and_scale(acc, 0, n << log_d, (0, 0))
    print("eval_constraints: DummyMidenAir 2^19 x 20, 2^22 points", flush=True)
    G.eval_constraints(airs.miden_program(), main, n, log_d, G7, (5, 8), acc)
    got = acc.download().reshape(-1, 2)
    want = O.eval_constraints(airs.miden_program().ops, main.download(), n, log_d, G7, (5, 8))
    assert np.array_equal(got, want)
    acc.free()
    tree.free()


@pytest.mark.parametrize("seed,n,logb,log_d,nops", [(1, 64, 2, 2, 400), (2, 1 << 10, 1, 1, 1500), (3, 4, 1, 0, 200),
                                                     (4, 8, 1, 1, 300), (5, 2, 3, 3, 300), (6, 1 << 9, 3, 3, 3000)])
def test_random_programs_match_oracle(device, seed, n, logb, log_d, nops):
    # every input kind, mixed base/extension arithmetic, long live ranges; the domains
    # run from 4 points (smaller than one 32 x 32 tile) to 2^12
    main_w, aux_w, npub, n_ext, per_cols = 7, 4, 3, 2, 2
    prog = airs.random_program(seed, nops, main_w, aux_w // 2, npub, n_ext, per_cols)
    rng = np.random.default_rng(seed)
    main = rng.integers(0, P, size=(n, main_w), dtype=np.uint64)
    aux = rng.integers(0, P, size=(n, aux_w), dtype=np.uint64)
    prow = max(1, min(n, 4))
    per = rng.integers(0, P, size=(prow, per_cols), dtype=np.uint64)
    table = O.coset_lde_batch(per, log_d, pow(G7, n // prow, P), bit_reverse=False)
    pub = [int(v) for v in rng.integers(0, P, size=npub, dtype=np.uint64)]
    ext = [int(v) for v in rng.integers(0, P, size=2 * n_ext, dtype=np.uint64)]
    got, want = _run_both(device, prog, main, n, logb, log_d, tuple(int(v) for v in rng.integers(0, P, size=2, dtype=np.uint64)), aux=aux,
                          periodic=table, publics=pub, ext=ext, seed=seed)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("live", [70, 300])
def test_many_live_values_match_oracle(device, live):
    n = 256
    trace = np.random.default_rng(live).integers(0, P, size=(n, 3), dtype=np.uint64)
    got, want = _run_both(device, airs.wide_program(live), trace, n, 2, 2, (3, 4), seed=live)
    assert np.array_equal(got, want)


def test_bad_programs_and_shapes_are_rejected(device):
    n = 64
    main = device.alloc_matrix(2 * n, 2)
    main.fill_random(1)
    acc = device.alloc_matrix(4 * n, 2)  # room for n d points at d = 4 (the short-main case below)
    ops = airs.fibonacci_program().ops
    bad = ops.copy()
    bad[9]["b"] = 50  # a product that reads a value defined later
    for prog, kw, code in [
        (bad, dict(publics=[0, 0, 0]), 1),
        (np.concatenate([ops[:1], ops[1:2]]).copy(), dict(), 6),  # asserts nothing
        (ops, dict(publics=[0]), 1),                       # a public value that is not supplied
    ]:
        with pytest.raises(G.P3GError) as e:
            G.eval_constraints(prog, main, n, 1, G7, (1, 0), acc, **kw)
        assert e.value.code == code, (code, e.value)
    # an operand that names an assertion
    bad = ops.copy()
    bad[12]["b"] = 10
    assert bad[10]["op"] == G.air.ASSERT_ZERO and bad[12]["op"] == G.air.SUB
    with pytest.raises(G.P3GError) as e:
        G.eval_constraints(bad, main, n, 1, G7, (1, 0), acc, publics=[0, 0, 0])
    assert e.value.code == 1
    # a shift inside the subgroup of order n d (Z_H vanishes there)
    for shift in (1, pow(G7, (P - 1) // 2, P)):
        with pytest.raises(G.P3GError) as e:
            G.eval_constraints(ops, main, n, 1, shift, (1, 0), acc, publics=[0, 0, 0])
        assert e.value.code == 1
    # inputs the program reads but the call does not supply
    with pytest.raises(G.P3GError) as e:
        G.eval_constraints(airs.running_program(), main, n, 1, G7, (1, 0), acc, ext_inputs=[1, 2, 3, 4])
    assert e.value.code == 1  # no aux matrix
    # quotient degree and domain limits: rejected on the host before anything touches the
    # accumulator (a view that claims enough rows passes the binding's own size check)
    claimed = G.DeviceMatrix(device, acc.ptr, n << 26, 2, owned=False)
    for log_d in (21, 26):
        with pytest.raises(G.P3GError) as e:
            G.eval_constraints(ops, main, n, log_d, G7, (1, 0), claimed, publics=[0, 0, 0])
        assert e.value.code == 8
    # the main matrix must hold the n d points
    with pytest.raises(G.P3GError) as e:
        G.eval_constraints(ops, main, n, 2, G7, (1, 0), acc, publics=[0, 0, 0])
    assert e.value.code == 1
    main.free()
    acc.free()


def test_two_instances_as_prove_multi(device):
    # prove_multi's quotient loop (prover/mod.rs:355-434) over two AIRs in ascending height,
    # all on the device: one commit_traces tree holds both LDEs (the smaller on the nested
    # shift g^(N/n), commit.rs:205-221); the accumulator is extended to each instance's
    # n*D (scaled by beta), each instance's constraints are evaluated in place on its tree
    # matrix with its own coset shift, then Z_H is divided out on the largest quotient
    # coset and the quotient committed.  Every step must equal the oracle's.
    D, logD, logb = 8, 3, 3
    n0, n1 = 1 << 6, 1 << 8
    fib, pub = airs.fibonacci_trace(n0, 3, 5)
    miden = airs.miden_trace(n1, 12, 77)
    alpha, beta = (1234, 5678), (9012, 3456)
    shift0 = pow(G7, n1 // n0, P)
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    tree = G.commit_traces(cfg, [fib, miden], logb)
    mats = [G.DeviceMatrix(device, *tree.matrix(j), owned=False) for j in range(2)]
    acc = device.alloc_matrix(n1 * D, 2)
    G.cyclic_extend_and_scale(acc, 0, n0 * D, beta)
    G.eval_constraints(airs.fibonacci_program(), mats[0], n0, logD, shift0, alpha, acc, publics=pub)
    G.cyclic_extend_and_scale(acc, n0 * D, n1 * D, beta)
    G.eval_constraints(airs.miden_program(), mats[1], n1, logD, G7, alpha, acc)
    G.divide_by_vanishing(acc, n1, D, G7)
    got = acc.download().reshape(-1, 2)
    q = device.upload(got.reshape(n1, 2 * D))
    qt = G.commit_quotient(cfg, q, D, logb)

    lde0 = O.coset_lde_batch(fib, logb, shift0)
    lde1 = O.coset_lde_batch(miden, logb, G7)
    assert np.array_equal(mats[0].download(), lde0) and np.array_equal(mats[1].download(), lde1)
    want = O.cyclic_extend_scale(np.zeros((0, 2), np.uint64), beta, n0 * D)
    want = O.eval_constraints(airs.fibonacci_program().ops, lde0, n0, logD, shift0, alpha, acc=want, publics=pub)
    want = O.cyclic_extend_scale(want, beta, n1 * D)
    want = O.eval_constraints(airs.miden_program().ops, lde1, n1, logD, G7, alpha, acc=want)
    want = O.divide_by_vanishing(want, D, G7)
    assert np.array_equal(got, want)
    assert qt.root().tolist() == O.commit_quotient(O.HASH_P2, want, D, logb).root().tolist()
    for x in (qt, q, acc, tree):
        x.free()


@pytest.mark.parametrize("n,logb,log_d", [(1 << 8, 2, 2), (1 << 12, 1, 1)])
def test_poseidon2_air_matches_oracle(device, n, logb, log_d):
    # a 3,047-op AIR over 248 columns (several 1,024-op device functions), on a valid trace:
    # word for word against the oracle, and numerator / Z_H is of degree <= 2n - 3
    rc, dg = O.p2_constants(), O.p2_diag()
    inp = np.random.default_rng(n).integers(0, P, size=(n, 12), dtype=np.uint64)
    trace = airs.poseidon2_trace(inp, rc, dg)
    prog = airs.poseidon2_program(rc, dg)
    got, want = _run_both(device, prog, trace, n, logb, log_d, (31, 37), seed=n)
    assert np.array_equal(got, want)
    acc = device.upload(np.zeros((n << log_d, 2), np.uint64))
    main = device.upload(O.coset_lde_batch(trace, logb, G7))
    G.eval_constraints(prog, main, n, log_d, G7, (31, 37), acc)
    G.divide_by_vanishing(acc, n, 1 << log_d, G7)
    q = acc.download().reshape(-1, 2)
    co = np.stack([O.ntt(np.ascontiguousarray(q[:, c]), inverse=True) for c in range(2)], axis=1)
    assert not co[2 * n - 2:].any()
    acc.free()
    main.free()


@pytest.mark.parametrize("periods,heights", airs.TINY_CASES)
def test_tiny_air_cases_match_oracle(device, periods, heights):
    # the reference's TinyAir cases (tiny_air.rs) through prove_multi's quotient loop on the
    # device: main traces and aux traces in one tree each (nested coset shifts), every
    # instance evaluated in place with its periodic table, one division by Z_H.  Word for
    # word the oracle's, and the quotient is of degree < 4N (valid traces).
    log_d, logb = 3, 3
    D, N = 1 << log_d, max(heights)
    challenge = (5, 9)
    alpha, beta = (321, 123), (987, 654)
    prog = airs.tiny_program(len(periods))
    traces = [airs.tiny_trace(h, idx + 2, challenge) for idx, h in enumerate(heights)]
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    mtree = G.commit_traces(cfg, [m for m, _ in traces], logb)
    atree = G.commit_traces(cfg, [a for _, a in traces], logb)
    acc = device.alloc_matrix(N * D, 2)
    want = np.zeros((0, 2), np.uint64)
    m = 0
    for idx, (h, (main, aux)) in enumerate(zip(heights, traces)):
        shift = pow(G7, N // h, P)
        G.cyclic_extend_and_scale(acc, m, h * D, beta)
        want = O.cyclic_extend_scale(want, beta, h * D)
        per = airs.tiny_periodic_table(periods, h, log_d, shift) if periods else None
        dm = G.DeviceMatrix(device, *mtree.matrix(idx), owned=False)
        da = G.DeviceMatrix(device, *atree.matrix(idx), owned=False)
        G.eval_constraints(prog, dm, h, log_d, shift, alpha, acc, aux=da, periodic=per, publics=[idx + 2],
                           ext_inputs=list(challenge))
        want = O.eval_constraints(prog.ops, O.coset_lde_batch(main, logb, shift), h, log_d, shift, alpha, acc=want,
                                  aux=O.coset_lde_batch(aux, logb, shift), periodic=per, publics=[idx + 2],
                                  ext_inputs=list(challenge))
        m = h * D
    G.divide_by_vanishing(acc, N, D, G7)
    want = O.divide_by_vanishing(want, D, G7)
    got = acc.download().reshape(-1, 2)
    assert np.array_equal(got, want)
    co = np.stack([O.ntt(np.ascontiguousarray(got[:, c]), inverse=True) for c in range(2)], axis=1)
    assert not co[4 * N:].any()
    for x in (acc, mtree, atree):
        x.free()


@pytest.mark.parametrize("n", [8, 1 << 12])
def test_bus_air_matches_oracle(device, n):
    # the reference's BusTestAir (bus_test.rs, bus_identity_check's values) on committed
    # main and aux trees: word for word the oracle's, and the quotient is of degree < 4n
    log_d, logb = 3, 3
    main, aux, vals = airs.bus_trace(n, 2, 42, 67, (11, 13), (17, 19))
    ext = [11, 13, 17, 19, vals[0][0], vals[0][1], vals[1][0], vals[1][1]]
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    mt, at = G.commit_traces(cfg, [main], logb), G.commit_traces(cfg, [aux], logb)
    dm = G.DeviceMatrix(device, *mt.matrix(0), owned=False)
    da = G.DeviceMatrix(device, *at.matrix(0), owned=False)
    acc = device.upload(np.zeros((n << log_d, 2), np.uint64))
    G.eval_constraints(airs.bus_program(), dm, n, log_d, G7, (5, 6), acc, aux=da, publics=[2, 42, 67], ext_inputs=ext)
    want = O.eval_constraints(airs.bus_program().ops, O.coset_lde_batch(main, logb, G7), n, log_d, G7, (5, 6),
                              aux=O.coset_lde_batch(aux, logb, G7), publics=[2, 42, 67], ext_inputs=ext)
    got = acc.download().reshape(-1, 2)
    assert np.array_equal(got, want)
    G.divide_by_vanishing(acc, n, 1 << log_d, G7)
    q = acc.download().reshape(-1, 2)
    co = np.stack([O.ntt(np.ascontiguousarray(q[:, c]), inverse=True) for c in range(2)], axis=1)
    assert not co[4 * n:].any()
    for x in (acc, mt, at):
        x.free()


def test_padding_air_multi_trace_matches_oracle(device):
    # the reference's multi_trace_with_aux_padding (multi_aux_alignment.rs:118-150): widths one
    # more than the LMCS alignment (main 9 columns, aux 9 extension columns), instances of
    # heights 8 and 16 through the quotient loop, read in place from the aligned trees
    cfg = G.LmcsConfig(device, G.HASH_POSEIDON2)
    width = aux_width = cfg.alignment() + 1
    log_d, logb = 1, 1
    D, N = 1 << log_d, 16
    challenge, alpha, beta = (3, 4), (8, 9), (10, 11)
    heights = [8, 16]
    traces = [airs.padding_trace(h, width, aux_width, idx + 2, challenge) for idx, h in enumerate(heights)]
    mtree = G.commit_traces(cfg, [m for m, _ in traces], logb)
    atree = G.commit_traces(cfg, [a for _, a in traces], logb)
    prog = airs.padding_program()
    acc = device.alloc_matrix(N * D, 2)
    want = np.zeros((0, 2), np.uint64)
    m = 0
    for idx, (h, (main, aux)) in enumerate(zip(heights, traces)):
        shift = pow(G7, N // h, P)
        G.cyclic_extend_and_scale(acc, m, h * D, beta)
        want = O.cyclic_extend_scale(want, beta, h * D)
        dm = G.DeviceMatrix(device, *mtree.matrix(idx), owned=False)
        da = G.DeviceMatrix(device, *atree.matrix(idx), owned=False)
        G.eval_constraints(prog, dm, h, log_d, shift, alpha, acc, aux=da, publics=[idx + 2], ext_inputs=list(challenge))
        want = O.eval_constraints(prog.ops, O.coset_lde_batch(main, logb, shift), h, log_d, shift, alpha, acc=want,
                                  aux=O.coset_lde_batch(aux, logb, shift), publics=[idx + 2], ext_inputs=list(challenge))
        m = h * D
    G.divide_by_vanishing(acc, N, D, G7)
    want = O.divide_by_vanishing(want, D, G7)
    got = acc.download().reshape(-1, 2)
    assert np.array_equal(got, want)
    # constraint degree 2: the quotient has degree < N
    co = np.stack([O.ntt(np.ascontiguousarray(got[:, c]), inverse=True) for c in range(2)], axis=1)
    assert not co[N:].any()
    for x in (acc, mtree, atree):
        x.free()
