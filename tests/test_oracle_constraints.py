"""Pins of the oracle's constraint evaluation (evaluate_constraints_into,
p3-miden-lifted-stark/src/prover/constraints/mod.rs:63-179), without the reference:

  - by hand: the Fibonacci AIR's five constraints written out directly in big
    integers (points, selectors, current and next rows, powers of alpha), not
    through the program interpreter, at several points of gJ;
  - by degree: for a valid trace every constraint vanishes on the trace domain H,
    so numerator / Z_H (divide_by_vanishing) is a polynomial of the AIR's
    quotient degree; its coefficients above that degree, read off an inverse
    DFT over the coset, must be zero.  A wrong next-row step, a wrong selector,
    a wrong row order or a wrong coset would leave a rational function of full
    degree.  One corrupted trace value makes the same check fail.
"""
import numpy as np
import pytest

import oracle as O
import airs

P = airs.P
G7 = 7


def _lde(trace, log_blowup, shift=G7):
    return O.coset_lde_batch(np.asarray(trace, dtype=np.uint64), log_blowup, shift)  # bit-reversed rows


def _bitrev(i, bits):
    return int(format(i, f"0{bits}b")[::-1], 2) if bits else 0


def _periodic_table(n, log_d, shift=G7):
    """PeriodicLde::build (periodic.rs:49-81): the period-4 column's coset LDE on the
    max-period subgroup's shift, natural order, 4 * 2^log_d rows."""
    per = np.array(airs.PERIODIC, dtype=np.uint64).reshape(4, 1)
    pshift = pow(shift, n // 4, P)
    return O.coset_lde_batch(per, log_d, pshift, bit_reverse=False)


def _coset_coeffs(vals_ef, shift):
    """coefficients (times shift^k) of the polynomial taking vals on shift*<w>, per coordinate"""
    out = []
    for c in range(2):
        v = np.ascontiguousarray(vals_ef[:, c])
        out.append(O.ntt(v, inverse=True))
    return np.stack(out, axis=1)


def _quotient_top_zero(num, n, d, first_zero):
    q = O.divide_by_vanishing(num, d, G7)
    co = _coset_coeffs(q, G7)
    return bool(np.all(co[first_zero:] == 0))


def test_fibonacci_by_hand_at_points():
    n, logb, log_d = 16, 2, 1
    trace, pub = airs.fibonacci_trace(n, 2, 3)
    lde = _lde(trace, logb)
    prog = airs.fibonacci_program()
    alpha = (123456789, 987654321)
    acc0 = np.arange(2 * n * 2, dtype=np.uint64).reshape(-1, 2)
    got = O.eval_constraints(prog.ops, lde, n, log_d, G7, alpha, acc=acc0, publics=pub)
    M, logM, d = n << log_d, (n << log_d).bit_length() - 1, 1 << log_d
    wM = O.two_adic_generator(logM)
    wn_inv = pow(O.two_adic_generator(4), P - 2, P)
    for i in (0, 1, 2, 7, M - 2, M - 1):
        x = G7 * pow(wM, i, P) % P
        cur = [int(v) for v in lde[_bitrev(i, logM)]]
        nxt = [int(v) for v in lde[_bitrev((i + d) % M, logM)]]
        zh = (pow(x, n, P) - 1) % P
        first = zh * pow(x - 1, P - 2, P) % P
        last = zh * pow((x - wn_inv) % P, P - 2, P) % P
        trans = (x - wn_inv) % P
        cs = [first * (cur[0] - pub[0]), first * (cur[1] - pub[1]), trans * (nxt[0] - cur[1]),
              trans * (nxt[1] - cur[0] - cur[1]), last * (cur[1] - pub[2])]
        folded = (0, 0)
        K = len(cs)
        for k, c in enumerate(cs):
            apow = (1, 0)
            for _ in range(K - 1 - k):
                apow = airs.ef_mul(apow, alpha)
            t = airs.ef_mul(apow, (c % P, 0))
            folded = ((folded[0] + t[0]) % P, (folded[1] + t[1]) % P)
        want = ((int(acc0[i, 0]) + folded[0]) % P, (int(acc0[i, 1]) + folded[1]) % P)
        assert tuple(int(v) for v in got[i]) == want, i


@pytest.mark.parametrize("n,logb,log_d", [(16, 1, 1), (64, 2, 1), (256, 3, 2)])
def test_fibonacci_quotient_is_low_degree(n, logb, log_d):
    trace, pub = airs.fibonacci_trace(n, 5, 8)
    lde = _lde(trace, logb)
    num = O.eval_constraints(airs.fibonacci_program().ops, lde, n, log_d, G7, (31, 41), publics=pub)
    # boundary quotients have degree n - 2, transition quotients degree 0
    assert _quotient_top_zero(num, n, 1 << log_d, n - 1)
    bad = trace.copy()
    bad[n // 2, 0] = (int(bad[n // 2, 0]) + 1) % P
    num_bad = O.eval_constraints(airs.fibonacci_program().ops, _lde(bad, logb), n, log_d, G7, (31, 41), publics=pub)
    assert not _quotient_top_zero(num_bad, n, 1 << log_d, n - 1)
    wrong_pub = [pub[0], pub[1], (pub[2] + 1) % P]
    num_wp = O.eval_constraints(airs.fibonacci_program().ops, lde, n, log_d, G7, (31, 41), publics=wrong_pub)
    assert not _quotient_top_zero(num_wp, n, 1 << log_d, n - 1)


@pytest.mark.parametrize("n,width,log_d", [(32, 9, 3), (32, 12, 4), (128, 11, 4)])
def test_miden_dummy_quotient_is_low_degree(n, width, log_d):
    # 9 columns of degree n - 1 multiplied: 9n - 9, over Z_H: degree 8n - 9
    trace = airs.miden_trace(n, width, n + width)
    num = O.eval_constraints(airs.miden_program().ops, _lde(trace, log_d), n, log_d, G7, (2, 3))
    assert _quotient_top_zero(num, n, 1 << log_d, 8 * n - 8)
    trace[3, 0] = 1
    num_bad = O.eval_constraints(airs.miden_program().ops, _lde(trace, log_d), n, log_d, G7, (2, 3))
    assert not _quotient_top_zero(num_bad, n, 1 << log_d, 8 * n - 8)


@pytest.mark.parametrize("n,logb", [(16, 1), (64, 2)])
def test_running_sum_with_periodic_and_extension_columns_is_low_degree(n, logb):
    log_d = 1
    r = (1234567, 7654321)
    main, aux, final = airs.running_trace(n, r, n)
    per = _periodic_table(n, log_d)
    ext = [r[0], r[1], final[0], final[1]]
    num = O.eval_constraints(airs.running_program().ops, _lde(main, logb), n, log_d, G7, (5, 6),
                             aux=_lde(aux, logb), periodic=per, ext_inputs=ext)
    # w - v p has degree < 2n; the extension boundary / transition constraints as for Fibonacci
    assert _quotient_top_zero(num, n, 2, n)
    bad = aux.copy()
    bad[n // 3, 1] = (int(bad[n // 3, 1]) + 5) % P
    num_bad = O.eval_constraints(airs.running_program().ops, _lde(main, logb), n, log_d, G7, (5, 6),
                                 aux=_lde(bad, logb), periodic=per, ext_inputs=ext)
    assert not _quotient_top_zero(num_bad, n, 2, n)
    badm = main.copy()
    badm[5, 1] = (int(badm[5, 1]) + 1) % P  # w != v p on one row
    num_badm = O.eval_constraints(airs.running_program().ops, _lde(badm, logb), n, log_d, G7, (5, 6),
                                  aux=_lde(aux, logb), periodic=per, ext_inputs=ext)
    assert not _quotient_top_zero(num_badm, n, 2, n)


def test_malformed_programs_are_rejected():
    lde = _lde(np.zeros((8, 2), np.uint64), 1)
    p = airs.fibonacci_program()
    ops = p.ops.copy()
    ops[3]["a"] = 99  # operand beyond its own index
    with pytest.raises(ValueError):
        O.eval_constraints(ops, lde, 8, 1, G7, (1, 0), publics=[0, 0, 0])
    with pytest.raises(ValueError):  # a public value that is not supplied
        O.eval_constraints(p.ops, lde, 8, 1, G7, (1, 0), publics=[0])
    ops = p.ops.copy()
    ops[12]["b"] = 10  # an operand that is an assertion
    with pytest.raises(ValueError):
        O.eval_constraints(ops, lde, 8, 1, G7, (1, 0), publics=[0, 0, 0])
    for shift in (1, P - 1):  # inside the subgroup of order n d: Z_H vanishes on the coset
        with pytest.raises(ValueError):
            O.eval_constraints(p.ops, lde, 8, 1, shift, (1, 0), publics=[0, 0, 0])


def test_poseidon2_air_trace_is_the_permutation_and_its_quotient_is_low_degree():
    # a Poseidon2-GL12 permutation per row (p3gpu.example_airs.poseidon2_program, 3,047 ops,
    # 236 constraints of degree 3 over 248 columns): the trace's last round is the oracle's
    # permutation of each row, and the quotient of a valid trace has degree <= 2n - 3
    n, logb, log_d = 32, 2, 2
    rc, dg = O.p2_constants(), O.p2_diag()
    rng = np.random.default_rng(7)
    inp = rng.integers(0, P, size=(n, 12), dtype=np.uint64)
    trace = airs.poseidon2_trace(inp, rc, dg)
    out = airs.poseidon2_output_columns()
    for r in (0, 5, n - 1):
        assert trace[r, out].tolist() == O.p2_permute(inp[r]).tolist()
    prog = airs.poseidon2_program(rc, dg)
    num = O.eval_constraints(prog.ops, _lde(trace, logb), n, log_d, G7, (3, 5))
    assert _quotient_top_zero(num, n, 1 << log_d, 2 * n - 2)
    for col in (20, 150, 247):  # an x^3 register, a partial-round output, an output cell
        bad = trace.copy()
        bad[9, col] = (int(bad[9, col]) + 1) % P
        num_bad = O.eval_constraints(prog.ops, _lde(bad, logb), n, log_d, G7, (3, 5))
        assert not _quotient_top_zero(num_bad, n, 1 << log_d, 2 * n - 2), col


def _tiny_quotient(periods, heights, publics=None, log_d=3, logb=3, challenge=(5, 9)):
    """prove_multi's quotient loop (prover/mod.rs:355-432) over TinyAir instances with the
    oracle: nested coset shifts, beta extension, evaluation, one division by Z_H"""
    D, N = 1 << log_d, max(heights)
    prog = airs.tiny_program(len(periods))
    acc = np.zeros((0, 2), np.uint64)
    for idx, h in enumerate(heights):
        main, aux = airs.tiny_trace(h, idx + 2, challenge)
        shift = pow(G7, N // h, P)
        acc = O.cyclic_extend_scale(acc, (987, 654), h * D)
        per = airs.tiny_periodic_table(periods, h, log_d, shift) if periods else None
        pub = [idx + 2] if publics is None else publics[idx]
        acc = O.eval_constraints(prog.ops, _lde(main, logb, shift), h, log_d, shift, (321, 123), acc=acc,
                                 aux=_lde(aux, logb, shift), periodic=per, publics=pub, ext_inputs=list(challenge))
    return O.divide_by_vanishing(acc, D, G7), N, D


@pytest.mark.parametrize("periods,heights", airs.TINY_CASES)
def test_tiny_air_cases_give_a_low_degree_quotient(periods, heights):
    # the reference's TinyAir cases: valid traces, so the combined quotient of the instances
    # (lifted to the largest height) is a polynomial of degree < 4N (constraint degree 5)
    q, N, D = _tiny_quotient(periods, heights)
    co = _coset_coeffs(q, G7)
    assert not co[4 * N:].any()


def test_tiny_air_wrong_public_value_or_periodic_pattern_is_caught():
    # a wrong start value in the second instance
    q, N, D = _tiny_quotient([2, 4], [4, 8], publics=[[2], [4]])
    assert _coset_coeffs(q, G7)[4 * N:].any()
    # a periodic column that is 0 on the last row (period 4 pattern [1, 0, 0, 0])
    orig = airs.tiny_periodic_columns
    try:
        airs.tiny_periodic_columns = lambda periods: [[1] + [0] * (p - 1) for p in periods]
        q, N, D = _tiny_quotient([4], [8])
    finally:
        airs.tiny_periodic_columns = orig
    assert _coset_coeffs(q, G7)[4 * N:].any()


def _bus_quotient(n, publics, ext_fix=None, log_d=3, logb=3):
    main, aux, vals = airs.bus_trace(n, 2, 42, 67, (11, 13), (17, 19))
    ext = [11, 13, 17, 19, vals[0][0], vals[0][1], vals[1][0], vals[1][1]]
    if ext_fix:
        ext = ext_fix(ext)
    num = O.eval_constraints(airs.bus_program().ops, _lde(main, logb), n, log_d, G7, (5, 6), aux=_lde(aux, logb),
                             publics=publics, ext_inputs=ext)
    return O.divide_by_vanishing(num, 1 << log_d, G7)


def test_bus_air_gives_a_low_degree_quotient_and_catches_wrong_inputs():
    # the reference's BusTestAir (bus_test.rs) with its bus_identity_check values: start 2,
    # pi_0 = 42, pi_1 = 67, height 8 (and 64).  Valid: degree < 4n.  A wrong pi_0, a wrong
    # aux value or a wrong challenge breaks a constraint.
    for n in (8, 64):
        assert not _coset_coeffs(_bus_quotient(n, [2, 42, 67]), G7)[4 * n:].any()
    n = 8
    assert _coset_coeffs(_bus_quotient(n, [2, 99, 67]), G7)[4 * n:].any()
    assert _coset_coeffs(_bus_quotient(n, [2, 42, 67], lambda e: e[:6] + [(e[6] + 1) % P, e[7]]), G7)[4 * n:].any()
    assert _coset_coeffs(_bus_quotient(n, [2, 42, 67], lambda e: [12] + e[1:]), G7)[4 * n:].any()
