"""Example AIRs written against p3gpu.AirProgram (the reference's AirBuilder vocabulary),
with trace generators: the constraint-evaluation tests and the bench use them.

  fibonacci: two columns, a' = b, b' = a + b; a_0, b_0 and the last b are public values
             (degree 2, quotient degree 2);
  miden:     the reference's DummyMidenAir (p3-miden-lifted-examples/src/miden.rs:38-64):
             local[0] * ... * local[8] == 0 on every row (degree 9, quotient degree 8),
             column 0 zero, the rest random (generate_dummy_trace, :122-144);
  running:   a main column v, a column w = v * p with p a period-4 periodic column, and an
             extension aux column s, the running sum of r * v with r the permutation
             randomness, ending in the permutation value (degree 2);
  tiny:      the reference's TinyAir test AIR (x -> x^4 on main and aux, periodic columns);
  bus:       the reference's BusTestAir (two aux columns, randomness and aux values);
  padding:   the reference's PaddingAir (widths one more than the LMCS alignment);
  poseidon2: a Poseidon2-GL12 permutation per row (3,047 ops over 248 columns).
"""
import numpy as np

from .air import AirProgram

P = (1 << 64) - (1 << 32) + 1
W = 7  # the extension's non-residue: F[x]/(x^2 - 7)


def fibonacci_program():
    p = AirProgram()
    a, b = p.main(0), p.main(1)
    a2, b2 = p.main(0, next=True), p.main(1, next=True)
    first, last, trans = p.is_first_row(), p.is_last_row(), p.is_transition()
    p.assert_zero(first * (a - p.public(0)))
    p.assert_zero(first * (b - p.public(1)))
    p.assert_zero(trans * (a2 - b))
    p.assert_zero(trans * (b2 - a - b))
    p.assert_zero(last * (b - p.public(2)))
    return p


def fibonacci_trace(n, a0=1, b0=1):
    t = np.zeros((n, 2), np.uint64)
    a, b = a0, b0
    for i in range(n):
        t[i] = (a, b)
        a, b = b, (a + b) % P
    return t, [a0, b0, int(t[n - 1, 1])]


def miden_program():
    p = AirProgram()
    prod = p.main(0)
    for j in range(1, 9):
        prod = prod * p.main(j)
    p.assert_zero(prod)
    return p


def miden_trace(n, width, seed):
    rng = np.random.default_rng(seed)
    t = rng.integers(0, P, size=(n, width), dtype=np.uint64)
    t[:, 0] = 0
    return t


PERIODIC = [3, 5, 7, 11]


def running_program():
    p = AirProgram()
    v, w, v2 = p.main(0), p.main(1), p.main(0, next=True)
    s, s2 = p.aux(0), p.aux(0, next=True)
    r, final = p.ext_input(0), p.ext_input(1)
    p.assert_zero(w - v * p.periodic(0))
    p.assert_zero_ext(p.is_first_row() * (s - r * v))
    p.assert_zero_ext(p.is_transition() * (s2 - s - r * v2))
    p.assert_zero_ext(p.is_last_row() * (s - final))
    return p


def ef_mul(a, b):
    return ((a[0] * b[0] + W * a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def running_trace(n, r, seed):
    """main (n x 2), aux (n x 2: s as c0, c1), final value"""
    rng = np.random.default_rng(seed)
    v = rng.integers(0, P, size=n, dtype=np.uint64)
    main = np.zeros((n, 2), np.uint64)
    aux = np.zeros((n, 2), np.uint64)
    s = (0, 0)
    for i in range(n):
        main[i] = (int(v[i]), int(v[i]) * PERIODIC[i % 4] % P)
        rv = ef_mul(r, (int(v[i]), 0))
        s = ((s[0] + rv[0]) % P, (s[1] + rv[1]) % P)
        aux[i] = s
    return main, aux, s


def random_program(seed, nops, main_w, aux_cols=0, npub=0, n_ext=0, per_cols=0, selectors=True):
    """A random straight-line program over every kind of input: leaves (main and aux
    columns on both rows, periodic columns, public values, extension inputs, constants,
    selectors) and ADD/SUB/MUL/NEG of earlier values (mixed base and extension operands),
    with an assertion about every eighth op."""
    rng = np.random.default_rng(seed)
    p = AirProgram()
    vals = []
    leaves = [lambda: p.main(int(rng.integers(main_w)), next=bool(rng.integers(2))),
              lambda: p.const(int(rng.integers(0, P, dtype=np.uint64)))]
    if aux_cols:
        leaves.append(lambda: p.aux(int(rng.integers(aux_cols)), next=bool(rng.integers(2))))
    if npub:
        leaves.append(lambda: p.public(int(rng.integers(npub))))
    if n_ext:
        leaves.append(lambda: p.ext_input(int(rng.integers(n_ext))))
    if per_cols:
        leaves.append(lambda: p.periodic(int(rng.integers(per_cols))))
    if selectors:
        leaves += [p.is_first_row, p.is_last_row, p.is_transition]
    while len(p._ops) < nops:
        if len(vals) < 4 or rng.random() < 0.3:
            vals.append(leaves[int(rng.integers(len(leaves)))]())
            continue
        # operands mostly from the last 32 values (as an AIR's constraints use their own
        # subexpressions), one in twenty from anywhere earlier (long live ranges)
        def pick():
            far = rng.random() < 0.05
            return vals[int(rng.integers(len(vals)))] if far else vals[-1 - int(rng.integers(min(32, len(vals))))]
        a, b = pick(), pick()
        kind = int(rng.integers(4))
        vals.append(a + b if kind == 0 else a - b if kind == 1 else a * b if kind == 2 else -a)
        if rng.random() < 0.125:
            (p.assert_zero if rng.integers(2) else p.assert_zero_ext)(vals[-1])
    p.assert_zero(vals[-1])
    return p


def wide_program(live):
    """`live` values alive at once: a product over main(0) + k for k < live, formed
    only after every factor exists."""
    p = AirProgram()
    vals = [p.main(0) + k for k in range(live)]
    total = vals[0]
    for v in vals[1:]:
        total = total * v
    p.assert_zero(total)
    return p


# --------------------------------------------------------------------------- Poseidon2-GL12
P2_FULL, P2_PARTIAL, P2_W = 4, 22, 12  # half of the full rounds, partial rounds, width
P2_COLS = P2_W + 2 * P2_FULL * 2 * P2_W + 2 * P2_PARTIAL  # 248


def _p2_columns():
    """Column offsets: inputs | per beginning full round: x^3 registers, post-round state |
    per partial round: x^3 register, lane-0 S-box output | per ending full round: as the
    beginning."""
    begin = [P2_W + 2 * P2_W * r for r in range(P2_FULL)]
    part0 = P2_W + 2 * P2_W * P2_FULL
    partial = [part0 + 2 * r for r in range(P2_PARTIAL)]
    end0 = part0 + 2 * P2_PARTIAL
    end = [end0 + 2 * P2_W * r for r in range(P2_FULL)]
    return begin, partial, end


def _m4(p, x):
    a, b, c, d = x
    ab, cd = a + b, c + d
    abcd = ab + cd
    return [abcd + a + b + b, abcd + b + c + c, abcd + c + d + d, abcd + d + a + a]


def _external(p, s):
    """circ(2 M4, M4, M4) on 12 lanes: M4 on each block of 4, then every lane plus the sum of
    its position across the blocks"""
    y = _m4(p, s[0:4]) + _m4(p, s[4:8]) + _m4(p, s[8:12])
    sums = [y[k] + y[4 + k] + y[8 + k] for k in range(4)]
    return [y[i] + sums[i & 3] for i in range(12)]


def _internal(p, s, diag):
    """diag(d) + 1 1^T"""
    total = s[0]
    for v in s[1:]:
        total = total + v
    return [s[i] * p.const(int(diag[i])) + total for i in range(12)]


def poseidon2_program(rc, diag):
    """One Poseidon2-GL12 permutation per row (S-box x^7, 8 full and 22 partial rounds, the
    external layer circ(2 M4, M4, M4) first and after every full round, the internal layer
    diag(d) + 1 1^T after every partial round), shaped like Plonky3's Poseidon2Air (the
    reference's poseidon2.rs instantiates it over BabyBear): one x^3 register per S-box,
    the state stored after every full round and the lane-0 S-box output of every partial
    round, the other partial-round lanes carried as expressions.  Every constraint has
    degree 3: register = x^3, output = register^2 * x.  rc: 118 round constants (4 x 12
    beginning, 4 x 12 ending, 22 partial), diag: the internal diagonal."""
    p = AirProgram()
    begin, partial, end = _p2_columns()
    rc = [int(v) for v in rc]
    s = _external(p, [p.main(i) for i in range(P2_W)])

    def full(s, base, consts):
        ys = []
        for i in range(P2_W):
            x = s[i] + p.const(consts[i])
            cube = p.main(base + i)
            p.assert_zero(cube - x * x * x)
            ys.append(cube * cube * x)
        out = _external(p, ys)
        post = [p.main(base + P2_W + i) for i in range(P2_W)]
        for i in range(P2_W):
            p.assert_zero(post[i] - out[i])
        return post

    for r in range(P2_FULL):
        s = full(s, begin[r], rc[P2_W * r:P2_W * (r + 1)])
    for r in range(P2_PARTIAL):
        x = s[0] + p.const(rc[2 * P2_W * P2_FULL + r])
        cube, out = p.main(partial[r]), p.main(partial[r] + 1)
        p.assert_zero(cube - x * x * x)
        p.assert_zero(out - cube * cube * x)
        s = _internal(p, [out] + s[1:], diag)
    off = P2_W * P2_FULL
    for r in range(P2_FULL):
        s = full(s, end[r], rc[off + P2_W * r:off + P2_W * (r + 1)])
    return p


def poseidon2_trace(inputs, rc, diag):
    """The trace of poseidon2_program for rows of 12 inputs (n x 12); the state after the
    last round is the permutation of the row's inputs."""
    inputs = np.asarray(inputs, dtype=np.uint64)
    n = inputs.shape[0]
    t = np.zeros((n, P2_COLS), dtype=object)
    rc = [int(v) for v in rc]
    d = [int(v) for v in diag]
    s = [inputs[:, i].astype(object) for i in range(P2_W)]
    t[:, :P2_W] = inputs.astype(object)

    def m4(x):
        a, b, c, e = x
        abcd = a + b + c + e
        return [abcd + a + 2 * b, abcd + b + 2 * c, abcd + c + 2 * e, abcd + e + 2 * a]

    def ext(s):
        y = m4(s[0:4]) + m4(s[4:8]) + m4(s[8:12])
        sums = [y[k] + y[4 + k] + y[8 + k] for k in range(4)]
        return [(y[i] + sums[i & 3]) % P for i in range(12)]

    begin, partial, end = _p2_columns()
    s = ext(s)

    def full(s, base, consts):
        ys = []
        for i in range(P2_W):
            x = (s[i] + consts[i]) % P
            cube = x * x % P * x % P
            t[:, base + i] = cube
            ys.append(cube * cube % P * x % P)
        s = ext(ys)
        for i in range(P2_W):
            t[:, base + P2_W + i] = s[i]
        return s

    for r in range(P2_FULL):
        s = full(s, begin[r], rc[P2_W * r:P2_W * (r + 1)])
    for r in range(P2_PARTIAL):
        x = (s[0] + rc[2 * P2_W * P2_FULL + r]) % P
        cube = x * x % P * x % P
        out = cube * cube % P * x % P
        t[:, partial[r]] = cube
        t[:, partial[r] + 1] = out
        s = [out] + s[1:]
        total = sum(s) % P
        s = [(s[i] * d[i] + total) % P for i in range(12)]
    off = P2_W * P2_FULL
    for r in range(P2_FULL):
        s = full(s, end[r], rc[off + P2_W * r:off + P2_W * (r + 1)])
    return t.astype(np.uint64)


def poseidon2_output_columns():
    """the columns holding the permutation's output (the last round's post state)"""
    _, _, end = _p2_columns()
    return list(range(end[-1] + P2_W, end[-1] + 2 * P2_W))


# --------------------------------------------------------------------------- TinyAir
def tiny_program(num_periodic):
    """The reference's TinyAir (p3-miden-lifted-stark/tests/tiny_air.rs:27-118), over
    Goldilocks: main[0] starts at public value 0 and each row is the previous one^4; every
    periodic column (pattern [1, 0, ..., 0, 1] per period) is 1 on the first and the last
    row; the aux column starts at the permutation randomness and each row is the previous
    one^4.  Constraints in the AIR's emission order."""
    p = AirProgram()
    local, nxt = p.main(0), p.main(0, next=True)
    first, last, trans = p.is_first_row(), p.is_last_row(), p.is_transition()
    p.assert_zero(first * (local - p.public(0)))
    sq = local * local
    p.assert_zero(trans * (nxt - sq * sq))
    for c in range(num_periodic):
        v = p.periodic(c)
        p.assert_zero(first * (v - 1))
        p.assert_zero(last * (v - 1))
    a, a2 = p.aux(0), p.aux(0, next=True)
    p.assert_zero_ext(first * (a - p.ext_input(0)))
    sq = a * a
    p.assert_zero_ext(trans * (a2 - sq * sq))
    return p


def tiny_periodic_columns(periods):
    """TinyAir::new's periodic columns: 1 at the first and last position of each period."""
    cols = []
    for q in periods:
        c = [0] * q
        c[0] = c[q - 1] = 1
        cols.append(c)
    return cols


def tiny_trace(n, start, challenge):
    """main (n x 1): start^(4^i); aux (n x 2, c0 and c1): challenge^(4^i) in F[x]/(x^2 - 7)
    (TinyAuxBuilder, tiny_air.rs:121-140)."""
    main = np.zeros((n, 1), np.uint64)
    aux = np.zeros((n, 2), np.uint64)
    x, e = int(start) % P, (int(challenge[0]) % P, int(challenge[1]) % P)
    for i in range(n):
        main[i, 0] = x
        aux[i] = e
        x = pow(x, 4, P)
        e = ef_mul(e, e)
        e = ef_mul(e, e)
    return main, aux


# --------------------------------------------------------------------------- BusTestAir
def bus_program():
    """The reference's BusTestAir (p3-miden-lifted-stark/tests/bus_test.rs:43-150) over
    Goldilocks: main[0] starts at public value 0 and each row is the previous one^4; two aux
    columns with two randomness values c0, c1 and two aux values: on the first row
    aux0 (pi_0 + c0) = 1 and aux1 = pi_1 + c1 (pi_0, pi_1 public values 1 and 2), both
    columns constant, and equal to the aux values on the last row.  Extension inputs:
    c0, c1, then the aux values."""
    p = AirProgram()
    pv0, pv1, pv2 = p.public(0), p.public(1), p.public(2)
    local, nxt = p.main(0), p.main(0, next=True)
    first, last, trans = p.is_first_row(), p.is_last_row(), p.is_transition()
    p.assert_zero(first * (local - pv0))
    sq = local * local
    p.assert_zero(trans * (nxt - sq * sq))
    c0, c1 = p.ext_input(0), p.ext_input(1)
    av0, av1 = p.ext_input(2), p.ext_input(3)
    a0, a1 = p.aux(0), p.aux(1)
    p.assert_zero_ext(first * (a0 * (pv1 + c0) - 1))
    p.assert_zero_ext(first * (a1 - (pv2 + c1)))
    p.assert_zero_ext(trans * (p.aux(0, next=True) - a0))
    p.assert_zero_ext(trans * (p.aux(1, next=True) - a1))
    p.assert_zero_ext(last * (a0 - av0))
    p.assert_zero_ext(last * (a1 - av1))
    return p


def ef_inv(a):
    """1 / (a0 + a1 X) in F[x]/(x^2 - 7): (a0 - a1 X) / (a0^2 - 7 a1^2)"""
    norm = (a[0] * a[0] - W * a[1] * a[1]) % P
    ni = pow(norm, P - 2, P)
    return (a[0] * ni % P, (P - a[1]) % P * ni % P)


def bus_trace(n, start, pi0, pi1, c0, c1):
    """main (n x 1): start^(4^i); aux (n x 4): the constant columns 1/(pi_0 + c0) and
    pi_1 + c1 as (c0, c1) pairs (BusTestAuxBuilder, bus_test.rs:154-183); the aux values."""
    main = np.zeros((n, 1), np.uint64)
    x = int(start) % P
    for i in range(n):
        main[i, 0] = x
        x = pow(x, 4, P)
    col0 = ef_inv(((int(pi0) + c0[0]) % P, c0[1] % P))
    col1 = ((int(pi1) + c1[0]) % P, c1[1] % P)
    aux = np.tile(np.array([col0[0], col0[1], col1[0], col1[1]], dtype=np.uint64), (n, 1))
    return main, aux, [col0, col1]


# --------------------------------------------------------------------------- PaddingAir
def padding_program():
    """The reference's PaddingAir (p3-miden-lifted-stark/tests/multi_aux_alignment.rs:23-80):
    main[0] equals public value 0 on the first row and is constant; aux column 0 equals the
    randomness on the first row and is constant.  The other columns (the AIR is made one
    wider than the LMCS alignment) are unconstrained."""
    p = AirProgram()
    first, trans = p.is_first_row(), p.is_transition()
    local = p.main(0)
    p.assert_zero(first * (local - p.public(0)))
    p.assert_zero(trans * (p.main(0, next=True) - local))
    a = p.aux(0)
    p.assert_zero_ext(first * (a - p.ext_input(0)))
    p.assert_zero_ext(trans * (p.aux(0, next=True) - a))
    return p


def padding_trace(n, width, aux_width, start, challenge):
    """main (n x width): start then zeros on every row; aux (n x 2 aux_width): the
    randomness then zeros on every row (PaddingAuxBuilder, multi_aux_alignment.rs:82-102)."""
    main = np.zeros((n, width), np.uint64)
    main[:, 0] = int(start) % P
    aux = np.zeros((n, 2 * aux_width), np.uint64)
    aux[:, 0], aux[:, 1] = int(challenge[0]) % P, int(challenge[1]) % P
    return main, aux
