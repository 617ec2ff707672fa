"""The device field operations of the hashing and NTT kernels on chosen
operands (p3g_field_check), against Python integers.

Random data reaches the rare branches of the v_mad_u64_u32 products (a borrow
after the carry fold, probability ~2^-32) essentially never, so the operands
here include constructed cases for each (carry, borrow) combination.  A
product a*b = L + 2^64 H (H = H0 + 2^32 H1) is computed as T + c 2^64 =
L + (2^32 - 1) H0, then T - H1 with borrow b, then the fold (c - b)(2^32 - 1):
  - c = 0, b = 1: a = b = 3 * 2^48 (L = 0, H0 = 0, H1 = 9);
  - c = 1, b = 1: a = alpha 2^32, b = beta with alpha beta = Y =
    2 + 2^32 (2^32 - 1) + 2^64 y2, so that L = 2^33, H0 = 2^32 - 1 and
    T = 1 < H1 = y2.
The TRY forms (the hashing kernels' shortcut) skip the borrow fold and flag
the wave instead; the test checks the flags against the same model.  Their
value is wrong only for c = 0, b = 1 (by + eps): with c = 1 the shortcut's
+ eps wraps past 2^64, which is -eps again.
"""
import numpy as np
import pytest

P = 0xFFFFFFFF00000001
EPS = (1 << 32) - 1
M32 = (1 << 32) - 1
M64 = (1 << 64) - 1


def _borrows(a: int, b: int, s: int = 0) -> bool:
    x = a * b + s
    lo, hi = x & M64, x >> 64
    t = (lo + EPS * (hi & M32)) & M64
    return t < (hi >> 32)


def _carry(a: int, b: int, s: int = 0) -> bool:
    x = a * b + s
    lo, hi = x & M64, x >> 64
    return lo + EPS * (hi & M32) >= 1 << 64


def _carry_borrow_pair():
    """a, b < 2^64 with both the carry and the borrow (c = 1, b = 1)."""
    for y2 in range(2, 10000):
        y = 2 + (1 << 32) * EPS + (y2 << 64)
        for alpha in range(6, 5000):
            if y % alpha == 0 and y // alpha < 1 << 64:
                a, b = alpha << 32, y // alpha
                assert _carry(a, b) and _borrows(a, b)
                return a, b
    raise AssertionError("no carry-and-borrow operands found")


EDGES = [0, 1, 2, 3, EPS - 1, EPS, EPS + 1, 1 << 32, (1 << 32) + 1, P - 2, P - 1, P, P + 1, M64 - 1, M64,
         1 << 63, (1 << 63) - 1, 3 << 48, EPS << 32, 0x123456789ABCDEF0]


def _operands(rng, n_random=4000):
    a = [x for x in EDGES for _ in EDGES]
    b = [y for _ in EDGES for y in EDGES]
    ca, cb = _carry_borrow_pair()
    a += [3 << 48, ca, cb, 5 << 48]
    b += [3 << 48, cb, ca, 5 << 48]
    a += [int(v) for v in rng.integers(0, M64, n_random, dtype=np.uint64, endpoint=True)]
    b += [int(v) for v in rng.integers(0, M64, n_random, dtype=np.uint64, endpoint=True)]
    return a, b


def test_constructed_cases_hit_every_fold():
    ca, cb = _carry_borrow_pair()
    assert not _carry(3 << 48, 3 << 48) and _borrows(3 << 48, 3 << 48)
    assert _carry(ca, cb) and _borrows(ca, cb)
    rng = np.random.default_rng(1)
    r = [int(v) for v in rng.integers(0, M64, 2000, dtype=np.uint64)]
    pairs = list(zip(r[::2], r[1::2]))
    assert any(_carry(a, b) for a, b in pairs) and any(not _carry(a, b) for a, b in pairs)


@pytest.mark.gpu
@pytest.mark.parametrize("with_sum", [False, True])
def test_products_exact_and_try(device, with_sum):
    rng = np.random.default_rng(7 + with_sum)
    a, b = _operands(rng)
    n = len(a)
    s = [int(v) for v in rng.integers(0, M64, n, dtype=np.uint64, endpoint=True)] if with_sum else [0] * n
    if with_sum:  # the constructed borrow cases keep their sum at zero
        for i in range(len(EDGES) ** 2, len(EDGES) ** 2 + 4):
            s[i] = 0
    A, B, S = (np.array(v, dtype=np.uint64) for v in (a, b, s))
    exact, f0 = device.field_check(2 if with_sum else 0, A, B, S if with_sum else None)
    tried, flags = device.field_check(3 if with_sum else 1, A, B, S if with_sum else None)
    assert not f0.any()
    want_flags = np.zeros_like(flags)
    for i in range(n):
        x = a[i] * b[i] + s[i]
        assert int(exact[i]) % P == x % P, (i, a[i], b[i], s[i])
        br = _borrows(a[i], b[i], s[i])
        if br:
            want_flags[i // 64] |= np.uint64(1 << (i % 64))
            # the skipped fold subtracts eps; with the carry as well, the shortcut's + eps
            # wraps past 2^64 (== -eps) and the value is right anyway
            want = x if _carry(a[i], b[i], s[i]) else x + EPS
            assert int(tried[i]) % P == want % P, i
        else:
            assert int(tried[i]) == int(exact[i]), i
    assert np.array_equal(flags, want_flags)
    assert want_flags.any()  # the constructed cases did borrow


@pytest.mark.gpu
def test_canonical_products(device):
    rng = np.random.default_rng(11)
    can = [x for x in EDGES if x < P] + [int(v) for v in rng.integers(0, P, 3000, dtype=np.uint64)]
    a = np.array([x for x in can[:40] for _ in can[:40]] + can, dtype=np.uint64)
    b = np.array([y for _ in can[:40] for y in can[:40]] + can[::-1], dtype=np.uint64)
    out, _ = device.field_check(4, a, b)
    for x, y, r in zip(a.tolist(), b.tolist(), out.tolist()):
        assert r == x * y % P


@pytest.mark.gpu
@pytest.mark.parametrize("e", range(96))
def test_powers_of_two(device, e):
    rng = np.random.default_rng(100 + e)
    vals = [x for x in EDGES if x < P] + [int(v) for v in rng.integers(0, P, 500, dtype=np.uint64)]
    if 1 <= e <= 32:
        # lo + eps hi in [p, 2^64) without a carry (hi = 1, lo = 2^64 - 2^e j): the reduction
        # after the mad
        for j in sorted({EPS // (1 << e) + 1, 2 * EPS // (1 << e)}):
            v = (1 << (65 - e)) - j
            if v < P and EPS < (j << e) <= 2 * EPS:
                vals.append(v)
    if 33 <= e <= 63:
        # v = h1 2^(96 - e): low product word 0 and h0 = 0, so the borrow fold runs
        vals += [h1 << (96 - e) for h1 in sorted({1, 2, 3, (1 << (e - 32)) - 1}) if h1 < 1 << (e - 32)]
    v = np.array(vals, dtype=np.uint64)
    out, _ = device.field_check(5, v, np.zeros_like(v), e=e)
    for x, r in zip(vals, out.tolist()):
        assert r == x * pow(2, e, P) % P, (e, x)


@pytest.mark.gpu
def test_half_word_sums(device):
    rng = np.random.default_rng(13)
    top = (1 << 40) - 1
    L = [0, 1, top, M32, 1 << 32, top - 1] + [int(v) for v in rng.integers(0, top, 2000, dtype=np.uint64)]
    H = [top, top, top, M32, 0, 1 << 39] + [int(v) for v in rng.integers(0, top, 2000, dtype=np.uint64)]
    out, _ = device.field_check(6, np.array(L, dtype=np.uint64), np.array(H, dtype=np.uint64))
    for lo, hi, r in zip(L, H, out.tolist()):
        assert r % P == (lo + (hi << 32)) % P


@pytest.mark.gpu
@pytest.mark.parametrize("op", [7, 8])
def test_canonical_sums_and_differences(device, op):
    """The NTT butterflies' add_c (64-bit adds and one compare) and sub_c (a borrow
    chain whose borrow adds p back as +1 on the low word and -1 + carry on the
    high word), on every pair of edge values and random canonical words."""
    rng = np.random.default_rng(20 + op)
    can = [x for x in EDGES if x < P] + [(1 << 32) - 2, P - EPS, EPS << 1]
    a = [x for x in can for _ in can] + [int(v) for v in rng.integers(0, P, 4000, dtype=np.uint64)]
    b = [y for _ in can for y in can] + [int(v) for v in rng.integers(0, P, 4000, dtype=np.uint64)]
    out, _ = device.field_check(op, np.array(a, dtype=np.uint64), np.array(b, dtype=np.uint64))
    for x, y, r in zip(a, b, out.tolist()):
        assert r == ((x + y) if op == 7 else (x - y)) % P, (op, x, y)


@pytest.mark.gpu
def test_canonical_reduction(device):
    rng = np.random.default_rng(31)
    vals = EDGES + [P + k for k in range(5)] + [M64 - k for k in range(5)]
    vals += [int(v) for v in rng.integers(P, M64, 500, dtype=np.uint64, endpoint=True)]
    vals += [int(v) for v in rng.integers(0, M64, 2000, dtype=np.uint64, endpoint=True)]
    out, _ = device.field_check(9, np.array(vals, dtype=np.uint64), np.zeros(len(vals), np.uint64))
    for x, r in zip(vals, out.tolist()):
        assert r == x % P, x


@pytest.mark.gpu
def test_field_check_rejects_bad_arguments(device):
    import p3gpu as G
    one = np.ones(4, np.uint64)
    for op, e, n in [(10, 0, 4), (5, 96, 4), (0, 0, 0)]:
        with pytest.raises(G.P3GError):
            device.field_check(op, one[:n], one[:n], e=e)
