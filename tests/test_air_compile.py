"""Constraint programs compiled to gfx950 kernels (p3g_air_compile) on the host, no GPU:
the example AIRs and large random programs compile, a second request is served from the
process cache, and malformed programs are rejected before any compile."""
import time

import numpy as np
import pytest

import p3gpu as G
import airs


@pytest.mark.parametrize("make", [airs.fibonacci_program, airs.miden_program, airs.running_program])
def test_example_airs_compile(make):
    assert G.air_compile(make()) > 1000


def test_random_and_wide_programs_compile_and_cache():
    p = airs.random_program(11, 2000, 9, 2, 3, 2, 2)
    size = G.air_compile(p)
    assert size > 1000
    t0 = time.perf_counter()
    assert G.air_compile(p) == size  # the same program again: no compile
    assert time.perf_counter() - t0 < 0.05
    assert G.air_compile(airs.wide_program(300)) > 1000


def test_malformed_programs_are_rejected():
    ops = airs.fibonacci_program().ops
    cases = []
    bad = ops.copy()
    bad[9]["b"] = 50  # reads a later value
    cases.append((bad, 1))
    bad = ops.copy()
    bad[12]["b"] = 10  # reads an assertion
    cases.append((bad, 1))
    bad = ops.copy()
    bad[0]["b"] = 2  # row neither current nor next
    cases.append((bad, 1))
    bad = ops.copy()
    bad[3]["op"] = 99
    cases.append((bad, 1))
    c = G.AirProgram()
    c.const(5)
    cases.append((c.ops, 6))  # asserts nothing
    bad = c.ops.copy()
    bad[0]["imm"] = airs.P  # non-canonical constant
    cases.append((bad, 1))
    cases.append((ops[:0], 6))  # empty
    for prog, code in cases:
        with pytest.raises(G.P3GError) as e:
            G.air_compile(prog)
        assert e.value.code == code, (code, e.value)
    with pytest.raises(TypeError):
        G.air_compile(np.zeros(4, np.uint64))
