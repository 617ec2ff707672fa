"""AIR constraints as a straight-line program for the GPU constraint evaluator.

``AirProgram`` plays the role of the reference's AIR builders (the AirBuilder /
ExtensionBuilder / PermutationAirBuilder / PeriodicAirBuilder vocabulary of
p3-miden-lifted-air, used by ProverConstraintFolder in p3-miden-lifted-stark/src/
prover/constraints/folder.rs): an AIR's ``eval`` written against it records every
value it computes as one op, in order, and every ``assert_zero`` /
``assert_zero_ext`` as a constraint.  The resulting op list (``program.ops``) is
what ``p3g_eval_constraints`` takes (include/p3gpu.h, p3g_air_op).

    p = AirProgram()
    a, b = p.main(0), p.main(1)
    a2, b2 = p.main(0, next=True), p.main(1, next=True)
    p.assert_zero(p.is_first_row() * (a - p.public(0)))
    p.assert_zero(p.is_transition() * (b2 - a - b))
"""
from __future__ import annotations

import numpy as np

P = 0xFFFFFFFF00000001

CONST, MAIN, AUX, PERIODIC, PUBLIC, EXT_INPUT, IS_FIRST_ROW, IS_LAST_ROW, IS_TRANSITION = range(9)
ADD, SUB, MUL, NEG, ASSERT_ZERO, ASSERT_ZERO_EXT = range(9, 15)

AIR_OP_DTYPE = np.dtype([("op", "<u4"), ("a", "<u4"), ("b", "<u4"), ("reserved", "<u4"), ("imm", "<u8")])


class Expr:
    """A value of the program (base or extension field); arithmetic records new ops."""

    __slots__ = ("prog", "slot")

    def __init__(self, prog: "AirProgram", slot: int):
        self.prog, self.slot = prog, slot

    def _other(self, x) -> "Expr":
        return x if isinstance(x, Expr) else self.prog.const(x)

    def __add__(self, x):
        return self.prog._op(ADD, self.slot, self._other(x).slot)

    def __radd__(self, x):
        return self._other(x).__add__(self)

    def __sub__(self, x):
        return self.prog._op(SUB, self.slot, self._other(x).slot)

    def __rsub__(self, x):
        return self._other(x).__sub__(self)

    def __mul__(self, x):
        return self.prog._op(MUL, self.slot, self._other(x).slot)

    def __rmul__(self, x):
        return self._other(x).__mul__(self)

    def __neg__(self):
        return self.prog._op(NEG, self.slot)


class AirProgram:
    def __init__(self):
        self._ops = []
        self.num_constraints = 0

    def _op(self, op: int, a: int = 0, b: int = 0, imm: int = 0) -> Expr:
        self._ops.append((op, a, b, 0, imm))
        return Expr(self, len(self._ops) - 1)

    # values
    def const(self, v: int) -> Expr:
        return self._op(CONST, imm=int(v) % P)

    def main(self, col: int, next: bool = False) -> Expr:
        """main().current_slice()[col] (or next_slice()[col])"""
        return self._op(MAIN, col, 1 if next else 0)

    def aux(self, col: int, next: bool = False) -> Expr:
        """permutation() window, extension column col"""
        return self._op(AUX, col, 1 if next else 0)

    def periodic(self, col: int) -> Expr:
        return self._op(PERIODIC, col)

    def public(self, i: int) -> Expr:
        return self._op(PUBLIC, i)

    def ext_input(self, i: int) -> Expr:
        """permutation_randomness()[i], then permutation_values() after the randomness"""
        return self._op(EXT_INPUT, i)

    def is_first_row(self) -> Expr:
        return self._op(IS_FIRST_ROW)

    def is_last_row(self) -> Expr:
        return self._op(IS_LAST_ROW)

    def is_transition(self) -> Expr:
        return self._op(IS_TRANSITION)

    # constraints
    def assert_zero(self, e) -> None:
        e = e if isinstance(e, Expr) else self.const(e)
        self._op(ASSERT_ZERO, e.slot)
        self.num_constraints += 1

    def assert_zero_ext(self, e) -> None:
        e = e if isinstance(e, Expr) else self.const(e)
        self._op(ASSERT_ZERO_EXT, e.slot)
        self.num_constraints += 1

    @property
    def ops(self) -> np.ndarray:
        return np.array(self._ops, dtype=AIR_OP_DTYPE)
