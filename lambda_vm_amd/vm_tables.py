"""VM table AIRs of the reference's zkVM, captured with the constraint IR.

EQ (prover/src/tables/eq.rs): `res = (a == b) XOR invert` for 64-bit a, b.
Columns (eq.rs:37-59): a (DWordWL, 2 words), b (DWordWL), invert (bit), res
(bit), diff = a - b (DWordHL, 4 halves), eq (bit), mu (multiplicity).
Transition constraints (eq.rs:240-266):
  0, 1  the ADD carry pair of b + diff = a (constraints/templates.rs:378-418,
        operands :136-185): carry_0 = (b.lo + diff.lo - a.lo) * 2^-32,
        carry_1 = (b.hi + diff.hi + carry_0 - a.hi) * 2^-32, each a bit;
  2     IS_BIT(invert) (templates.rs:283-299);
  3     res - (eq + invert - 2 eq invert);
  4     IS_BIT(mu);
then the LogUp constraints of its bus interactions (eq.rs:152-225): IS_HALF
on each diff half and ZERO[sum of the halves] -> eq, sent with multiplicity
mu; ALU[a (DWordWL), b (DWordWL), EQ + 64 invert, res, 0] received with mu.
Trace generation follows eq.rs:120-145 (one row per operation, padded with
zero rows to a power of two, at least 4).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import constraint_ir as ir
from .air import AirProgram, Boundary, emit_logup_constraints
from .logup import BusInteraction, BusValue, LinearTerm, Multiplicity, Packing, split_interactions

P = ir.P
INV_SHIFT_32 = 18446744065119617026  # 2^-32 mod p (templates.rs:26)

# bus ids (prover/src/tables/types.rs:255-336) and the ALU op code (:562-570)
BUS_IS_HALFWORD, BUS_ZERO, BUS_ALU = 1, 8, 25
ALU_OP_EQ = 3


class EqCols:
    A_0, A_1, B_0, B_1, INVERT, RES = 0, 1, 2, 3, 4, 5
    DIFF_0, DIFF_1, DIFF_2, DIFF_3, EQ, MU = 6, 7, 8, 9, 10, 11
    NUM_COLUMNS = 12


@dataclass
class EqOperation:
    a: int
    b: int
    invert: bool


def generate_eq_trace(ops) -> np.ndarray:
    c = EqCols
    n = max(4, 1 << max(len(ops) - 1, 0).bit_length())
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)
    for i, op in enumerate(ops):
        a, b = op.a & (2**64 - 1), op.b & (2**64 - 1)
        t[i, c.A_0], t[i, c.A_1] = a & 0xFFFFFFFF, a >> 32
        t[i, c.B_0], t[i, c.B_1] = b & 0xFFFFFFFF, b >> 32
        eq = a == b
        t[i, c.INVERT] = int(op.invert)
        t[i, c.RES] = int(eq ^ bool(op.invert))
        d = (a - b) & (2**64 - 1)
        for k in range(4):
            t[i, c.DIFF_0 + k] = (d >> (16 * k)) & 0xFFFF
        t[i, c.EQ] = int(eq)
        t[i, c.MU] = 1
    return t


def eq_bus_interactions():
    c = EqCols
    mu = Multiplicity("Column", (c.MU,))
    its = [BusInteraction.sender(BUS_IS_HALFWORD, mu, [BusValue.packed(Packing.DIRECT, col)])
           for col in (c.DIFF_0, c.DIFF_1, c.DIFF_2, c.DIFF_3)]
    its.append(BusInteraction.sender(BUS_ZERO, mu, [
        BusValue.of_linear([LinearTerm(1, col) for col in (c.DIFF_0, c.DIFF_1, c.DIFF_2, c.DIFF_3)]),
        BusValue.packed(Packing.DIRECT, c.EQ)]))
    its.append(BusInteraction.receiver(BUS_ALU, mu, [
        BusValue.packed(Packing.DWORD_WL, c.A_0),
        BusValue.packed(Packing.DWORD_WL, c.B_0),
        BusValue.of_linear([LinearTerm(ALU_OP_EQ), LinearTerm(64, c.INVERT)]),
        BusValue.packed(Packing.DIRECT, c.RES),
        BusValue.of_linear([LinearTerm(0)])]))
    return its


def _emit_is_bit(b: ir.IrBuilder, idx: int, col: int) -> None:
    x = b.main(0, col)
    b.emit(idx, x * (b.const_base(1) - x))


def eq_table_air() -> AirProgram:
    c = EqCols
    b = ir.IrBuilder()
    inv = b.const_base(INV_SHIFT_32)
    two16 = b.const_base(1 << 16)
    lhs_lo, lhs_hi = b.main(0, c.B_0), b.main(0, c.B_1)
    rhs_lo = b.main(0, c.DIFF_0) + b.main(0, c.DIFF_1) * two16
    rhs_hi = b.main(0, c.DIFF_2) + b.main(0, c.DIFF_3) * two16
    sum_lo, sum_hi = b.main(0, c.A_0), b.main(0, c.A_1)
    carry0 = (lhs_lo + rhs_lo - sum_lo) * inv
    carry1 = (lhs_hi + rhs_hi + carry0 - sum_hi) * inv
    one = b.const_base(1)
    b.emit(0, carry0 * (one - carry0))
    b.emit(1, carry1 * (one - carry1))
    _emit_is_bit(b, 2, c.INVERT)
    res, eq, invert = b.main(0, c.RES), b.main(0, c.EQ), b.main(0, c.INVERT)
    b.emit(3, res - (eq + invert - b.const_base(2) * eq * invert))
    _emit_is_bit(b, 4, c.MU)
    its = eq_bus_interactions()
    n_total = emit_logup_constraints(b, its, 5)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    air = AirProgram("eq", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                     boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                     max_degree=max(3 if pairs else 0, 1 + absorbed, 2))
    return air
