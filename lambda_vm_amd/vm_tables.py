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

BYTEWISE (prover/src/tables/bytewise.rs): full-word AND/OR/XOR of two 64-bit
values, byte by byte. Columns (bytewise.rs:29-42): a (8 bytes, DWordBL), b (8
bytes), op (the alu_op code), res (8 bytes), mu. No polynomial constraints of
its own; its bus interactions (bytewise.rs:127-203): BYTE_ALU[op, a_i, b_i,
res_i] sent with multiplicity mu for each byte, ARE_BYTES[mu, 0] sent once per
row, ALU[a, b, op, res] (DWordBL operands) received with mu. Trace generation
(bytewise.rs:96-121): duplicate operations merged with their count as mu,
split at MU_MAX = 255 (types.rs:528-550), padded to a power of two (at least
4). The reference merges through a hash map, so its row order is arbitrary;
here rows follow first occurrence.

IS_WH (prover/src/tables/is_wh.rs): range check of the top 48 bits of a 64-bit
value, one row per distinct (word, half) pair. Columns (is_wh.rs:41-54):
value[0] (the Half: bits 16..32), value[1], value[2] (the Word as two halves),
mu (a B20). No polynomial constraints; bus interactions (is_wh.rs:158-190):
IS_B20[mu] sent once per row, IS_HALFWORD[value[i]] sent with mu for i = 0..2,
IS_WH[value[1] + 2^16 value[2], value[0]] received with mu. Trace generation
(is_wh.rs:83-127): pairs counted, sorted by (word, half), split at 2^20 - 1,
padded with zero rows to a power of two (at least 4).

LT (prover/src/tables/lt.rs): signed or unsigned lhs < rhs for 64-bit values,
optionally inverted. Columns (lt.rs:41-87): lhs, rhs (DWordHHW: word, half,
half), signed, lt, lhs - rhs (DWordHL, 4 halves), the two sign bits, invert,
out, mu. Constraints (lt.rs:400-466): the carry pair of rhs + diff = lhs
(each carry a bit), lt = signed * signed_lt + (1 - signed) * carry_1 with
signed_lt = a(1 - b) + a carry_1 + (1 - b) carry_1 for the sign bits a, b,
out = lt XOR invert, invert and signed bits. Bus interactions (lt.rs:190-330):
MSB16[lhs[2], lhs_msb] and MSB16[rhs[2], rhs_msb], IS_HALFWORD of the four
difference halves and of lhs[1], rhs[1] (all with mu), ARE_BYTES[mu, 0] once,
ALU[lhs, rhs, LT + 32 signed + 64 invert, out, 0] received with mu. Trace
generation (lt.rs:133-175): deduplicated with counts as mu (split at 255).

STORE (prover/src/tables/store.rs): receives the CPU's store MEMORY op and
sends the MEMW write. Columns (store.rs:35-49): base_address (DWordWL),
timestamp (DWordWL), write2, write4, write8 (width bits, none set for one
byte), value (8 bytes, DWordBL), mu. Constraints (store.rs:259-283): IS_BIT
of write2, write4, write8 and mu; (sum of widths)(1 - sum); (sum of
widths)(1 - mu). Bus interactions (store.rs:129-248): MEMW[0, base_address,
value bytes, timestamp, write2, write4, write8] sent with mu (16 elements),
MEMORY[timestamp, base_address, value (DWordBL), 1 + 4 write2 + 8 write4 +
16 write8, 0, 0] received with mu, ARE_BYTES[value_i, 0] sent with mu for each
byte. Trace generation (store.rs:94-116): one row per operation (mu = 1), zero
rows to a power of two, at least 4.

BRANCH (prover/src/tables/branch.rs): next pc of a branch or jump, base +
offset with the low bit cleared (base = register for JALR, else pc). Columns
(branch.rs:38-88): pc, offset, register (DWordWL each), JALR, next_pc_high
(3 halves), next_pc_low (2 bytes, bit 0 cleared), unmasked_low_byte, mu.
Constraints (branch.rs:430-473): the carry pair of base + offset = unmasked
next pc, each carry a bit, on the pc path weighted by 1 - JALR and on the
register path by JALR; IS_BIT(JALR). Bus interactions (branch.rs:230-370):
ARE_BYTES[mu, next_pc_low_1] once per row, BYTE_ALU[AND, unmasked_low_byte,
254, next_pc_low_0] and IS_HALFWORD of the three high halves with mu,
BRANCH[next_pc (two words), pc, offset, register, JALR] received with mu.
Trace generation (branch.rs:165-222): deduplicated with counts as mu (split
at 255), zero rows to a power of two, at least 4.

MEMW_REGISTER (prover/src/tables/memw_register.rs): register reads and writes.
Columns (memw_register.rs:55-82): address (register index), timestamp
(DWordWL), value (2 words), old value (2 words), old timestamp low word (the
high word is shared with the timestamp), mu_read, mu_write. Constraints
(memw_register.rs:497-507): IS_BIT of both multiplicities and of their sum.
Bus interactions (memw_register.rs:280-480), all but the last two with
multiplicity mu_read + mu_write: IS_HALFWORD[timestamp_0 - old_timestamp_lo -
1]; per word the Memory token [1, 2 address + i, 0, old timestamp, old_i]
sent and [1, 2 address + i, 0, timestamp, value_i] received; the MEMW read
(24 elements, with the old value) received with mu_read and the MEMW write
(16 elements) with mu_write. Trace generation (memw_register.rs:210-247): one
row per access, zero rows to a power of two, at least 4.

LOAD (prover/src/tables/load.rs): loads of 1, 2, 4 or 8 bytes with sign or
zero extension. Columns (load.rs:36-68): base_address, timestamp (DWordWL),
read2, read4, read8, signed, res (8 bytes, already extended), sign_bit, mu.
Constraints (load.rs:505-590): IS_BIT of signed and the three width flags;
width sum a bit; a width implies mu; the bytes above the loaded width equal
signed * sign_bit * 255 (bytes 4..8 unless read8, 2..4 unless read4/read8,
byte 1 on one-byte loads); IS_BIT(mu). Bus interactions (load.rs:236-470):
the MEMW read (24 elements: res as old and as value) sent with mu; MSB8[the
top loaded byte, sign_bit] sent with mu - read2 - read4 - read8 (byte 0),
read2 (byte 1) or read4 (byte 3); MEMORY[timestamp, base_address, 0, 0,
2 signed + 4 read2 + 8 read4 + 16 read8, res (DWordBL)] received with mu.
Trace generation (load.rs:182-232): one row per operation (mu = 1), zero rows
to a power of two, at least 4.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import constraint_ir as ir
from .air import AirProgram, Boundary, emit_logup_constraints
from .logup import ONE, BusInteraction, BusValue, LinearTerm, Multiplicity, Packing, split_interactions

P = ir.P
INV_SHIFT_32 = 18446744065119617026  # 2^-32 mod p (templates.rs:26)

# bus ids (prover/src/tables/types.rs:255-336) and the ALU op code (:562-570)
BUS_ARE_BYTES, BUS_IS_HALFWORD, BUS_IS_B20, BUS_MSB16, BUS_ZERO, BUS_BYTE_ALU, BUS_ALU, BUS_IS_WH = (
    0, 1, 2, 7, 8, 24, 25, 32)
BUS_MSB8, BUS_MEMW, BUS_MEMORY, BUS_BRANCH, BUS_MEMORY_OP = 6, 14, 16, 17, 26
ALU_OP_AND, ALU_OP_OR, ALU_OP_XOR, ALU_OP_EQ, ALU_OP_LT = 0, 1, 2, 3, 4
MU_MAX = 255


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


class BytewiseCols:
    A = tuple(range(0, 8))
    B = tuple(range(8, 16))
    OP = 16
    RES = tuple(range(17, 25))
    MU = 25
    NUM_COLUMNS = 26


@dataclass(frozen=True)
class BytewiseOperation:
    a: int
    b: int
    op: int

    def compute_res(self) -> int:
        if self.op == ALU_OP_AND:
            return self.a & self.b
        if self.op == ALU_OP_OR:
            return self.a | self.b
        if self.op == ALU_OP_XOR:
            return self.a ^ self.b
        raise ValueError(f"BYTEWISE only handles AND/OR/XOR, got opcode {self.op}")


def dedup_rows(ops):
    """(op, mu) rows: each distinct op with its count, split at MU_MAX
    (types.rs:534-550), in order of first occurrence."""
    counts = {}
    for op in ops:
        counts[op] = counts.get(op, 0) + 1
    rows = []
    for op, left in counts.items():
        while left > 0:
            mu = min(left, MU_MAX)
            left -= mu
            rows.append((op, mu))
    return rows


def generate_bytewise_trace(ops) -> np.ndarray:
    c = BytewiseCols
    rows = dedup_rows(ops)
    n = max(4, 1 << max(len(rows) - 1, 0).bit_length())
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)
    for i, (op, mu) in enumerate(rows):
        res = op.compute_res()
        for k in range(8):
            t[i, c.A[k]] = (op.a >> (8 * k)) & 0xFF
            t[i, c.B[k]] = (op.b >> (8 * k)) & 0xFF
            t[i, c.RES[k]] = (res >> (8 * k)) & 0xFF
        t[i, c.OP] = op.op
        t[i, c.MU] = mu
    return t


def bytewise_bus_interactions():
    c = BytewiseCols
    mu = Multiplicity("Column", (c.MU,))
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    its = [BusInteraction.sender(BUS_BYTE_ALU, mu, [d(c.OP), d(c.A[i]), d(c.B[i]), d(c.RES[i])]) for i in range(8)]
    its.append(BusInteraction.sender(BUS_ARE_BYTES, ONE, [d(c.MU), BusValue.of_linear([LinearTerm(0)])]))
    its.append(BusInteraction.receiver(BUS_ALU, mu, [
        BusValue.packed(Packing.DWORD_BL, c.A[0]), BusValue.packed(Packing.DWORD_BL, c.B[0]), d(c.OP),
        BusValue.packed(Packing.DWORD_BL, c.RES[0])]))
    return its


def bytewise_table_air() -> AirProgram:
    """BYTEWISE as an AIR: only the LogUp constraints of its interactions."""
    c = BytewiseCols
    b = ir.IrBuilder()
    its = bytewise_bus_interactions()
    n_total = emit_logup_constraints(b, its, 0)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("bytewise", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3 if pairs else 0, 1 + absorbed))


class IsWhCols:
    VALUE_0, VALUE_1, VALUE_2, MU = 0, 1, 2, 3
    NUM_COLUMNS = 4


IS_WH_MU_MAX = (1 << 20) - 1


def is_wh_key(value: int) -> tuple:
    """(word, half) of a 64-bit value: bits 32..64 and 16..32 (is_wh.rs:68-76)."""
    return ((value >> 32) & 0xFFFFFFFF, (value >> 16) & 0xFFFF)


def dedup_is_wh_rows(keys):
    """((word, half), mu) rows sorted by key, split at 2^20 - 1 (is_wh.rs:83-99)."""
    counts = {}
    for k in keys:
        counts[k] = counts.get(k, 0) + 1
    rows = []
    for k in sorted(counts):
        left = counts[k]
        while left > 0:
            mu = min(left, IS_WH_MU_MAX)
            left -= mu
            rows.append((k, mu))
    return rows


def generate_is_wh_trace(rows) -> np.ndarray:
    c = IsWhCols
    n = max(4, 1 << max(len(rows) - 1, 0).bit_length())
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)
    for i, ((word, half), mu) in enumerate(rows):
        t[i, c.VALUE_0] = half
        t[i, c.VALUE_1] = word & 0xFFFF
        t[i, c.VALUE_2] = word >> 16
        t[i, c.MU] = mu
    return t


def is_wh_bus_interactions():
    c = IsWhCols
    mu = Multiplicity("Column", (c.MU,))
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    return [
        BusInteraction.sender(BUS_IS_B20, ONE, [d(c.MU)]),
        BusInteraction.sender(BUS_IS_HALFWORD, mu, [d(c.VALUE_0)]),
        BusInteraction.sender(BUS_IS_HALFWORD, mu, [d(c.VALUE_1)]),
        BusInteraction.sender(BUS_IS_HALFWORD, mu, [d(c.VALUE_2)]),
        BusInteraction.receiver(BUS_IS_WH, mu, [
            BusValue.of_linear([LinearTerm(1, c.VALUE_1), LinearTerm(1 << 16, c.VALUE_2)]), d(c.VALUE_0)]),
    ]


def is_wh_table_air() -> AirProgram:
    """IS_WH as an AIR: only the LogUp constraints of its interactions."""
    c = IsWhCols
    b = ir.IrBuilder()
    its = is_wh_bus_interactions()
    n_total = emit_logup_constraints(b, its, 0)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("is_wh", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3 if pairs else 0, 1 + absorbed))


class LtCols:
    LHS_0, LHS_1, LHS_2, RHS_0, RHS_1, RHS_2, SIGNED, LT = 0, 1, 2, 3, 4, 5, 6, 7
    SUB_0, SUB_1, SUB_2, SUB_3, LHS_MSB, RHS_MSB, INVERT, OUT, MU = 8, 9, 10, 11, 12, 13, 14, 15, 16
    NUM_COLUMNS = 17


@dataclass(frozen=True)
class LtOperation:
    lhs: int
    rhs: int
    signed: bool
    invert: bool = False

    def compute_lt(self) -> bool:
        if self.signed:
            s = lambda v: v - (1 << 64) if v >> 63 else v  # noqa: E731
            return s(self.lhs) < s(self.rhs)
        return self.lhs < self.rhs

    def compute_out(self) -> bool:
        return self.compute_lt() ^ self.invert


def generate_lt_trace(ops) -> np.ndarray:
    c = LtCols
    rows = dedup_rows(ops)
    n = max(4, 1 << max(len(rows) - 1, 0).bit_length())
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)
    for i, (op, mu) in enumerate(rows):
        for base, v in ((c.LHS_0, op.lhs), (c.RHS_0, op.rhs)):  # DWordHHW
            t[i, base] = v & 0xFFFFFFFF
            t[i, base + 1] = (v >> 32) & 0xFFFF
            t[i, base + 2] = (v >> 48) & 0xFFFF
        t[i, c.SIGNED] = int(op.signed)
        t[i, c.LT] = int(op.compute_lt())
        d = (op.lhs - op.rhs) & (2**64 - 1)
        for k in range(4):
            t[i, c.SUB_0 + k] = (d >> (16 * k)) & 0xFFFF
        t[i, c.LHS_MSB] = op.lhs >> 63
        t[i, c.RHS_MSB] = op.rhs >> 63
        t[i, c.INVERT] = int(op.invert)
        t[i, c.OUT] = int(op.compute_out())
        t[i, c.MU] = mu
    return t


def lt_bus_interactions():
    c = LtCols
    mu = Multiplicity("Column", (c.MU,))
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    its = [BusInteraction.sender(BUS_MSB16, mu, [d(c.LHS_2), d(c.LHS_MSB)]),
           BusInteraction.sender(BUS_MSB16, mu, [d(c.RHS_2), d(c.RHS_MSB)])]
    its += [BusInteraction.sender(BUS_IS_HALFWORD, mu, [d(col)])
            for col in (c.SUB_0, c.SUB_1, c.SUB_2, c.SUB_3, c.LHS_1, c.RHS_1)]
    its.append(BusInteraction.sender(BUS_ARE_BYTES, ONE, [d(c.MU), BusValue.of_linear([LinearTerm(0)])]))
    its.append(BusInteraction.receiver(BUS_ALU, mu, [
        BusValue.packed(Packing.DWORD_HHW, c.LHS_0), BusValue.packed(Packing.DWORD_HHW, c.RHS_0),
        BusValue.of_linear([LinearTerm(ALU_OP_LT), LinearTerm(32, c.SIGNED), LinearTerm(64, c.INVERT)]),
        d(c.OUT), BusValue.of_linear([LinearTerm(0)])]))
    return its


def lt_table_air() -> AirProgram:
    c = LtCols
    b = ir.IrBuilder()
    one, inv, two16 = b.const_base(1), b.const_base(INV_SHIFT_32), b.const_base(1 << 16)
    m = lambda col: b.main(0, col)  # noqa: E731
    carry0 = (m(c.RHS_0) + (m(c.SUB_0) + m(c.SUB_1) * two16) - m(c.LHS_0)) * inv
    lhs_hi = m(c.LHS_1) + m(c.LHS_2) * two16
    rhs_hi = m(c.RHS_1) + m(c.RHS_2) * two16
    sub_hi = m(c.SUB_2) + m(c.SUB_3) * two16
    carry1 = (rhs_hi + sub_hi + carry0 - lhs_hi) * inv
    b.emit(0, carry0 * (one - carry0))
    b.emit(1, carry1 * (one - carry1))
    a_msb, b_msb, signed = m(c.LHS_MSB), m(c.RHS_MSB), m(c.SIGNED)
    one_minus_b = one - b_msb
    signed_lt = a_msb * one_minus_b + a_msb * carry1 + one_minus_b * carry1
    expected = signed * signed_lt + (one - signed) * carry1
    b.emit(2, m(c.LT) - expected)
    lt, invert = m(c.LT), m(c.INVERT)
    b.emit(3, m(c.OUT) - (lt + invert - b.const_base(2) * lt * invert))
    b.emit(4, invert * (one - invert))
    b.emit(5, signed * (one - signed))
    its = lt_bus_interactions()
    n_total = emit_logup_constraints(b, its, 6)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("lt", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3, 1 + absorbed))


class StoreCols:
    BASE_ADDRESS_0, BASE_ADDRESS_1, TIMESTAMP_0, TIMESTAMP_1, WRITE2, WRITE4, WRITE8 = 0, 1, 2, 3, 4, 5, 6
    VALUE = tuple(range(7, 15))
    MU = 15
    NUM_COLUMNS = 16


@dataclass(frozen=True)
class StoreOperation:
    base_address: int
    timestamp: int
    value: int
    write2: bool = False
    write4: bool = False
    write8: bool = False

    @classmethod
    def of_width(cls, base_address: int, timestamp: int, value: int, nbytes: int) -> "StoreOperation":
        """store.rs:69-78: a width of 1 sets no flag."""
        return cls(base_address, timestamp, value, nbytes == 2, nbytes == 4, nbytes == 8)


def generate_store_trace(ops) -> np.ndarray:
    c = StoreCols
    n = max(4, 1 << max(len(ops) - 1, 0).bit_length())
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)
    for i, op in enumerate(ops):
        a, ts, v = op.base_address & (2**64 - 1), op.timestamp & (2**64 - 1), op.value & (2**64 - 1)
        t[i, c.BASE_ADDRESS_0], t[i, c.BASE_ADDRESS_1] = a & 0xFFFFFFFF, a >> 32
        t[i, c.TIMESTAMP_0], t[i, c.TIMESTAMP_1] = ts & 0xFFFFFFFF, ts >> 32
        t[i, c.WRITE2], t[i, c.WRITE4], t[i, c.WRITE8] = int(op.write2), int(op.write4), int(op.write8)
        for k in range(8):
            t[i, c.VALUE[k]] = (v >> (8 * k)) & 0xFF
        t[i, c.MU] = 1
    return t


def store_bus_interactions():
    c = StoreCols
    mu = Multiplicity("Column", (c.MU,))
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    zero = BusValue.of_linear([LinearTerm(0)])
    its = [BusInteraction.sender(BUS_MEMW, mu, [zero, BusValue.packed(Packing.DWORD_WL, c.BASE_ADDRESS_0)]
                                 + [d(col) for col in c.VALUE]
                                 + [BusValue.packed(Packing.DWORD_WL, c.TIMESTAMP_0), d(c.WRITE2), d(c.WRITE4),
                                    d(c.WRITE8)]),
           BusInteraction.receiver(BUS_MEMORY_OP, mu, [
               BusValue.packed(Packing.DWORD_WL, c.TIMESTAMP_0), BusValue.packed(Packing.DWORD_WL, c.BASE_ADDRESS_0),
               BusValue.packed(Packing.DWORD_BL, c.VALUE[0]),
               BusValue.of_linear([LinearTerm(1), LinearTerm(4, c.WRITE2), LinearTerm(8, c.WRITE4),
                                   LinearTerm(16, c.WRITE8)]),
               zero, zero])]
    its += [BusInteraction.sender(BUS_ARE_BYTES, mu, [d(col), zero]) for col in c.VALUE]
    return its


def store_table_air() -> AirProgram:
    c = StoreCols
    b = ir.IrBuilder()
    for idx, col in enumerate((c.WRITE2, c.WRITE4, c.WRITE8, c.MU)):
        _emit_is_bit(b, idx, col)
    one = b.const_base(1)
    width = b.main(0, c.WRITE2) + b.main(0, c.WRITE4) + b.main(0, c.WRITE8)
    b.emit(4, width * (one - width))
    b.emit(5, width * (one - b.main(0, c.MU)))
    its = store_bus_interactions()
    n_total = emit_logup_constraints(b, its, 6)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("store", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3 if pairs else 0, 1 + absorbed, 2))


class BranchCols:
    PC_0, PC_1, OFFSET_0, OFFSET_1, REGISTER_0, REGISTER_1, JALR = 0, 1, 2, 3, 4, 5, 6
    NEXT_PC_HIGH_0, NEXT_PC_HIGH_1, NEXT_PC_HIGH_2, NEXT_PC_LOW_0, NEXT_PC_LOW_1 = 7, 8, 9, 10, 11
    UNMASKED_LOW_BYTE, MU = 12, 13
    NUM_COLUMNS = 14


BRANCH_MASK = 254  # clears the low bit of the target (branch.rs:95-99)


@dataclass(frozen=True)
class BranchOperation:
    pc: int
    offset: int
    register: int
    jalr: bool

    def next_pc_unmasked(self) -> int:
        return ((self.register if self.jalr else self.pc) + self.offset) & (2**64 - 1)

    def next_pc(self) -> int:
        return self.next_pc_unmasked() & ~1 & (2**64 - 1)


def generate_branch_trace(ops) -> np.ndarray:
    c = BranchCols
    rows = dedup_rows(ops)
    n = max(4, 1 << max(len(rows) - 1, 0).bit_length())
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)
    for i, (op, mu) in enumerate(rows):
        for base, v in ((c.PC_0, op.pc), (c.OFFSET_0, op.offset), (c.REGISTER_0, op.register)):
            v &= 2**64 - 1
            t[i, base], t[i, base + 1] = v & 0xFFFFFFFF, v >> 32
        t[i, c.JALR] = int(op.jalr)
        nxt = op.next_pc()
        t[i, c.NEXT_PC_HIGH_0] = (nxt >> 16) & 0xFFFF
        t[i, c.NEXT_PC_HIGH_1] = (nxt >> 32) & 0xFFFF
        t[i, c.NEXT_PC_HIGH_2] = (nxt >> 48) & 0xFFFF
        t[i, c.NEXT_PC_LOW_0] = nxt & 0xFF
        t[i, c.NEXT_PC_LOW_1] = (nxt >> 8) & 0xFF
        t[i, c.UNMASKED_LOW_BYTE] = op.next_pc_unmasked() & 0xFF
        t[i, c.MU] = mu
    return t


def branch_bus_interactions():
    c = BranchCols
    mu = Multiplicity("Column", (c.MU,))
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    its = [BusInteraction.sender(BUS_ARE_BYTES, ONE, [d(c.MU), d(c.NEXT_PC_LOW_1)]),
           BusInteraction.sender(BUS_BYTE_ALU, mu, [BusValue.of_linear([LinearTerm(ALU_OP_AND)]),
                                                    d(c.UNMASKED_LOW_BYTE),
                                                    BusValue.of_linear([LinearTerm(BRANCH_MASK)]),
                                                    d(c.NEXT_PC_LOW_0)])]
    its += [BusInteraction.sender(BUS_IS_HALFWORD, mu, [d(col)])
            for col in (c.NEXT_PC_HIGH_0, c.NEXT_PC_HIGH_1, c.NEXT_PC_HIGH_2)]
    its.append(BusInteraction.receiver(BUS_BRANCH, mu, [
        BusValue.of_linear([LinearTerm(1, c.NEXT_PC_LOW_0), LinearTerm(1 << 8, c.NEXT_PC_LOW_1),
                            LinearTerm(1 << 16, c.NEXT_PC_HIGH_0)]),
        BusValue.of_linear([LinearTerm(1, c.NEXT_PC_HIGH_1), LinearTerm(1 << 16, c.NEXT_PC_HIGH_2)]),
        d(c.PC_0), d(c.PC_1), d(c.OFFSET_0), d(c.OFFSET_1), d(c.REGISTER_0), d(c.REGISTER_1), d(c.JALR)]))
    return its


def branch_table_air() -> AirProgram:
    c = BranchCols
    b = ir.IrBuilder()
    one, inv = b.const_base(1), b.const_base(INV_SHIFT_32)
    m = lambda col: b.main(0, col)  # noqa: E731
    unmasked0 = m(c.UNMASKED_LOW_BYTE) + m(c.NEXT_PC_LOW_1) * b.const_base(1 << 8) \
        + m(c.NEXT_PC_HIGH_0) * b.const_base(1 << 16)
    unmasked1 = m(c.NEXT_PC_HIGH_1) + m(c.NEXT_PC_HIGH_2) * b.const_base(1 << 16)
    jalr = m(c.JALR)
    for idx, (cond, base0) in enumerate(((one - jalr, c.PC_0), (jalr, c.REGISTER_0))):
        carry0 = (m(base0) + m(c.OFFSET_0) - unmasked0) * inv
        carry1 = (m(base0 + 1) + m(c.OFFSET_1) + carry0 - unmasked1) * inv
        b.emit(2 * idx, cond * carry0 * (one - carry0))
        b.emit(2 * idx + 1, cond * carry1 * (one - carry1))
    b.emit(4, jalr * (one - jalr))
    its = branch_bus_interactions()
    n_total = emit_logup_constraints(b, its, 5)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("branch", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3, 1 + absorbed))


class MemwRegisterCols:
    ADDRESS, TIMESTAMP_0, TIMESTAMP_1, VAL_0, VAL_1, OLD_0, OLD_1, OLD_TIMESTAMP_LO, MU_READ, MU_WRITE = range(10)
    NUM_COLUMNS = 10


@dataclass(frozen=True)
class RegisterAccess:
    """One register access (memw_register.rs:84-167): base_address = 2 x the
    register index; the old timestamp shares the high word of the timestamp."""
    base_address: int
    timestamp: int
    value: int
    old: int
    old_timestamp: int
    is_read: bool


def generate_memw_register_trace(accesses) -> np.ndarray:
    c = MemwRegisterCols
    n = max(4, 1 << max(len(accesses) - 1, 0).bit_length())
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)
    M64 = (1 << 64) - 1
    for i, a in enumerate(accesses):
        if a.base_address % 2:
            raise ValueError("register base address must be even")
        if not 0 <= a.base_address // 2 < 1 << 16:
            raise ValueError("register index must fit in 16 bits")
        # RegRow::from_memw (prover/src/tables/memw_register.rs:130-195): the old
        # access shares the timestamp's high word and lies 1..2^16 below it in
        # the low word (the IS_HALFWORD check of ts_lo - old_ts_lo - 1)
        if (a.old_timestamp >> 32) != (a.timestamp >> 32):
            raise ValueError("old timestamp must share the timestamp's high word")
        if not 1 <= (a.timestamp & 0xFFFFFFFF) - (a.old_timestamp & 0xFFFFFFFF) <= 1 << 16:
            raise ValueError("timestamp low-word delta must be in [1, 2^16]")
        value, old = a.value & M64, a.old & M64
        t[i, c.ADDRESS] = a.base_address // 2
        t[i, c.TIMESTAMP_0], t[i, c.TIMESTAMP_1] = a.timestamp & 0xFFFFFFFF, (a.timestamp >> 32) & 0xFFFFFFFF
        t[i, c.VAL_0], t[i, c.VAL_1] = value & 0xFFFFFFFF, value >> 32
        t[i, c.OLD_0], t[i, c.OLD_1] = old & 0xFFFFFFFF, old >> 32
        t[i, c.OLD_TIMESTAMP_LO] = a.old_timestamp & 0xFFFFFFFF
        t[i, c.MU_READ], t[i, c.MU_WRITE] = int(a.is_read), int(not a.is_read)
    return t


def memw_register_bus_interactions():
    c = MemwRegisterCols
    mu_sum = Multiplicity("Sum", (c.MU_READ, c.MU_WRITE))
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    k = lambda v: BusValue.of_linear([LinearTerm(v)])  # noqa: E731
    its = [BusInteraction.sender(BUS_IS_HALFWORD, mu_sum, [BusValue.of_linear(
        [LinearTerm(1, c.TIMESTAMP_0), LinearTerm(-1, c.OLD_TIMESTAMP_LO), LinearTerm(-1)])])]
    for i in range(2):
        its.append(BusInteraction.sender(BUS_MEMORY, mu_sum, [
            k(1), BusValue.of_linear([LinearTerm(2, c.ADDRESS), LinearTerm(i)]), k(0), d(c.OLD_TIMESTAMP_LO),
            d(c.TIMESTAMP_1), d(c.OLD_0 + i)]))
    for i in range(2):
        its.append(BusInteraction.receiver(BUS_MEMORY, mu_sum, [
            k(1), BusValue.of_linear([LinearTerm(2, c.ADDRESS), LinearTerm(i)]), k(0), d(c.TIMESTAMP_0),
            d(c.TIMESTAMP_1), d(c.VAL_0 + i)]))
    addr = BusValue.of_linear([LinearTerm(2, c.ADDRESS)])
    write = [k(1), addr, k(0), d(c.VAL_0), d(c.VAL_1)] + [k(0)] * 6 + [d(c.TIMESTAMP_0), d(c.TIMESTAMP_1), k(1), k(0),
                                                                         k(0)]
    its.append(BusInteraction.receiver(BUS_MEMW, Multiplicity("Column", (c.MU_READ,)),
                                       [d(c.OLD_0), d(c.OLD_1)] + [k(0)] * 6 + write))
    its.append(BusInteraction.receiver(BUS_MEMW, Multiplicity("Column", (c.MU_WRITE,)), list(write)))
    return its


def memw_register_table_air() -> AirProgram:
    c = MemwRegisterCols
    b = ir.IrBuilder()
    _emit_is_bit(b, 0, c.MU_READ)
    _emit_is_bit(b, 1, c.MU_WRITE)
    mu_sum = b.main(0, c.MU_READ) + b.main(0, c.MU_WRITE)
    b.emit(2, mu_sum * (b.const_base(1) - mu_sum))
    its = memw_register_bus_interactions()
    n_total = emit_logup_constraints(b, its, 3)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("memw_register", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False),
                      [0] * n_total, boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3 if pairs else 0, 1 + absorbed, 2))


class LoadCols:
    BASE_ADDRESS_0, BASE_ADDRESS_1, TIMESTAMP_0, TIMESTAMP_1, READ2, READ4, READ8, SIGNED = range(8)
    RES = tuple(range(8, 16))
    SIGN_BIT, MU = 16, 17
    NUM_COLUMNS = 18


@dataclass(frozen=True)
class LoadOperation:
    base_address: int
    timestamp: int
    width: int
    signed: bool
    res: tuple  # 8 bytes, already extended

    @classmethod
    def of_memory(cls, base_address: int, timestamp: int, width: int, signed: bool, raw: int) -> "LoadOperation":
        """The extended result of loading `width` bytes of `raw` (RISC-V LB..LD,
        LBU..LWU)."""
        v = raw & ((1 << (8 * width)) - 1)
        if signed and v >> (8 * width - 1):
            v |= (2**64 - 1) ^ ((1 << (8 * width)) - 1)
        return cls(base_address, timestamp, width, signed, tuple((v >> (8 * k)) & 0xFF for k in range(8)))

    def sign_bit(self) -> int:
        """load.rs:113-122: the top bit of the top loaded byte."""
        return (self.res[{1: 0, 2: 1, 4: 3, 8: 7}.get(self.width, 0)] >> 7) & 1


def generate_load_trace(ops) -> np.ndarray:
    c = LoadCols
    n = max(4, 1 << max(len(ops) - 1, 0).bit_length())
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)
    for i, op in enumerate(ops):
        a, ts = op.base_address & (2**64 - 1), op.timestamp & (2**64 - 1)
        t[i, c.BASE_ADDRESS_0], t[i, c.BASE_ADDRESS_1] = a & 0xFFFFFFFF, a >> 32
        t[i, c.TIMESTAMP_0], t[i, c.TIMESTAMP_1] = ts & 0xFFFFFFFF, ts >> 32
        t[i, c.READ2], t[i, c.READ4], t[i, c.READ8] = int(op.width == 2), int(op.width == 4), int(op.width == 8)
        t[i, c.SIGNED] = int(op.signed)
        for k in range(8):
            t[i, c.RES[k]] = op.res[k]
        t[i, c.SIGN_BIT] = op.sign_bit()
        t[i, c.MU] = 1
    return t


def load_bus_interactions():
    c = LoadCols
    mu = Multiplicity("Column", (c.MU,))
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    k = lambda v: BusValue.of_linear([LinearTerm(v)])  # noqa: E731
    res = [d(col) for col in c.RES]
    its = [BusInteraction.sender(BUS_MEMW, mu, res + [k(0), d(c.BASE_ADDRESS_0), d(c.BASE_ADDRESS_1)] + list(res)
                                 + [d(c.TIMESTAMP_0), d(c.TIMESTAMP_1), d(c.READ2), d(c.READ4), d(c.READ8)]),
           BusInteraction.sender(BUS_MSB8, Multiplicity("Linear", linear=(
               LinearTerm(1, c.MU), LinearTerm(-1, c.READ2), LinearTerm(-1, c.READ4), LinearTerm(-1, c.READ8))), [d(c.RES[0]), d(c.SIGN_BIT)]),
           BusInteraction.sender(BUS_MSB8, Multiplicity("Column", (c.READ2,)), [d(c.RES[1]), d(c.SIGN_BIT)]),
           BusInteraction.sender(BUS_MSB8, Multiplicity("Column", (c.READ4,)), [d(c.RES[3]), d(c.SIGN_BIT)]),
           BusInteraction.receiver(BUS_MEMORY_OP, mu, [
               BusValue.packed(Packing.DWORD_WL, c.TIMESTAMP_0), BusValue.packed(Packing.DWORD_WL, c.BASE_ADDRESS_0),
               k(0), k(0),
               BusValue.of_linear([LinearTerm(2, c.SIGNED), LinearTerm(4, c.READ2), LinearTerm(8, c.READ4),
                                   LinearTerm(16, c.READ8)]),
               BusValue.packed(Packing.DWORD_BL, c.RES[0])])]
    return its


def load_table_air() -> AirProgram:
    c = LoadCols
    b = ir.IrBuilder()
    one = b.const_base(1)
    m = lambda col: b.main(0, col)  # noqa: E731
    for idx, col in enumerate((c.SIGNED, c.READ2, c.READ4, c.READ8)):
        _emit_is_bit(b, idx, col)
    r2, r4, r8 = m(c.READ2), m(c.READ4), m(c.READ8)
    width = r2 + r4 + r8
    b.emit(4, width * (one - width))
    b.emit(5, width * (one - m(c.MU)))
    ext = m(c.SIGNED) * m(c.SIGN_BIT) * b.const_base(255)
    for off, i in enumerate(range(4, 8)):
        b.emit(6 + off, (one - r8) * (m(c.RES[i]) - ext))
    for off, i in enumerate(range(2, 4)):
        b.emit(10 + off, (one - r4 - r8) * (m(c.RES[i]) - ext))
    b.emit(12, (one - r2 - r4 - r8) * (m(c.RES[1]) - ext))
    _emit_is_bit(b, 13, c.MU)
    its = load_bus_interactions()
    n_total = emit_logup_constraints(b, its, 14)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("load", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3, 1 + absorbed))


# ---------------------------------------------------------------------------
# Preprocessed tables: the leading columns are fixed per table (a function of
# the row index only) and commit to a tree of their own; the multiplicity
# columns after them commit to the main tree (prover.rs:1194-1290).

BUS_HWSL, BUS_KECCAK_RC = 9, 23

# Keccak-f[1600] round constants (executor/src/vm/instruction/execution.rs:1001, FIPS 202)
KECCAK_ROUND_CONSTANTS = (
    0x0000000000000001, 0x0000000000008082, 0x800000000000808A, 0x8000000080008000,
    0x000000000000808B, 0x0000000080000001, 0x8000000080008081, 0x8000000000008009,
    0x000000000000008A, 0x0000000000000088, 0x0000000080008009, 0x000000008000000A,
    0x000000008000808B, 0x800000000000008B, 0x8000000000008089, 0x8000000000008003,
    0x8000000000008002, 0x8000000000000080, 0x000000000000800A, 0x800000008000000A,
    0x8000000080008081, 0x8000000000008080, 0x0000000080000001, 0x8000000080008008)


class KeccakRcCols:
    """prover/src/tables/keccak_rc.rs:27-35: the round index, the 8 bytes of its
    round constant (little-endian), the multiplicity."""
    ROUND = 0
    RC = tuple(range(1, 9))
    MU = 9
    NUM_COLUMNS = 10
    NUM_PRECOMPUTED = 9
    NUM_ROWS = 32  # 24 rounds padded to a power of two (keccak_rc.rs:41-45)


def keccak_rc_precomputed() -> np.ndarray:
    """The (32, 9) precomputed columns: [round, rc byte 0..7] per row, zero
    constants past round 23 (keccak_rc.rs:52-67)."""
    c = KeccakRcCols
    t = np.zeros((c.NUM_ROWS, c.NUM_PRECOMPUTED), dtype=np.uint64)
    for r in range(c.NUM_ROWS):
        rc = KECCAK_ROUND_CONSTANTS[r] if r < 24 else 0
        t[r, 0] = r
        for i in range(8):
            t[r, 1 + i] = (rc >> (8 * i)) & 0xFF
    return t


def generate_keccak_rc_trace(num_keccak_ops: int) -> np.ndarray:
    """The precomputed columns and MU = number of permutations on each real
    round row (keccak_rc.rs:191-227)."""
    c = KeccakRcCols
    t = np.zeros((c.NUM_ROWS, c.NUM_COLUMNS), dtype=np.uint64)
    t[:, :c.NUM_PRECOMPUTED] = keccak_rc_precomputed()
    t[:24, c.MU] = num_keccak_ops
    return t


def keccak_rc_bus_interactions():
    """One receiver on the KeccakRc bus: [round, rc[0..8]] with MU (keccak_rc.rs:236-250)."""
    c = KeccakRcCols
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    return [BusInteraction.receiver(BUS_KECCAK_RC, Multiplicity("Column", (c.MU,)),
                                    [d(c.ROUND)] + [d(col) for col in c.RC])]


def keccak_rc_table_air() -> AirProgram:
    """KECCAK_RC as an AIR: only the LogUp constraints of its receiver; the
    first nine columns are preprocessed."""
    c = KeccakRcCols
    b = ir.IrBuilder()
    its = keccak_rc_bus_interactions()
    n_total = emit_logup_constraints(b, its, 0)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("keccak_rc", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3 if pairs else 0, 1 + absorbed),
                      num_precomputed=c.NUM_PRECOMPUTED, precomputed=keccak_rc_precomputed)


class BitwiseCols:
    """prover/src/tables/bitwise.rs:46-95: X, Y (bytes), Z (4 bits), AND, OR,
    XOR, MSB8, MSB16, ZERO, SLL, SLLC (precomputed), then one multiplicity
    column per lookup type."""
    X, Y, Z, AND, OR, XOR, MSB8, MSB16, ZERO, SLL, SLLC = range(11)
    MU_MSB8, MU_MSB16, MU_ZERO, MU_ARE_BYTES, MU_IS_HALF, MU_IS_B20, MU_HWSL = range(11, 18)
    MU_BYTE_ALU_AND, MU_BYTE_ALU_OR, MU_BYTE_ALU_XOR = 18, 19, 20
    NUM_COLUMNS = 21
    NUM_PRECOMPUTED = 11
    NUM_ROWS = 1 << 20


# lookup type -> multiplicity column (bitwise.rs:441-471)
BITWISE_MU_COLUMN = {"msb8": BitwiseCols.MU_MSB8, "msb16": BitwiseCols.MU_MSB16, "zero": BitwiseCols.MU_ZERO,
                     "are_bytes": BitwiseCols.MU_ARE_BYTES, "is_half": BitwiseCols.MU_IS_HALF,
                     "is_b20": BitwiseCols.MU_IS_B20, "hwsl": BitwiseCols.MU_HWSL,
                     "and": BitwiseCols.MU_BYTE_ALU_AND, "or": BitwiseCols.MU_BYTE_ALU_OR,
                     "xor": BitwiseCols.MU_BYTE_ALU_XOR}


def bitwise_row_index(x: int, y: int, z: int) -> int:
    """Row of (X, Y, Z) = x + 256 y + 65536 z (bitwise.rs:396-400)."""
    return x + 256 * y + 65536 * z


def bitwise_precomputed() -> np.ndarray:
    """The (2^20, 11) precomputed columns of every (X, Y, Z) (bitwise.rs:117-160):
    row x + 256 y + 65536 z holds X, Y, Z, X&Y, X|Y, X^Y, MSB of X, MSB of the
    halfword X + 256 Y, [X = Y = Z = 0], SLL = (halfword << Z) mod 2^16 and
    SLLC = halfword >> (16 - Z) (0 when Z = 0)."""
    idx = np.arange(BitwiseCols.NUM_ROWS, dtype=np.uint64)
    x, y, z = idx & np.uint64(0xFF), (idx >> np.uint64(8)) & np.uint64(0xFF), (idx >> np.uint64(16)) & np.uint64(0xF)
    hw = x + y * np.uint64(256)
    t = np.empty((BitwiseCols.NUM_ROWS, BitwiseCols.NUM_PRECOMPUTED), dtype=np.uint64)
    t[:, 0], t[:, 1], t[:, 2] = x, y, z
    t[:, 3], t[:, 4], t[:, 5] = x & y, x | y, x ^ y
    t[:, 6] = (x >> np.uint64(7)) & np.uint64(1)
    t[:, 7] = (hw >> np.uint64(15)) & np.uint64(1)
    t[:, 8] = (idx == 0).astype(np.uint64)
    t[:, 9] = (hw << z) & np.uint64(0xFFFF)
    zz = np.where(z == 0, np.uint64(16), z)  # z = 0: halfword >> 16 = 0
    t[:, 10] = np.where(z == 0, np.uint64(0), hw >> (np.uint64(16) - zz))
    return t


def generate_bitwise_trace(lookups) -> np.ndarray:
    """lookups: (type, x, y, z) tuples (type a BITWISE_MU_COLUMN key); the
    precomputed columns and, per lookup, one more count in its type's
    multiplicity column at row (x, y, z) (bitwise.rs:402-427)."""
    c = BitwiseCols
    t = np.zeros((c.NUM_ROWS, c.NUM_COLUMNS), dtype=np.uint64)
    t[:, :c.NUM_PRECOMPUTED] = bitwise_precomputed()
    for kind, x, y, z in lookups:
        t[bitwise_row_index(x, y, z), BITWISE_MU_COLUMN[kind]] += 1
    return t


def bitwise_bus_interactions():
    """The ten receivers of bitwise.rs:728-936, in that order."""
    c = BitwiseCols
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    mu = lambda col: Multiplicity("Column", (col,))  # noqa: E731
    hw = BusValue.of_linear([LinearTerm(1, c.X), LinearTerm(256, c.Y)])
    b20 = BusValue.of_linear([LinearTerm(1, c.X), LinearTerm(256, c.Y), LinearTerm(65536, c.Z)])
    const = lambda v: BusValue.of_linear([LinearTerm(v)])  # noqa: E731
    return [
        BusInteraction.receiver(BUS_MSB8, mu(c.MU_MSB8), [d(c.X), d(c.MSB8)]),
        BusInteraction.receiver(BUS_MSB16, mu(c.MU_MSB16), [hw, d(c.MSB16)]),
        BusInteraction.receiver(BUS_ZERO, mu(c.MU_ZERO), [b20, d(c.ZERO)]),
        BusInteraction.receiver(BUS_ARE_BYTES, mu(c.MU_ARE_BYTES), [d(c.X), d(c.Y)]),
        BusInteraction.receiver(BUS_IS_HALFWORD, mu(c.MU_IS_HALF), [hw]),
        BusInteraction.receiver(BUS_IS_B20, mu(c.MU_IS_B20), [b20]),
        BusInteraction.receiver(BUS_HWSL, mu(c.MU_HWSL), [hw, d(c.Z), d(c.SLL), d(c.SLLC)]),
        BusInteraction.receiver(BUS_BYTE_ALU, mu(c.MU_BYTE_ALU_AND), [const(ALU_OP_AND), d(c.X), d(c.Y), d(c.AND)]),
        BusInteraction.receiver(BUS_BYTE_ALU, mu(c.MU_BYTE_ALU_OR), [const(ALU_OP_OR), d(c.X), d(c.Y), d(c.OR)]),
        BusInteraction.receiver(BUS_BYTE_ALU, mu(c.MU_BYTE_ALU_XOR), [const(ALU_OP_XOR), d(c.X), d(c.Y), d(c.XOR)]),
    ]


def bitwise_table_air() -> AirProgram:
    """BITWISE as an AIR: only the LogUp constraints of its ten receivers; the
    first eleven columns are preprocessed."""
    c = BitwiseCols
    b = ir.IrBuilder()
    its = bitwise_bus_interactions()
    n_total = emit_logup_constraints(b, its, 0)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("bitwise", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3 if pairs else 0, 1 + absorbed),
                      num_precomputed=c.NUM_PRECOMPUTED, precomputed=bitwise_precomputed)


# REGISTER (prover/src/tables/register.rs): one row per register word address
# (x0..x31 at 0..63, x254 at 508/509, x255 = pc at 510/511), padded to 128.
STACK_TOP = 0xFFFFFFFFFFFFFFF0  # executor/src/vm/registers.rs:3
CPU_PADDING_PC = 1              # prover/src/tables/cpu.rs:45


class RegisterCols:
    """register.rs:81-103: OFFSET (word address), INIT, FINI (words), IS_PC
    (fixed), then the final timestamp as two words."""
    OFFSET, INIT, FINI, IS_PC, TIMESTAMP_LO, TIMESTAMP_HI = range(6)
    NUM_COLUMNS = 6
    NUM_PRECOMPUTED = 4
    NUM_ADDRESSES = 68
    NUM_ROWS = 128


def register_word_address_list() -> list:
    """register.rs:124-139."""
    return list(range(64)) + [508, 509, 510, 511]


def register_init_from_entry_point(entry_point: int) -> list:
    """The program-start register image (register.rs:160-184): the stack
    pointer's words hold STACK_TOP, the pc's the entry point, the rest 0."""
    vals = {4: STACK_TOP & 0xFFFFFFFF, 5: STACK_TOP >> 32, 510: entry_point & 0xFFFFFFFF, 511: entry_point >> 32}
    return [vals.get(a, 0) for a in register_word_address_list()]


def halted_fini(committed_bytes: int) -> list:
    """The final registers of every run that halts (register.rs:201-206): zero
    but x254 (the committed byte count) and the pc's low word (the padding pc)."""
    fini = [0] * RegisterCols.NUM_ADDRESSES
    fini[64] = committed_bytes
    fini[66] = CPU_PADDING_PC
    return fini


def fini_from_final_state(final_state: dict, init: list) -> list:
    """FINI in address order: the final value of an accessed word, else its
    init (register.rs:323-335). final_state: word address -> (timestamp, value)."""
    return [final_state[a][1] if a in final_state else init[i] for i, a in enumerate(register_word_address_list())]


def register_precomputed(init: list, fini: list) -> np.ndarray:
    """The (128, 4) fixed columns OFFSET, INIT, FINI, IS_PC (register.rs:340-362)."""
    c = RegisterCols
    t = np.zeros((c.NUM_ROWS, c.NUM_PRECOMPUTED), dtype=np.uint64)
    for i, a in enumerate(register_word_address_list()):
        t[i] = (a, init[i], fini[i], int(a in (510, 511)))
    return t


def generate_register_trace(final_state: dict, init: list) -> np.ndarray:
    """register.rs:254-293: per address its fixed columns, FINI = its final
    value and the final timestamp (1 and FINI = INIT when never accessed);
    padding rows carry timestamp 1 so their tokens cancel."""
    c = RegisterCols
    t = np.zeros((c.NUM_ROWS, c.NUM_COLUMNS), dtype=np.uint64)
    fini = fini_from_final_state(final_state, init)
    t[:, :c.NUM_PRECOMPUTED] = register_precomputed(init, fini)
    for i, a in enumerate(register_word_address_list()):
        ts = final_state[a][0] if a in final_state else 1
        t[i, c.TIMESTAMP_LO], t[i, c.TIMESTAMP_HI] = ts & 0xFFFFFFFF, ts >> 32
    t[c.NUM_ADDRESSES:, c.TIMESTAMP_LO] = 1
    return t


def register_bus_interactions(final_pc_ts: int):
    """register.rs:399-462: the initial Memory token (timestamp 1) received
    once per row; the final token sent, at the fixed final pc timestamp on
    the pc rows (multiplicity IS_PC) and at the row's timestamp elsewhere
    (multiplicity 1 - IS_PC)."""
    c = RegisterCols
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    k = lambda v: BusValue.of_linear([LinearTerm(v)])  # noqa: E731
    return [
        BusInteraction.receiver(BUS_MEMORY, ONE, [k(1), d(c.OFFSET), k(0), k(1), k(0), d(c.INIT)]),
        BusInteraction.sender(BUS_MEMORY, Multiplicity("Column", (c.IS_PC,)),
                              [k(1), d(c.OFFSET), k(0), k(final_pc_ts), k(0), d(c.FINI)]),
        BusInteraction.sender(BUS_MEMORY, Multiplicity("Negated", (c.IS_PC,)),
                              [k(1), d(c.OFFSET), k(0), d(c.TIMESTAMP_LO), d(c.TIMESTAMP_HI), d(c.FINI)]),
    ]


def register_table_air(init: list, fini: list, final_pc_ts: int) -> AirProgram:
    """REGISTER as an AIR: only the LogUp constraints of its three Memory
    interactions; OFFSET, INIT, FINI and IS_PC are fixed (a monolithic proof
    fixes FINI = halted_fini, an epoch's FINI is the next epoch's INIT)."""
    c = RegisterCols
    b = ir.IrBuilder()
    its = register_bus_interactions(final_pc_ts)
    n_total = emit_logup_constraints(b, its, 0)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    init, fini = list(init), list(fini)
    return AirProgram("register", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3 if pairs else 0, 1 + absorbed),
                      num_precomputed=c.NUM_PRECOMPUTED, precomputed=lambda: register_precomputed(init, fini))


# PAGE (prover/src/tables/page.rs) and GLOBAL_MEMORY (global_memory.rs): one
# row per byte of a page. Both fix OFFSET and INIT (a private-input page fixes
# OFFSET only, page.rs:86-95); the address is the constant page base plus OFFSET.
PAGE_SIZE = 1 << 18            # page.rs:48 DEFAULT_PAGE_SIZE
GENESIS_EPOCH = 0              # local_to_global.rs:74
BUS_GLOBAL_MEMORY = 31         # types.rs:368


class PageCols:
    """page.rs:61-80."""
    OFFSET, INIT, FINI, TIMESTAMP_LO, TIMESTAMP_HI = range(5)
    NUM_COLUMNS = 5


class GlobalMemoryCols:
    """global_memory.rs:55-80."""
    OFFSET, INIT, FINI, FINI_EPOCH = range(4)
    NUM_COLUMNS = 4


def page_precomputed(init_values, num_fixed: int = 2, page_size: int = PAGE_SIZE) -> np.ndarray:
    """The fixed columns of a page (page.rs:466-493): OFFSET = 0..page_size-1 and,
    unless the page holds private input (num_fixed = 1), INIT = the initial
    bytes, zero past their end."""
    t = np.zeros((page_size, num_fixed), dtype=np.uint64)
    t[:, 0] = np.arange(page_size, dtype=np.uint64)
    if num_fixed > 1 and init_values is not None:
        b = np.frombuffer(bytes(init_values), dtype=np.uint8)
        if len(b) > page_size:
            raise ValueError("initial data exceeds the page")
        t[: len(b), 1] = b
    return t


def _page_base_check(page_base: int, page_size: int):
    if page_base % page_size:
        raise ValueError("page base must be page-aligned")


def generate_page_trace(page_base: int, init_values, final_state: dict, page_size: int = PAGE_SIZE) -> np.ndarray:
    """page.rs:253-300: per offset its INIT byte and, if the address was accessed,
    its final (timestamp, value); else timestamp 0 and FINI = INIT.
    final_state: byte address -> (timestamp, value)."""
    _page_base_check(page_base, page_size)
    c = PageCols
    t = np.zeros((page_size, c.NUM_COLUMNS), dtype=np.uint64)
    t[:, :2] = page_precomputed(init_values, 2, page_size)
    t[:, c.FINI] = t[:, c.INIT]
    for addr, (ts, val) in final_state.items():
        off = addr - page_base
        if not 0 <= off < page_size:
            raise ValueError("final state holds an address outside the page")
        if not 0 <= val < 256:
            raise ValueError("page values are bytes")
        t[off, c.FINI] = val
        t[off, c.TIMESTAMP_LO], t[off, c.TIMESTAMP_HI] = ts & 0xFFFFFFFF, ts >> 32
    return t


def _page_address(page_base: int, offset_col: int):
    """address_lo = page_base_lo + OFFSET, address_hi = page_base_hi (page.rs:588-598)."""
    lo = BusValue.of_linear([LinearTerm(page_base & 0xFFFFFFFF), LinearTerm(1, offset_col)])
    hi = BusValue.of_linear([LinearTerm(page_base >> 32)])
    return lo, hi


def page_bus_interactions(page_base: int):
    """page.rs:600-667: ARE_BYTES[init, fini] sent once per row; the initial
    Memory token [0, address, 0, 0, init] received; the final token
    [0, address, timestamp, fini] sent."""
    c = PageCols
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    k = lambda v: BusValue.of_linear([LinearTerm(v)])  # noqa: E731
    lo, hi = _page_address(page_base, c.OFFSET)
    return [
        BusInteraction.sender(BUS_ARE_BYTES, ONE, [d(c.INIT), d(c.FINI)]),
        BusInteraction.receiver(BUS_MEMORY, ONE, [k(0), lo, hi, k(0), k(0), d(c.INIT)]),
        BusInteraction.sender(BUS_MEMORY, ONE, [k(0), lo, hi, d(c.TIMESTAMP_LO), d(c.TIMESTAMP_HI), d(c.FINI)]),
    ]


def _bus_only_air(name, num_columns, its, num_fixed, precomputed) -> AirProgram:
    b = ir.IrBuilder()
    n_total = emit_logup_constraints(b, its, 0)
    prog = b.finish(num_columns)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram(name, num_columns, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3 if pairs else 0, 1 + absorbed),
                      num_precomputed=num_fixed, precomputed=precomputed)


def page_table_air(page_base: int, init_values=None, private: bool = False, page_size: int = PAGE_SIZE) -> AirProgram:
    """PAGE as an AIR: only the LogUp constraints of its three interactions,
    OFFSET and INIT fixed (OFFSET alone for a private-input page, whose INIT is
    a committed trace column)."""
    _page_base_check(page_base, page_size)
    k = 1 if private else 2
    data = None if init_values is None else bytes(init_values)
    return _bus_only_air("page", PageCols.NUM_COLUMNS, page_bus_interactions(page_base), k,
                         lambda: page_precomputed(data, k, page_size))


def generate_global_memory_trace(page_base: int, init_values, final_state: dict,
                                 page_size: int = PAGE_SIZE) -> np.ndarray:
    """global_memory.rs:107-149: per offset its genesis byte and, if a later
    epoch touched the address, its final (value, epoch); else FINI = INIT at
    GENESIS_EPOCH. final_state: byte address -> (value, epoch)."""
    _page_base_check(page_base, page_size)
    c = GlobalMemoryCols
    t = np.zeros((page_size, c.NUM_COLUMNS), dtype=np.uint64)
    t[:, :2] = page_precomputed(init_values, 2, page_size)
    t[:, c.FINI] = t[:, c.INIT]
    t[:, c.FINI_EPOCH] = GENESIS_EPOCH
    for addr, (val, epoch) in final_state.items():
        off = addr - page_base
        if not 0 <= off < page_size:
            raise ValueError("final state holds an address outside the page")
        if not 0 <= val < 256:
            raise ValueError("page values are bytes")
        t[off, c.FINI], t[off, c.FINI_EPOCH] = val, epoch
    return t


def global_memory_bus_interactions(page_base: int):
    """global_memory.rs:165-213: the genesis token [address, init, GENESIS]
    sent, the finalization token [address, fini, fini_epoch] received, on the
    GlobalMemory bus."""
    c = GlobalMemoryCols
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    lo, hi = _page_address(page_base, c.OFFSET)
    return [
        BusInteraction.sender(BUS_GLOBAL_MEMORY, ONE, [lo, hi, d(c.INIT), BusValue.of_linear([LinearTerm(GENESIS_EPOCH)])]),
        BusInteraction.receiver(BUS_GLOBAL_MEMORY, ONE, [lo, hi, d(c.FINI), d(c.FINI_EPOCH)]),
    ]


def global_memory_table_air(page_base: int, init_values=None, private: bool = False,
                            page_size: int = PAGE_SIZE) -> AirProgram:
    """GLOBAL_MEMORY as an AIR: the LogUp constraints of its two interactions;
    the fixed columns are PAGE's (so is their commitment)."""
    _page_base_check(page_base, page_size)
    k = 1 if private else 2
    data = None if init_values is None else bytes(init_values)
    return _bus_only_air("global_memory", GlobalMemoryCols.NUM_COLUMNS, global_memory_bus_interactions(page_base), k,
                         lambda: page_precomputed(data, k, page_size))


# CPU (prover/src/tables/cpu.rs): its one fixed column and the height rules
# the fixed column depends on (the rest of the CPU AIR is not captured here).
CPU_TIMESTAMP_STRIDE = 4        # cpu.rs:58
CPU_MAX_ROWS = (1 << 30) - 1    # cpu.rs:64


def cpu_timestamp_precomputed(height: int) -> np.ndarray:
    """The fixed TIMESTAMP column of a CPU table of `height` rows: row i holds
    4·(i + 1) (cpu.rs:161-165, :619-622)."""
    return (CPU_TIMESTAMP_STRIDE * np.arange(1, height + 1, dtype=np.uint64)).reshape(height, 1)


def cpu_final_pc_timestamp(total_rows: int) -> int:
    """cpu.rs:72-74: the last row 4·N writes next_pc at 4·N + 1."""
    return CPU_TIMESTAMP_STRIDE * total_rows + 1


def cpu_timestamp_bases(heights) -> list:
    """cpu.rs:78-87: table k adds 4·(h_0 + ... + h_{k-1}) to its timestamps."""
    out, rows = [], 0
    for h in heights:
        out.append(CPU_TIMESTAMP_STRIDE * rows)
        rows += h
    return out


def cpu_table_heights(rows: int, max_rows_cpu: int) -> list:
    """cpu.rs:116-121: full chunks of max_rows_cpu, the last padded to a power
    of two (at least 4)."""
    step = max(1, max_rows_cpu)
    return [max(4, 1 << (min(rows - s, max_rows_cpu) - 1).bit_length()) for s in range(0, rows, step)]


def cpu_validate_heights(heights, max_rows_cpu: int) -> None:
    """cpu.rs:92-111: each height a power of two in [4, max_rows_cpu], at most
    CPU_MAX_ROWS rows in total; ValueError otherwise."""
    total = 0
    for h in heights:
        if h <= 0 or h & (h - 1) or not 4 <= h <= max_rows_cpu:
            raise ValueError(f"CPU table height must be a power of two in [4, {max_rows_cpu}], got {h}")
        total += h
    if total > CPU_MAX_ROWS:
        raise ValueError(f"{total} CPU rows exceed the per-epoch maximum {CPU_MAX_ROWS}")


# HINT (prover/src/tables/hint.rs): the receiver of the `hint` ecall, one row
# per call. The hinted bytes are unconstrained; where they land is not: x12
# (the output address), x10 (the selector) and x11 (the input address) are
# read through the memory argument at the ecall timestamp, the selector and
# both address low limbs are range-checked through ALU LT lookups, the 32
# bytes go to memory as four 8-byte MEMW writes and are range-checked in
# pairs, and mu is a bit.
BUS_ECALL = 19                              # types.rs:310
HINT_SYSCALL_NUMBER = 2**64 - 1 - 31        # executor/src/vm/instruction/execution.rs:171
HINT_SELECTOR_BOUND = 3                     # execution.rs:182
HINT_ADDR_LIMB_BOUND = (1 << 32) - 31       # hint.rs:85


class HintCols:
    """hint.rs:87-117."""
    TIMESTAMP_0, TIMESTAMP_1, ADDR_OUT_0, ADDR_OUT_1 = 0, 1, 2, 3
    OUT = tuple(range(4, 36))
    MU, SEL_0, SEL_1, ADDR_IN_0, ADDR_IN_1 = 36, 37, 38, 39, 40
    NUM_COLUMNS = 41


@dataclass(frozen=True)
class HintOperation:
    """hint.rs:120-131."""
    timestamp: int
    out_addr: int
    out_bytes: bytes
    hint_id: int
    in_addr: int


def generate_hint_trace(ops) -> np.ndarray:
    """hint.rs:136-162: one row per call in program order (mu = 1), zero rows
    of padding to a power of two, at least 4."""
    c = HintCols
    n = max(4, 1 << max(0, (len(ops) - 1).bit_length()))
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)
    for r, op in enumerate(ops):
        if not 0 <= op.timestamp < 2**32:
            raise ValueError("HINT timestamps fit 32 bits")
        if len(op.out_bytes) != 32:
            raise ValueError("a hint writes 32 bytes")
        for col, v in ((c.TIMESTAMP_0, op.timestamp), (c.ADDR_OUT_0, op.out_addr), (c.SEL_0, op.hint_id),
                       (c.ADDR_IN_0, op.in_addr)):
            v &= 2**64 - 1
            t[r, col], t[r, col + 1] = v & 0xFFFFFFFF, v >> 32
        t[r, c.OUT[0]:c.OUT[-1] + 1] = np.frombuffer(bytes(op.out_bytes), dtype=np.uint8)
        t[r, c.MU] = 1
    return t


def hint_bus_interactions():
    """hint.rs:215-317, in its order: the Ecall receiver, the MEMW register reads
    of x12, x10, x11, the three ALU LT range checks, the four MEMW writes of the
    output at out_addr + 8i, the sixteen ARE_BYTES pairs; all with multiplicity mu."""
    c = HintCols
    mu = Multiplicity("Column", (c.MU,))
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    k = lambda v: BusValue.of_linear([LinearTerm(v)])  # noqa: E731

    def register_read(reg, lo, hi):
        # hint.rs:188-205: [old[8], 1, 2*reg, 0, value[8], ts, w2=1, w4=0, w8=0], old == value
        value = [d(lo), d(hi)] + [k(0)] * 6
        return value + [k(1), k(2 * reg), k(0)] + value + [d(c.TIMESTAMP_0), d(c.TIMESTAMP_1), k(1), k(0), k(0)]

    its = [BusInteraction.receiver(BUS_ECALL, mu, [d(c.TIMESTAMP_0), d(c.TIMESTAMP_1),
                                                   k(HINT_SYSCALL_NUMBER & 0xFFFFFFFF), k(HINT_SYSCALL_NUMBER >> 32)])]
    its += [BusInteraction.sender(BUS_MEMW, mu, register_read(reg, lo, hi))
            for reg, lo, hi in ((12, c.ADDR_OUT_0, c.ADDR_OUT_1), (10, c.SEL_0, c.SEL_1), (11, c.ADDR_IN_0, c.ADDR_IN_1))]
    its.append(BusInteraction.sender(BUS_ALU, mu, [BusValue.packed(Packing.DWORD_WL, c.SEL_0), k(HINT_SELECTOR_BOUND),
                                                   k(0), k(ALU_OP_LT), k(1), k(0)]))
    its += [BusInteraction.sender(BUS_ALU, mu, [d(lo), k(0), k(HINT_ADDR_LIMB_BOUND), k(0), k(ALU_OP_LT), k(1), k(0)])
            for lo in (c.ADDR_IN_0, c.ADDR_OUT_0)]
    for i in range(4):
        base_lo = BusValue.of_linear([LinearTerm(1, c.ADDR_OUT_0), LinearTerm(8 * i)])
        its.append(BusInteraction.sender(BUS_MEMW, mu, [k(0), base_lo, d(c.ADDR_OUT_1)]
                                         + [d(c.OUT[8 * i + b]) for b in range(8)]
                                         + [d(c.TIMESTAMP_0), d(c.TIMESTAMP_1), k(0), k(0), k(1)]))
    its += [BusInteraction.sender(BUS_ARE_BYTES, mu, [d(c.OUT[2 * i]), d(c.OUT[2 * i + 1])]) for i in range(16)]
    return its


def hint_table_air() -> AirProgram:
    """HINT as an AIR: IS_BIT(mu) (hint.rs:355-372) and the LogUp constraints of
    its 27 interactions."""
    c = HintCols
    b = ir.IrBuilder()
    _emit_is_bit(b, 0, c.MU)
    its = hint_bus_interactions()
    n_total = emit_logup_constraints(b, its, 1)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("hint", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3 if pairs else 0, 1 + absorbed, 2))


# KECCAK core (prover/src/tables/keccak.rs): one row per keccak-f call. It
# receives the ecall, reads x10 (the state address), hands the input state to
# the round chip and takes its output back on the Keccak bus, reads and writes
# the 25 lanes through MEMW at state_addr + 8·lane, and range-checks the lane
# addresses (IS_WH + IS_HALFWORD), the alignment (BYTE_ALU AND 7 = 0) and the
# address bytes (ARE_BYTES).
BUS_KECCAK = 22                        # types.rs:318
KECCAK_SYSCALL_NUMBER = 2**64 - 1 - 1  # executor/src/vm/instruction/execution.rs:53


class KeccakCols:
    """keccak.rs:31-76: timestamp (2), addr (8 bytes), input and output state
    (200 bytes each, lane x + 5y at 8·(x + 5y)), state_ptr (25 × DWordWHH), mu."""
    TIMESTAMP_0, TIMESTAMP_1, ADDR = 0, 1, 2
    INPUT_STATE = 10
    OUTPUT_STATE = 210
    STATE_PTR = 410
    MU = 485
    NUM_COLUMNS = 486

    @staticmethod
    def state_ptr(lane: int, limb: int) -> int:
        return KeccakCols.STATE_PTR + 3 * lane + limb


@dataclass(frozen=True)
class KeccakOperation:
    """keccak.rs:82-88: lanes indexed x + 5y."""
    timestamp: int
    state_addr: int
    input: tuple
    output: tuple


def keccak_is_wh_keys(ops) -> list:
    """keccak.rs:92-104: the IS_WH key of every lane address of every call."""
    return [is_wh_key(op.state_addr + 8 * lane) for op in ops for lane in range(25)]


def generate_keccak_trace(ops) -> np.ndarray:
    """keccak.rs:110-167: one row per call (mu = 1); padding rows hold only
    state_ptr[lane] half_lo = 8·lane."""
    c = KeccakCols
    n = max(4, 1 << max(0, (len(ops) - 1).bit_length()))
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)

    def le_bytes(v):
        return np.frombuffer(int(v).to_bytes(8, "little"), dtype=np.uint8)

    for r, op in enumerate(ops):
        if op.state_addr + 8 * 24 >= 2**64:
            raise ValueError("keccak state address range overflows")
        t[r, c.TIMESTAMP_0], t[r, c.TIMESTAMP_1] = op.timestamp & 0xFFFFFFFF, op.timestamp >> 32
        t[r, c.ADDR:c.ADDR + 8] = le_bytes(op.state_addr)
        for lane in range(25):
            t[r, c.INPUT_STATE + 8 * lane:c.INPUT_STATE + 8 * lane + 8] = le_bytes(op.input[lane])
            t[r, c.OUTPUT_STATE + 8 * lane:c.OUTPUT_STATE + 8 * lane + 8] = le_bytes(op.output[lane])
            ptr = op.state_addr + 8 * lane
            t[r, c.state_ptr(lane, 0)], t[r, c.state_ptr(lane, 1)] = ptr & 0xFFFF, (ptr >> 16) & 0xFFFF
            t[r, c.state_ptr(lane, 2)] = ptr >> 32
        t[r, c.MU] = 1
    for r in range(len(ops), n):
        for lane in range(25):
            t[r, c.state_ptr(lane, 0)] = 8 * lane
    return t


def keccak_bus_interactions():
    """keccak.rs:175-440, in its order (84 interactions, all with multiplicity
    mu): Ecall receiver; MEMW read of x10 with the address from its bytes;
    Keccak send [ts, 0, input bytes] and receive [ts, 24, output bytes] (the
    state bytes x-major: lane x + 5y for x, then y); per
    lane IS_WH[word_hi, half_mid] and IS_HALFWORD[half_lo]; BYTE_ALU[AND,
    addr[0], 7, 0]; four ARE_BYTES pairs of the address bytes; per lane the
    MEMW read-write [input bytes, 0, address, output bytes, ts, 0, 0, 1]."""
    c = KeccakCols
    mu = Multiplicity("Column", (c.MU,))
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    k = lambda v: BusValue.of_linear([LinearTerm(v)])  # noqa: E731
    ts = [d(c.TIMESTAMP_0), d(c.TIMESTAMP_1)]
    its = [BusInteraction.receiver(BUS_ECALL, mu, ts + [k(KECCAK_SYSCALL_NUMBER & 0xFFFFFFFF),
                                                         k(KECCAK_SYSCALL_NUMBER >> 32)])]
    addr_lo = BusValue.of_linear([LinearTerm(256 ** i, c.ADDR + i) for i in range(4)])
    addr_hi = BusValue.of_linear([LinearTerm(256 ** i, c.ADDR + 4 + i) for i in range(4)])
    word = [addr_lo, addr_hi] + [k(0)] * 6
    its.append(BusInteraction.sender(BUS_MEMW, mu, word + [k(1), k(20), k(0)] + word + ts + [k(1), k(0), k(0)]))
    # state bytes in x-major order: lane x + 5y for x, then y, then byte (keccak.rs:290-334)
    order = [8 * (x + 5 * y) + byte for x in range(5) for y in range(5) for byte in range(8)]
    its.append(BusInteraction.sender(BUS_KECCAK, mu, ts + [k(0)] + [d(c.INPUT_STATE + i) for i in order]))
    its.append(BusInteraction.receiver(BUS_KECCAK, mu, ts + [k(24)] + [d(c.OUTPUT_STATE + i) for i in order]))
    for lane in range(25):
        its.append(BusInteraction.sender(BUS_IS_WH, mu, [d(c.state_ptr(lane, 2)), d(c.state_ptr(lane, 1))]))
        its.append(BusInteraction.sender(BUS_IS_HALFWORD, mu, [d(c.state_ptr(lane, 0))]))
    its.append(BusInteraction.sender(BUS_BYTE_ALU, mu, [k(ALU_OP_AND), d(c.ADDR), k(7), k(0)]))
    its += [BusInteraction.sender(BUS_ARE_BYTES, mu, [d(c.ADDR + 2 * i), d(c.ADDR + 2 * i + 1)]) for i in range(4)]
    for lane in range(25):
        lo = BusValue.of_linear([LinearTerm(1, c.state_ptr(lane, 0)), LinearTerm(1 << 16, c.state_ptr(lane, 1))])
        its.append(BusInteraction.sender(BUS_MEMW, mu, [d(c.INPUT_STATE + 8 * lane + b) for b in range(8)]
                                         + [k(0), lo, d(c.state_ptr(lane, 2))]
                                         + [d(c.OUTPUT_STATE + 8 * lane + b) for b in range(8)]
                                         + ts + [k(0), k(0), k(1)]))
    return its


def keccak_table_air() -> AirProgram:
    """KECCAK as an AIR (keccak.rs:454-524): 0-49 the gated ADD carry pair of
    addr (DWordBL) + 8·lane = state_ptr[lane] (DWordWHH) per lane, each
    mu·carry·(1 - carry); 50 mu·carry_1 of addr + 192 = state_ptr[24] (no
    overflow); 51 mu a bit; then the LogUp constraints of its 84 interactions."""
    c = KeccakCols
    b = ir.IrBuilder()
    inv = b.const_base(INV_SHIFT_32)
    one = b.const_base(1)
    m = lambda col: b.main(0, col)  # noqa: E731
    mu = m(c.MU)
    addr_lo = m(c.ADDR) + m(c.ADDR + 1) * b.const_base(256) + m(c.ADDR + 2) * b.const_base(65536) \
        + m(c.ADDR + 3) * b.const_base(1 << 24)
    addr_hi = m(c.ADDR + 4) + m(c.ADDR + 5) * b.const_base(256) + m(c.ADDR + 6) * b.const_base(65536) \
        + m(c.ADDR + 7) * b.const_base(1 << 24)
    carry1_top = None
    for lane in range(25):
        ptr_lo = m(c.state_ptr(lane, 0)) + m(c.state_ptr(lane, 1)) * b.const_base(1 << 16)
        ptr_hi = m(c.state_ptr(lane, 2))
        carry0 = (addr_lo + b.const_base(8 * lane) - ptr_lo) * inv
        carry1 = (addr_hi + carry0 - ptr_hi) * inv
        b.emit(2 * lane, mu * carry0 * (one - carry0))
        b.emit(2 * lane + 1, mu * carry1 * (one - carry1))
        carry1_top = carry1
    b.emit(50, mu * carry1_top)
    b.emit(51, mu * (one - mu))
    its = keccak_bus_interactions()
    n_total = emit_logup_constraints(b, its, 52)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("keccak", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3 if pairs else 0, 1 + absorbed, 3))


# MEMW_A (prover/src/tables/memw_aligned.rs): aligned memory and register
# accesses whose cells share one old timestamp. Per access it range-checks the
# address span (IS_HALFWORD), asserts old_ts < ts (ALU LT), consumes the old
# token and produces the new one on the Memory bus for each cell the width
# covers, and receives the MEMW read or write.
@dataclass(frozen=True)
class MemwOperation:
    """memw.rs:105-126 (aligned: one old timestamp). value/old: eight cells —
    bytes for memory, 32-bit register halves (two cells, width 2) for registers."""
    is_register: bool
    base_address: int
    value: tuple
    timestamp: int
    width: int
    is_read: bool
    old: tuple
    old_timestamp: int

    def write_flags(self):
        """memw.rs:165-173."""
        return {1: (0, 0, 0), 2: (1, 0, 0), 4: (0, 1, 0), 8: (0, 0, 1)}.get(self.width, (0, 0, 0))


class MemwAlignedCols:
    """memw_aligned.rs:53-83."""
    IS_REGISTER = 0
    BASE_ADDRESS = (1, 2, 3)
    VALUE = tuple(range(4, 12))
    TIMESTAMP_0, TIMESTAMP_1, WRITE2, WRITE4, WRITE8 = 12, 13, 14, 15, 16
    OLD = tuple(range(17, 25))
    OLD_TIMESTAMP_0, OLD_TIMESTAMP_1, MU_READ, MU_WRITE = 25, 26, 27, 28
    NUM_COLUMNS = 29


def generate_memw_aligned_trace(ops) -> np.ndarray:
    """memw_aligned.rs:91-132: one row per access, padded with zero rows to a
    power of two (at least 4)."""
    c = MemwAlignedCols
    n = max(4, 1 << max(0, (len(ops) - 1).bit_length()))
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)
    for r, op in enumerate(ops):
        ba = op.base_address & (2**64 - 1)
        t[r, c.IS_REGISTER] = int(op.is_register)
        t[r, c.BASE_ADDRESS[0]], t[r, c.BASE_ADDRESS[1]], t[r, c.BASE_ADDRESS[2]] = \
            ba & 0xFFFF, (ba >> 16) & 0xFFFF, ba >> 32
        t[r, c.VALUE[0]:c.VALUE[-1] + 1] = list(op.value)
        t[r, c.TIMESTAMP_0], t[r, c.TIMESTAMP_1] = op.timestamp & 0xFFFFFFFF, op.timestamp >> 32
        t[r, c.WRITE2], t[r, c.WRITE4], t[r, c.WRITE8] = op.write_flags()
        t[r, c.OLD[0]:c.OLD[-1] + 1] = list(op.old)
        t[r, c.OLD_TIMESTAMP_0], t[r, c.OLD_TIMESTAMP_1] = op.old_timestamp & 0xFFFFFFFF, op.old_timestamp >> 32
        t[r, c.MU_READ], t[r, c.MU_WRITE] = int(op.is_read), int(not op.is_read)
    return t


def memw_aligned_bus_interactions():
    """memw_aligned.rs:138-661, in its order (20): IS_HALFWORD[ba0 + w2 + 3 w4 +
    7 w8] and ALU[old_ts, ts, LT, 1, 0] with mu_read + mu_write; the Memory
    tokens of cell 0 (mu_read + mu_write), cell 1 (w2 + w4 + w8), cells 2-3
    (w4 + w8) and 4-7 (w8), each the old token [is_register, address + i,
    old_ts, old_i] sent and the new one [.., ts, value_i] received; the MEMW read
    receiver [old, is_register, address, value, ts, w2, w4, w8] (mu_read) and
    write receiver [is_register, address, value, ts, w2, w4, w8] (mu_write)."""
    c = MemwAlignedCols
    mu_sum = Multiplicity("Sum", (c.MU_READ, c.MU_WRITE))
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    k = lambda v: BusValue.of_linear([LinearTerm(v)])  # noqa: E731
    its = [BusInteraction.sender(BUS_IS_HALFWORD, mu_sum, [BusValue.of_linear(
               [LinearTerm(1, c.BASE_ADDRESS[0]), LinearTerm(1, c.WRITE2), LinearTerm(3, c.WRITE4),
                LinearTerm(7, c.WRITE8)])]),
           BusInteraction.sender(BUS_ALU, mu_sum, [BusValue.packed(Packing.DWORD_WL, c.OLD_TIMESTAMP_0),
                                                   BusValue.packed(Packing.DWORD_WL, c.TIMESTAMP_0),
                                                   k(ALU_OP_LT), k(1), k(0)])]
    hi = d(c.BASE_ADDRESS[2])

    def lo(i):
        terms = [LinearTerm(1, c.BASE_ADDRESS[0]), LinearTerm(1 << 16, c.BASE_ADDRESS[1])]
        return BusValue.of_linear(terms + ([LinearTerm(i)] if i else []))

    cover = [mu_sum, Multiplicity("Sum3", (c.WRITE2, c.WRITE4, c.WRITE8)),
             Multiplicity("Sum", (c.WRITE4, c.WRITE8)), Multiplicity("Sum", (c.WRITE4, c.WRITE8))] + \
        [Multiplicity("Column", (c.WRITE8,))] * 4
    for i in range(8):
        its.append(BusInteraction.sender(BUS_MEMORY, cover[i], [d(c.IS_REGISTER), lo(i), hi, d(c.OLD_TIMESTAMP_0),
                                                                d(c.OLD_TIMESTAMP_1), d(c.OLD[i])]))
        its.append(BusInteraction.receiver(BUS_MEMORY, cover[i], [d(c.IS_REGISTER), lo(i), hi, d(c.TIMESTAMP_0),
                                                                  d(c.TIMESTAMP_1), d(c.VALUE[i])]))
    tail = [d(c.TIMESTAMP_0), d(c.TIMESTAMP_1), d(c.WRITE2), d(c.WRITE4), d(c.WRITE8)]
    access = [d(c.IS_REGISTER), lo(0), hi] + [d(v) for v in c.VALUE] + tail
    its.append(BusInteraction.receiver(BUS_MEMW, Multiplicity("Column", (c.MU_READ,)),
                                       [d(o) for o in c.OLD] + access))
    its.append(BusInteraction.receiver(BUS_MEMW, Multiplicity("Column", (c.MU_WRITE,)), list(access)))
    return its


def memw_aligned_table_air() -> AirProgram:
    """MEMW_A as an AIR (memw_aligned.rs:671-699): 0 mu_sum a bit, 1 w2·(1 -
    mu_sum), 2-3 mu_read and mu_write bits, 4-6 the width flags bits, 7 their
    sum a bit; then the LogUp constraints of its 20 interactions."""
    c = MemwAlignedCols
    b = ir.IrBuilder()
    one = b.const_base(1)
    mu_sum = b.main(0, c.MU_READ) + b.main(0, c.MU_WRITE)
    w2 = b.main(0, c.WRITE2) + b.main(0, c.WRITE4) + b.main(0, c.WRITE8)
    b.emit(0, mu_sum * (one - mu_sum))
    b.emit(1, w2 * (one - mu_sum))
    for idx, col in enumerate((c.MU_READ, c.MU_WRITE, c.WRITE2, c.WRITE4, c.WRITE8)):
        _emit_is_bit(b, 2 + idx, col)
    b.emit(7, w2 * (one - w2))
    its = memw_aligned_bus_interactions()
    n_total = emit_logup_constraints(b, its, 8)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("memw_aligned", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False),
                      [0] * n_total, boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3 if pairs else 0, 1 + absorbed, 2))


# MUL (prover/src/tables/mul.rs): 64 x 64 -> 128-bit products, signed or
# unsigned per operand, deduplicated with lo/hi multiplicities. The product
# is pinned by four raw convolution columns (sums of 16-bit limb products of
# the sign-extended operands) whose carries into the 32-bit result words are
# virtual and range-checked through IS_B20.
ALU_OP_MUL = 7        # types.rs:569
MUL_SIGN_FILL = 0xFFFF


def _inv2(k: int) -> int:
    return pow(2, P - 1 - k, P)


class MulCols:
    """mul.rs:57-125."""
    LHS = (0, 1, 2, 3)
    LHS_SIGNED = 4
    RHS = (5, 6, 7, 8)
    RHS_SIGNED = 9
    LO = (10, 11, 12, 13)
    HI = (14, 15, 16, 17)
    LHS_IS_NEGATIVE, RHS_IS_NEGATIVE = 18, 19
    RAW_PRODUCT = (20, 21, 22, 23)
    MU_LO, MU_HI = 24, 25
    NUM_COLUMNS = 26


@dataclass(frozen=True)
class MulOperation:
    """mul.rs:149-246."""
    lhs: int
    lhs_signed: bool
    rhs: int
    rhs_signed: bool

    def _value(self, v, signed):
        v &= 2**64 - 1
        return v - 2**64 if signed and v >> 63 else v

    def product(self):
        p = (self._value(self.lhs, self.lhs_signed) * self._value(self.rhs, self.rhs_signed)) % 2**128
        return p & (2**64 - 1), p >> 64

    def lhs_is_negative(self):
        return self.lhs_signed and (self.lhs & (2**64 - 1)) >> 63 == 1

    def rhs_is_negative(self):
        return self.rhs_signed and (self.rhs & (2**64 - 1)) >> 63 == 1

    @staticmethod
    def _extended(v, neg):
        fill = MUL_SIGN_FILL if neg else 0
        return [(v >> (16 * i)) & 0xFFFF for i in range(4)] + [fill] * 4

    def raw_products(self):
        """mul.rs:204-231: raw[i] = sum_{k<2} 2^(16k) sum_j lhs_ext[j] rhs_ext[2i+k-j], mod 2^64."""
        a = self._extended(self.lhs & (2**64 - 1), self.lhs_is_negative())
        b = self._extended(self.rhs & (2**64 - 1), self.rhs_is_negative())
        out = []
        for i in range(4):
            s = sum((a[j] * b[2 * i + k - j]) << (16 * k) for k in range(2) for j in range(2 * i + k + 1))
            out.append(s % 2**64)
        return out


def dedup_mul_rows(ops):
    """mul.rs:260-295: (operation, wants_hi) pairs merged per operation with lo/hi
    counts, split at MU_MAX; rows follow first occurrence (the reference's map
    order is arbitrary)."""
    counts = {}
    for op, wants_hi in ops:
        lo, hi = counts.get(op, (0, 0))
        counts[op] = (lo, hi + 1) if wants_hi else (lo + 1, hi)
    rows = []
    for op, (lo, hi) in counts.items():
        while lo > 0 or hi > 0:
            rl, rh = min(lo, MU_MAX), min(hi, MU_MAX)
            lo, hi = lo - rl, hi - rh
            rows.append((op, rl, rh))
    return rows


def generate_mul_trace(ops) -> np.ndarray:
    """mul.rs:330-383."""
    c = MulCols
    rows = dedup_mul_rows(ops)
    n = max(4, 1 << max(0, (len(rows) - 1).bit_length()))
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)
    for r, (op, mu_lo, mu_hi) in enumerate(rows):
        lo, hi = op.product()
        for cols, v in ((c.LHS, op.lhs & (2**64 - 1)), (c.RHS, op.rhs & (2**64 - 1)), (c.LO, lo), (c.HI, hi)):
            for i, col in enumerate(cols):
                t[r, col] = (v >> (16 * i)) & 0xFFFF
        t[r, c.LHS_SIGNED], t[r, c.RHS_SIGNED] = int(op.lhs_signed), int(op.rhs_signed)
        t[r, c.LHS_IS_NEGATIVE], t[r, c.RHS_IS_NEGATIVE] = int(op.lhs_is_negative()), int(op.rhs_is_negative())
        t[r, c.RAW_PRODUCT[0]:c.RAW_PRODUCT[-1] + 1] = op.raw_products()
        t[r, c.MU_LO], t[r, c.MU_HI] = mu_lo, mu_hi
    return t


def mul_bus_interactions():
    """mul.rs:390-728, in its order (25): MSB16 of each operand's top half (with
    its signed flag), ARE_BYTES[mu_lo, mu_hi] once, IS_HALFWORD of the eight
    operand halves and the eight result halves and IS_B20 of the four virtual
    carries (all with mu_lo + mu_hi), then the ALU receivers of the lo result
    (mu_lo) and the hi result (mu_hi, flags + 128)."""
    c = MulCols
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    row = Multiplicity("Sum", (c.MU_LO, c.MU_HI))
    its = [BusInteraction.sender(BUS_MSB16, Multiplicity("Column", (c.LHS_SIGNED,)), [d(c.LHS[3]), d(c.LHS_IS_NEGATIVE)]),
           BusInteraction.sender(BUS_MSB16, Multiplicity("Column", (c.RHS_SIGNED,)), [d(c.RHS[3]), d(c.RHS_IS_NEGATIVE)]),
           BusInteraction.sender(BUS_ARE_BYTES, ONE, [d(c.MU_LO), d(c.MU_HI)])]
    its += [BusInteraction.sender(BUS_IS_HALFWORD, row, [d(col)]) for col in c.LHS + c.RHS + c.LO + c.HI]
    # carry[i] = sum_{j<=i} 2^(-32(i-j+1)) raw[j] - sum over result halves 2^(16 h - 32(i+1)) res_h (h < 2i+2)
    res = c.LO + c.HI
    for i in range(4):
        terms = [LinearTerm(_inv2(32 * (i - j + 1)), c.RAW_PRODUCT[j], True) for j in range(i, -1, -1)]
        terms += [LinearTerm(P - _inv2(32 * (i + 1) - 16 * h), res[h], True) for h in range(2 * i + 2)]
        its.append(BusInteraction.sender(BUS_IS_B20, row, [BusValue.of_linear(terms)]))

    def flags(hi):
        return BusValue.of_linear([LinearTerm(ALU_OP_MUL + hi), LinearTerm(32, c.LHS_SIGNED),
                                   LinearTerm(64, c.RHS_SIGNED)])

    for mu, hi, out in ((c.MU_LO, 0, c.LO[0]), (c.MU_HI, 128, c.HI[0])):
        its.append(BusInteraction.receiver(BUS_ALU, Multiplicity("Column", (mu,)), [
            BusValue.packed(Packing.DWORD_HL, c.LHS[0]), BusValue.packed(Packing.DWORD_HL, c.RHS[0]), flags(hi),
            BusValue.packed(Packing.DWORD_HL, out)]))
    return its


def mul_table_air() -> AirProgram:
    """MUL as an AIR (mul.rs:790-818): 0-1 the signed flags are bits, 2-3 an
    unsigned operand is not negative, 4-7 raw_product[i] equals the convolution
    of the sign-extended halves; then the LogUp constraints of its 25
    interactions."""
    c = MulCols
    b = ir.IrBuilder()
    one = b.const_base(1)
    m = lambda col: b.main(0, col)  # noqa: E731
    _emit_is_bit(b, 0, c.LHS_SIGNED)
    _emit_is_bit(b, 1, c.RHS_SIGNED)
    b.emit(2, (one - m(c.LHS_SIGNED)) * m(c.LHS_IS_NEGATIVE))
    b.emit(3, (one - m(c.RHS_SIGNED)) * m(c.RHS_IS_NEGATIVE))
    fill = b.const_base(MUL_SIGN_FILL)
    a = [m(col) for col in c.LHS] + [fill * m(c.LHS_IS_NEGATIVE)] * 4
    bb = [m(col) for col in c.RHS] + [fill * m(c.RHS_IS_NEGATIVE)] * 4
    for i in range(4):
        s = None
        for k in range(2):
            idx = 2 * i + k
            inner = None
            for j in range(idx + 1):
                t = a[j] * bb[idx - j]
                inner = t if inner is None else inner + t
            inner = inner if k == 0 else inner * b.const_base(1 << 16)
            s = inner if s is None else s + inner
        b.emit(4 + i, m(c.RAW_PRODUCT[i]) - s)
    its = mul_bus_interactions()
    n_total = emit_logup_constraints(b, its, 8)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("mul", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3 if pairs else 0, 1 + absorbed, 2))


# MEMW (prover/src/tables/memw.rs): the general memory access — unaligned, or
# with cells of different old timestamps. The base address is a DWordWL, cell
# i > 0 lives at base + i with a carry bit into the high word, and every cell
# carries its own old timestamp (and its own old_ts < ts check).
class MemwCols:
    """memw.rs:51-98."""
    IS_REGISTER, BASE_ADDRESS_0, BASE_ADDRESS_1 = 0, 1, 2
    VALUE = tuple(range(3, 11))
    TIMESTAMP_0, TIMESTAMP_1, WRITE2, WRITE4, WRITE8 = 11, 12, 13, 14, 15
    OLD = tuple(range(16, 24))
    CARRY = tuple(range(24, 31))
    OLD_TIMESTAMP_START = 31
    MU_READ, MU_WRITE = 47, 48
    NUM_COLUMNS = 49

    @staticmethod
    def old_timestamp(i: int) -> int:
        return MemwCols.OLD_TIMESTAMP_START + 2 * i


@dataclass(frozen=True)
class MemwGeneralOperation:
    """memw.rs:105-126: one old timestamp per cell."""
    is_register: bool
    base_address: int
    value: tuple
    timestamp: int
    width: int
    is_read: bool
    old: tuple
    old_timestamp: tuple

    def write_flags(self):
        return {1: (0, 0, 0), 2: (1, 0, 0), 4: (0, 1, 0), 8: (0, 0, 1)}.get(self.width, (0, 0, 0))


def generate_memw_trace(ops) -> np.ndarray:
    """memw.rs:177-233: carry[i] = 1 when the low word of base + i + 1 wraps."""
    c = MemwCols
    n = max(4, 1 << max(0, (len(ops) - 1).bit_length()))
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)
    for r, op in enumerate(ops):
        ba = op.base_address & (2**64 - 1)
        lo = ba & 0xFFFFFFFF
        t[r, c.IS_REGISTER] = int(op.is_register)
        t[r, c.BASE_ADDRESS_0], t[r, c.BASE_ADDRESS_1] = lo, ba >> 32
        t[r, c.VALUE[0]:c.VALUE[-1] + 1] = list(op.value)
        t[r, c.TIMESTAMP_0], t[r, c.TIMESTAMP_1] = op.timestamp & 0xFFFFFFFF, op.timestamp >> 32
        t[r, c.WRITE2], t[r, c.WRITE4], t[r, c.WRITE8] = op.write_flags()
        t[r, c.OLD[0]:c.OLD[-1] + 1] = list(op.old)
        for i in range(7):
            t[r, c.CARRY[i]] = int(lo + i + 1 >= 1 << 32)
        for i in range(8):
            ots = op.old_timestamp[i]
            t[r, c.old_timestamp(i)], t[r, c.old_timestamp(i) + 1] = ots & 0xFFFFFFFF, ots >> 32
        t[r, c.MU_READ], t[r, c.MU_WRITE] = int(op.is_read), int(not op.is_read)
    return t


def memw_bus_interactions():
    """memw.rs:242-828, in its order (26): per cell the old Memory token sent
    and the new one received (cell 0 with mu_read + mu_write, cell 1 with w2 +
    w4 + w8, cells 2-3 with w4 + w8, cells 4-7 with w8; cell i > 0 at
    [base_lo + i - 2^32 carry[i-1], base_hi + carry[i-1]]); the MEMW read and
    write receivers; per cell ALU[old_ts_i, ts, LT, 1, 0] with the cell's
    multiplicity."""
    c = MemwCols
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    k = lambda v: BusValue.of_linear([LinearTerm(v)])  # noqa: E731
    cover = [Multiplicity("Sum", (c.MU_READ, c.MU_WRITE)), Multiplicity("Sum3", (c.WRITE2, c.WRITE4, c.WRITE8)),
             Multiplicity("Sum", (c.WRITE4, c.WRITE8)), Multiplicity("Sum", (c.WRITE4, c.WRITE8))] + \
        [Multiplicity("Column", (c.WRITE8,))] * 4

    def address(i):
        if i == 0:
            return [d(c.BASE_ADDRESS_0), d(c.BASE_ADDRESS_1)]
        return [BusValue.of_linear([LinearTerm(1, c.BASE_ADDRESS_0), LinearTerm(i), LinearTerm(-(1 << 32), c.CARRY[i - 1])]),
                BusValue.of_linear([LinearTerm(1, c.BASE_ADDRESS_1), LinearTerm(1, c.CARRY[i - 1])])]

    its = []
    for i in range(8):
        ots = c.old_timestamp(i)
        its.append(BusInteraction.sender(BUS_MEMORY, cover[i], [d(c.IS_REGISTER)] + address(i) +
                                         [d(ots), d(ots + 1), d(c.OLD[i])]))
        its.append(BusInteraction.receiver(BUS_MEMORY, cover[i], [d(c.IS_REGISTER)] + address(i) +
                                           [d(c.TIMESTAMP_0), d(c.TIMESTAMP_1), d(c.VALUE[i])]))
    access = [d(c.IS_REGISTER), d(c.BASE_ADDRESS_0), d(c.BASE_ADDRESS_1)] + [d(v) for v in c.VALUE] + \
        [d(c.TIMESTAMP_0), d(c.TIMESTAMP_1), d(c.WRITE2), d(c.WRITE4), d(c.WRITE8)]
    its.append(BusInteraction.receiver(BUS_MEMW, Multiplicity("Column", (c.MU_READ,)), [d(o) for o in c.OLD] + access))
    its.append(BusInteraction.receiver(BUS_MEMW, Multiplicity("Column", (c.MU_WRITE,)), list(access)))
    for i in range(8):
        its.append(BusInteraction.sender(BUS_ALU, cover[i], [BusValue.packed(Packing.DWORD_WL, c.old_timestamp(i)),
                                                             BusValue.packed(Packing.DWORD_WL, c.TIMESTAMP_0),
                                                             k(ALU_OP_LT), k(1), k(0)]))
    return its


def memw_table_air() -> AirProgram:
    """MEMW as an AIR (memw.rs:852-902): 0 mu_sum a bit, 1 w2·(1 - mu_sum),
    2-3 mu_read and mu_write bits, 4-10 the carries bits, 11-13 the width flags
    bits, 14 their sum a bit; then the LogUp constraints of its 26 interactions."""
    c = MemwCols
    b = ir.IrBuilder()
    one = b.const_base(1)
    mu_sum = b.main(0, c.MU_READ) + b.main(0, c.MU_WRITE)
    w2 = b.main(0, c.WRITE2) + b.main(0, c.WRITE4) + b.main(0, c.WRITE8)
    b.emit(0, mu_sum * (one - mu_sum))
    b.emit(1, w2 * (one - mu_sum))
    for idx, col in enumerate((c.MU_READ, c.MU_WRITE) + c.CARRY + (c.WRITE2, c.WRITE4, c.WRITE8)):
        _emit_is_bit(b, 2 + idx, col)
    b.emit(14, w2 * (one - w2))
    its = memw_bus_interactions()
    n_total = emit_logup_constraints(b, its, 15)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("memw", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3 if pairs else 0, 1 + absorbed, 2))


# SHIFT (prover/src/tables/shift.rs): SLL/SRL/SRA and their word forms. An
# intra-limb shift by bit_shift = shift mod 16 through paired HWSL lookups
# ([SLL, SLLC] of each input half), then a whole-limb shift by the one-hot
# limb_shift (its last entry virtual).
ALU_OP_SHIFT = 5   # types.rs:567


class ShiftCols:
    """shift.rs:31-99."""
    IN = (0, 1, 2, 3)
    SHIFT_AMOUNT, DIRECTION, SIGNED, WORD_INSTR = 4, 5, 6, 7
    OUT_0, OUT_1 = 8, 9
    IS_NEGATIVE, BIT_SHIFT, ZBS = 10, 11, 12
    X = (13, 14, 15, 16, 17)
    Y = (18, 19, 20, 21)
    LIMB_SHIFT_RAW = (22, 23, 24)
    MU = 25
    SHIFT_B1, SHIFT_H1, SHIFT_HIGH = 26, 27, 28
    NUM_COLUMNS = 29


@dataclass(frozen=True)
class ShiftOperation:
    """shift.rs:105-145: the input value, the full shift operand (only its low
    byte shifts), left (direction 0) or right, arithmetic, word instruction."""
    value: int
    shift_amount: int
    direction: bool
    signed: bool
    word_instr: bool

    def halves(self):
        v = self.value & (2**64 - 1)
        return [(v >> (16 * i)) & 0xFFFF for i in range(4)]

    def result(self) -> int:
        """shift.rs:170-205 (the architectural result, word forms sign-extended)."""
        v, s = self.value & (2**64 - 1), self.shift_amount & 0xFF
        sh = s % (32 if self.word_instr else 64)
        if self.word_instr:
            lo = v & 0xFFFFFFFF
            if not self.direction:
                r = (lo << sh) & 0xFFFFFFFF
            elif not self.signed:
                r = lo >> sh
            else:
                r = ((lo - (1 << 32) if lo >> 31 else lo) >> sh) & 0xFFFFFFFF
            return (r - (1 << 32) if r >> 31 else r) & (2**64 - 1)
        if not self.direction:
            return (v << sh) & (2**64 - 1)
        if not self.signed:
            return v >> sh
        return ((v - 2**64 if v >> 63 else v) >> sh) & (2**64 - 1)

    def aux(self):
        """shift.rs:218-339: is_negative, bit_shift, zbs, X, Y, limb_shift, out words."""
        h = self.halves()
        s = self.shift_amount & 0xFF
        left = not self.direction
        is_neg = self.signed and (h[3] >> 15) & 1 == 1
        ext = 0xFFFF if is_neg else 0
        bit_shift = (s & 15) if left else ((256 - s) & 15)
        zbs = bit_shift == 0
        x, y = [0] * 5, [0] * 4
        if zbs:
            for i in range(4):
                if left:
                    x[i] = h[i]
                else:
                    y[i] = h[i]
        else:
            for i in range(4):
                x[i] = (h[i] << bit_shift) & 0xFFFF
                y[i] = h[i] >> (16 - bit_shift)
            x[4] = (ext << bit_shift) & 0xFFFF
        limb = ((s >> 4) & 1) if self.word_instr else ((s >> 4) & 3)
        ls = [int(j == limb) for j in range(4)]
        il = lambda i: x[0] if i == 0 else (x[i] + y[i - 1]) & 0xFFFF  # noqa: E731
        ir_ = lambda i: (y[i] + x[i + 1]) & 0xFFFF  # noqa: E731
        shifted = []
        for i in range(4):
            val = 0
            if left:
                for j in range(i + 1):
                    if ls[j]:
                        val = (val + il(i - j)) & 0xFFFF
            else:
                for j in range(4 - i):
                    if ls[j]:
                        val = (val + ir_(i + j)) & 0xFFFF
                for j in range(4 - i, 4):
                    if ls[j]:
                        val = (val + ext) & 0xFFFF
            shifted.append(val)
        out = (shifted[0] | shifted[1] << 16, shifted[2] | shifted[3] << 16)
        return is_neg, bit_shift, zbs, x, y, ls, out


def generate_shift_trace(ops) -> np.ndarray:
    """shift.rs:353-412: one row per operation (mu = 1, no deduplication);
    padding rows hold only zbs = 1."""
    c = ShiftCols
    n = max(4, 1 << max(0, (len(ops) - 1).bit_length()))
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)
    for r, op in enumerate(ops):
        is_neg, bit_shift, zbs, x, y, ls, out = op.aux()
        sa = op.shift_amount & (2**64 - 1)
        t[r, c.IN[0]:c.IN[-1] + 1] = op.halves()
        t[r, c.SHIFT_AMOUNT], t[r, c.SHIFT_B1] = sa & 0xFF, (sa >> 8) & 0xFF
        t[r, c.SHIFT_H1], t[r, c.SHIFT_HIGH] = (sa >> 16) & 0xFFFF, sa >> 32
        t[r, c.DIRECTION], t[r, c.SIGNED], t[r, c.WORD_INSTR] = int(op.direction), int(op.signed), int(op.word_instr)
        t[r, c.OUT_0], t[r, c.OUT_1] = out
        t[r, c.IS_NEGATIVE], t[r, c.BIT_SHIFT], t[r, c.ZBS] = int(is_neg), bit_shift, int(zbs)
        t[r, c.X[0]:c.X[-1] + 1] = x
        t[r, c.Y[0]:c.Y[-1] + 1] = y
        t[r, c.LIMB_SHIFT_RAW[0]:c.LIMB_SHIFT_RAW[-1] + 1] = ls[:3]
        t[r, c.MU] = 1
    t[len(ops):, c.ZBS] = 1
    return t


def shift_bus_interactions():
    """shift.rs:417-716, in its order (17): MSB16[in[3], is_negative] (signed);
    BYTE_ALU[AND, shift, 15, bit_shift] (mu - direction) and [AND, 256 - 16 zbs -
    shift, 15, bit_shift] (direction); ZERO[bit_shift, zbs] (mu); HWSL[in[i],
    bit_shift, X[i], Y[i]] and HWSL[65535 is_negative, bit_shift, X[4],
    65535 is_negative - X[4]] (1 - zbs); BYTE_ALU[AND, shift, 48 - 32 word_instr,
    48 - 48 ls0 - 32 ls1 - 16 ls2] (mu); the ALU receiver [in, shift amount,
    SHIFT + word_instr + 32 signed + 64 direction, out] (mu); ARE_BYTES[shift
    bytes 0-1] once; IS_HALFWORD of shift bits 16-31 and of the four input
    halves (mu)."""
    c = ShiftCols
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    k = lambda v: BusValue.of_linear([LinearTerm(v)])  # noqa: E731
    mu = Multiplicity("Column", (c.MU,))
    not_zbs = Multiplicity("Negated", (c.ZBS,))
    its = [BusInteraction.sender(BUS_MSB16, Multiplicity("Column", (c.SIGNED,)), [d(c.IN[3]), d(c.IS_NEGATIVE)]),
           BusInteraction.sender(BUS_BYTE_ALU, Multiplicity("Diff", (c.MU, c.DIRECTION)),
                                 [k(ALU_OP_AND), d(c.SHIFT_AMOUNT), k(15), d(c.BIT_SHIFT)]),
           BusInteraction.sender(BUS_BYTE_ALU, Multiplicity("Column", (c.DIRECTION,)), [
               k(ALU_OP_AND), BusValue.of_linear([LinearTerm(256), LinearTerm(-16, c.ZBS), LinearTerm(-1, c.SHIFT_AMOUNT)]),
               k(15), d(c.BIT_SHIFT)]),
           BusInteraction.sender(BUS_ZERO, mu, [d(c.BIT_SHIFT), d(c.ZBS)])]
    its += [BusInteraction.sender(BUS_HWSL, not_zbs, [d(c.IN[i]), d(c.BIT_SHIFT), d(c.X[i]), d(c.Y[i])])
            for i in range(4)]
    its.append(BusInteraction.sender(BUS_HWSL, not_zbs, [
        BusValue.of_linear([LinearTerm(65535, c.IS_NEGATIVE)]), d(c.BIT_SHIFT), d(c.X[4]),
        BusValue.of_linear([LinearTerm(65535, c.IS_NEGATIVE), LinearTerm(-1, c.X[4])])]))
    its.append(BusInteraction.sender(BUS_BYTE_ALU, mu, [
        k(ALU_OP_AND), d(c.SHIFT_AMOUNT), BusValue.of_linear([LinearTerm(48), LinearTerm(-32, c.WORD_INSTR)]),
        BusValue.of_linear([LinearTerm(48), LinearTerm(-48, c.LIMB_SHIFT_RAW[0]), LinearTerm(-32, c.LIMB_SHIFT_RAW[1]),
                            LinearTerm(-16, c.LIMB_SHIFT_RAW[2])])]))
    its.append(BusInteraction.receiver(BUS_ALU, mu, [
        BusValue.packed(Packing.DWORD_HL, c.IN[0]),
        BusValue.of_linear([LinearTerm(1, c.SHIFT_AMOUNT), LinearTerm(1 << 8, c.SHIFT_B1), LinearTerm(1 << 16, c.SHIFT_H1)]),
        d(c.SHIFT_HIGH),
        BusValue.of_linear([LinearTerm(ALU_OP_SHIFT), LinearTerm(1, c.WORD_INSTR), LinearTerm(32, c.SIGNED),
                            LinearTerm(64, c.DIRECTION)]),
        BusValue.packed(Packing.DWORD_WL, c.OUT_0)]))
    its.append(BusInteraction.sender(BUS_ARE_BYTES, ONE, [d(c.SHIFT_AMOUNT), d(c.SHIFT_B1)]))
    its.append(BusInteraction.sender(BUS_IS_HALFWORD, mu, [d(c.SHIFT_H1)]))
    its += [BusInteraction.sender(BUS_IS_HALFWORD, mu, [d(col)]) for col in c.IN]
    return its


def shift_table_air() -> AirProgram:
    """SHIFT as an AIR (shift.rs:823-940): 0 direction implies mu; 1-4 zbs
    pins X[i] to in[i]·(mu - direction), 5 zbs·X[4], 6-9 zbs pins Y[i] to
    in[i]·direction; 10-13 limb_shift bits; 14-15 out words equal the shifted
    halves; 16-18 direction, signed, word_instr bits; 19 zbs a bit; 20 mu a
    bit; 21 an unsigned input is not negative; then the LogUp constraints of
    its 17 interactions."""
    c = ShiftCols
    b = ir.IrBuilder()
    m = lambda col: b.main(0, col)  # noqa: E731
    one = b.const_base(1)

    def ls(i):
        return m(c.LIMB_SHIFT_RAW[i]) if i < 3 else \
            one - m(c.LIMB_SHIFT_RAW[0]) - m(c.LIMB_SHIFT_RAW[1]) - m(c.LIMB_SHIFT_RAW[2])

    def intra_left(i):
        return m(c.X[0]) if i == 0 else m(c.X[i]) + m(c.Y[i - 1])

    def intra_right(i):
        return m(c.Y[i]) + m(c.X[i + 1])

    def shifted_half(i):
        left = m(c.MU) - m(c.DIRECTION)
        right = m(c.DIRECTION)
        extension = m(c.IS_NEGATIVE) * b.const_base(65535)
        lp = None
        for j in range(i + 1):
            t = ls(j) * intra_left(i - j)
            lp = t if lp is None else lp + t
        rp = None
        for j in range(4 - i):
            t = ls(j) * intra_right(i + j)
            rp = t if rp is None else rp + t
        es = None
        for j in range(4 - i, 4):
            es = ls(j) if es is None else es + ls(j)
        right_part = rp + extension * es if es is not None else rp
        return left * lp + right * right_part

    b.emit(0, m(c.DIRECTION) * (one - m(c.MU)))
    for i in range(4):
        b.emit(1 + i, m(c.ZBS) * (m(c.X[i]) - m(c.IN[i]) * (m(c.MU) - m(c.DIRECTION))))
    b.emit(5, m(c.ZBS) * m(c.X[4]))
    for i in range(4):
        b.emit(6 + i, m(c.ZBS) * (m(c.Y[i]) - m(c.IN[i]) * m(c.DIRECTION)))
    for i in range(4):
        b.emit(10 + i, ls(i) * (one - ls(i)))
    for i in range(2):
        b.emit(14 + i, m(c.OUT_0 + i) - shifted_half(2 * i) - shifted_half(2 * i + 1) * b.const_base(1 << 16))
    for off, col in enumerate((c.DIRECTION, c.SIGNED, c.WORD_INSTR)):
        _emit_is_bit(b, 16 + off, col)
    _emit_is_bit(b, 19, c.ZBS)
    _emit_is_bit(b, 20, c.MU)
    b.emit(21, (one - m(c.SIGNED)) * m(c.IS_NEGATIVE))
    its = shift_bus_interactions()
    n_total = emit_logup_constraints(b, its, 22)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("shift", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3 if pairs else 0, 1 + absorbed, 3))


# DVRM (prover/src/tables/dvrm.rs): 64-bit DIV/DIVU/REM/REMU. n = q·d + r is
# proven through lookups: MUL[d, q] lo = n - r and hi = its sign extension,
# LT[|r|, |d|] = 1 unless d = 0, NEG templates tie abs_r/abs_d to r/d, and the
# ZERO checks pin div_by_zero and overflow (MIN / -1). Deduplicated by
# (n, d, signed) with separate quotient and remainder multiplicities.
ALU_OP_DIVREM = 8   # types.rs:570
DVRM_SIGN_FILL = 0xFFFF


class DvrmCols:
    """dvrm.rs:48-140."""
    N = (0, 1, 2, 3)
    D = (4, 5, 6, 7)
    SIGNED = 8
    Q = (9, 10, 11, 12)
    R = (13, 14, 15, 16)
    DIV_BY_ZERO, OVERFLOW = 17, 18
    ABS_R = (19, 20)
    ABS_D = (21, 22)
    N_SUB_R = (23, 24, 25, 26)
    SIGN_N_SUB_R, SIGN_N, SIGN_D, SIGN_Q, SIGN_R = 27, 28, 29, 30, 31
    MU_Q, MU_R = 32, 33
    NUM_COLUMNS = 34


@dataclass(frozen=True)
class DvrmOperation:
    """dvrm.rs:157-272: quotient and remainder as RISC-V defines them (division
    by zero: q = 2^64 - 1, r = n; signed MIN / -1: q = MIN, r = 0; signed
    division truncates toward zero)."""
    n: int
    d: int
    signed: bool

    def is_div_by_zero(self) -> bool:
        return self.d == 0

    def is_overflow(self) -> bool:
        return self.signed and self.n == 1 << 63 and self.d == 2**64 - 1

    @staticmethod
    def _s(v):
        return v - 2**64 if v >> 63 else v

    def quotient(self) -> int:
        if self.is_div_by_zero():
            return 2**64 - 1
        if self.is_overflow():
            return self.n
        if self.signed:
            a, b = self._s(self.n), self._s(self.d)
            q = abs(a) // abs(b)
            return (-q if (a < 0) != (b < 0) else q) & (2**64 - 1)
        return self.n // self.d

    def remainder(self) -> int:
        if self.is_div_by_zero():
            return self.n
        if self.is_overflow():
            return 0
        if self.signed:
            a, b = self._s(self.n), self._s(self.d)
            r = abs(a) % abs(b)
            return (-r if a < 0 else r) & (2**64 - 1)
        return self.n % self.d

    def sign_n(self) -> bool:
        return self.signed and self.n >> 63 == 1

    def sign_d(self) -> bool:
        return self.signed and self.d >> 63 == 1

    def sign_q(self) -> bool:
        return self.signed and not self.is_overflow()

    def sign_r(self) -> bool:
        return self.signed and self.remainder() >> 63 == 1

    def abs_r(self) -> int:
        r = self.remainder()
        return (2**64 - r) & (2**64 - 1) if self.sign_r() else r

    def abs_d(self) -> int:
        return (2**64 - self.d) & (2**64 - 1) if self.sign_d() else self.d

    def n_sub_r(self) -> int:
        return (self.n - self.remainder()) & (2**64 - 1)

    def sign_n_sub_r(self) -> bool:
        return self.signed and self.n_sub_r() >> 63 == 1


def dedup_dvrm_rows(ops):
    """dvrm.rs:287-320: (operation, wants_remainder) pairs merged per (n, d,
    signed) into (mu_q, mu_r) counts, split at MU_MAX; rows follow first
    occurrence (the reference's map order is arbitrary)."""
    counts = {}
    for op, wants_r in ops:
        q, r = counts.get(op, (0, 0))
        counts[op] = (q, r + 1) if wants_r else (q + 1, r)
    rows = []
    for op, (q, r) in counts.items():
        while q > 0 or r > 0:
            rq, rr = min(q, MU_MAX), min(r, MU_MAX)
            q, r = q - rq, r - rr
            rows.append((op, rq, rr))
    return rows


def generate_dvrm_trace(ops) -> np.ndarray:
    """dvrm.rs:322-385: one row per deduplicated operation, zero rows to a power
    of two (at least 4)."""
    c = DvrmCols
    rows = dedup_dvrm_rows(ops)
    n = max(4, 1 << max(0, (len(rows) - 1).bit_length()))
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)
    for i, (op, mu_q, mu_r) in enumerate(rows):
        for cols, v in ((c.N, op.n), (c.D, op.d), (c.Q, op.quotient()), (c.R, op.remainder()),
                        (c.N_SUB_R, op.n_sub_r())):
            for k, col in enumerate(cols):
                t[i, col] = (v >> (16 * k)) & 0xFFFF
        for cols, v in ((c.ABS_R, op.abs_r()), (c.ABS_D, op.abs_d())):
            t[i, cols[0]], t[i, cols[1]] = v & 0xFFFFFFFF, v >> 32
        t[i, c.SIGNED] = int(op.signed)
        t[i, c.DIV_BY_ZERO], t[i, c.OVERFLOW] = int(op.is_div_by_zero()), int(op.is_overflow())
        t[i, c.SIGN_N_SUB_R], t[i, c.SIGN_N] = int(op.sign_n_sub_r()), int(op.sign_n())
        t[i, c.SIGN_D], t[i, c.SIGN_Q], t[i, c.SIGN_R] = int(op.sign_d()), int(op.sign_q()), int(op.sign_r())
        t[i, c.MU_Q], t[i, c.MU_R] = mu_q, mu_r
    return t


def _neg_carry_zero(x, absx, sign_col):
    """The two ZERO sends of the NEG template x + abs_x = 2^64 (dvrm.rs:673-
    865): ZERO[x0 + x1; 1 - carry0] and ZERO[x0 + .. + x3; 1 - carry1] with
    carry0 = 2^-32 (x.lo + abs.lo), carry1 = 2^-32 (x.hi + abs.hi + carry0)."""
    neg = lambda k: P - _inv2(k)  # noqa: E731
    mu = Multiplicity("Column", (sign_col,))
    c0 = [LinearTerm(1), LinearTerm(neg(32), absx[0], True), LinearTerm(neg(32), x[0], True),
          LinearTerm(neg(16), x[1], True)]
    c1 = [LinearTerm(1), LinearTerm(neg(32), absx[1], True), LinearTerm(neg(32), x[2], True),
          LinearTerm(neg(16), x[3], True), LinearTerm(neg(64), absx[0], True), LinearTerm(neg(64), x[0], True),
          LinearTerm(neg(48), x[1], True)]
    return [BusInteraction.sender(BUS_ZERO, mu, [BusValue.of_linear([LinearTerm(1, x[0]), LinearTerm(1, x[1])]),
                                                 BusValue.of_linear(c0)]),
            BusInteraction.sender(BUS_ZERO, mu, [BusValue.of_linear([LinearTerm(1, col) for col in x]),
                                                 BusValue.of_linear(c1)])]


def dvrm_bus_interactions():
    """dvrm.rs:393-1023, in its order (35): ARE_BYTES[mu_q, mu_r] once;
    IS_HALFWORD of n, d, r, n - r and q halves (mu_q + mu_r); MSB16 of n[3],
    r[3], d[3] (signed); ALU LT[abs_r, abs_d] = 1 - div_by_zero; ALU MUL lo
    [d, q] = n - r and hi = its sign extension (flags MUL + 32 signed + 64
    sign_q, + 128 for hi); the NEG ZERO pairs of r (sign_r) and d (sign_d);
    ZERO[overflow sum; overflow] and ZERO[d halves; div_by_zero]; the ALU
    receivers of q (mu_q) and r (mu_r, flags + 128)."""
    c = DvrmCols
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    hl = lambda col: BusValue.packed(Packing.DWORD_HL, col)  # noqa: E731
    k = lambda v: BusValue.of_linear([LinearTerm(v)])  # noqa: E731
    row = Multiplicity("Sum", (c.MU_Q, c.MU_R))
    signed = Multiplicity("Column", (c.SIGNED,))
    its = [BusInteraction.sender(BUS_ARE_BYTES, ONE, [d(c.MU_Q), d(c.MU_R)])]
    its += [BusInteraction.sender(BUS_IS_HALFWORD, row, [d(col)]) for col in c.N + c.D + c.R + c.N_SUB_R + c.Q]
    its += [BusInteraction.sender(BUS_MSB16, signed, [d(top), d(sign)])
            for top, sign in ((c.N[3], c.SIGN_N), (c.R[3], c.SIGN_R), (c.D[3], c.SIGN_D))]
    its.append(BusInteraction.sender(BUS_ALU, row, [
        BusValue.packed(Packing.DWORD_WL, c.ABS_R[0]), BusValue.packed(Packing.DWORD_WL, c.ABS_D[0]), k(ALU_OP_LT),
        BusValue.of_linear([LinearTerm(1), LinearTerm(-1, c.DIV_BY_ZERO)]), k(0)]))

    def mul_flags(hi):
        return BusValue.of_linear([LinearTerm(ALU_OP_MUL + hi), LinearTerm(32, c.SIGNED), LinearTerm(64, c.SIGN_Q)])

    ext = BusValue.of_linear([LinearTerm(DVRM_SIGN_FILL + (DVRM_SIGN_FILL << 16), c.SIGN_N_SUB_R)])
    its.append(BusInteraction.sender(BUS_ALU, row, [hl(c.D[0]), hl(c.Q[0]), mul_flags(0), hl(c.N_SUB_R[0])]))
    its.append(BusInteraction.sender(BUS_ALU, row, [hl(c.D[0]), hl(c.Q[0]), mul_flags(128), ext, ext]))
    its += _neg_carry_zero(c.R, c.ABS_R, c.SIGN_R)
    its += _neg_carry_zero(c.D, c.ABS_D, c.SIGN_D)
    overflow_sum = [LinearTerm(1, col) for col in c.N] + [LinearTerm(-32769, c.SIGN_N), LinearTerm(1 + 4 * 65535)]
    overflow_sum += [LinearTerm(-1, col) for col in c.D]
    its.append(BusInteraction.sender(BUS_ZERO, row, [BusValue.of_linear(overflow_sum), d(c.OVERFLOW)]))
    its.append(BusInteraction.sender(BUS_ZERO, row, [BusValue.of_linear([LinearTerm(1, col) for col in c.D]),
                                                     d(c.DIV_BY_ZERO)]))
    for mu, extra, out in ((c.MU_Q, 0, c.Q[0]), (c.MU_R, 128, c.R[0])):
        its.append(BusInteraction.receiver(BUS_ALU, Multiplicity("Column", (mu,)), [
            hl(c.N[0]), hl(c.D[0]), BusValue.of_linear([LinearTerm(ALU_OP_DIVREM + extra), LinearTerm(32, c.SIGNED)]),
            hl(out)]))
    return its


def dvrm_table_air() -> AirProgram:
    """DVRM as an AIR (dvrm.rs:1116-1209): 0 signed a bit; 1 a nonzero remainder
    has the numerator's sign; 2-3 abs_r = r and 4-5 abs_d = d when not
    negative; 6 sign_q = signed (1 - overflow); 7-10 the carries of the
    sign-extended n - r + r = n are bits; 11 sign_n_sub_r a bit; 12-14 no sign
    bits on unsigned rows; 15-18 division by zero forces q halves to 65535;
    then the LogUp constraints of its 35 interactions."""
    c = DvrmCols
    b = ir.IrBuilder()
    m = lambda col: b.main(0, col)  # noqa: E731
    one, two16 = b.const_base(1), b.const_base(1 << 16)
    inv32 = b.const_base(INV_SHIFT_32)

    def ext_word(halves, sign, i):
        if i < 2:
            return m(halves[2 * i]) + m(halves[2 * i + 1]) * two16
        return m(sign) * b.const_base(DVRM_SIGN_FILL + (DVRM_SIGN_FILL << 16))

    def carry(i):
        s = ext_word(c.N_SUB_R, c.SIGN_N_SUB_R, i) + ext_word(c.R, c.SIGN_R, i)
        if i > 0:
            s = s + carry(i - 1)
        return (s - ext_word(c.N, c.SIGN_N, i)) * inv32

    _emit_is_bit(b, 0, c.SIGNED)
    r_sum = m(c.R[0]) + m(c.R[1]) + m(c.R[2]) + m(c.R[3])
    b.emit(1, r_sum * (m(c.SIGN_R) - m(c.SIGN_N)))
    for off, (absc, lo, hi) in enumerate(((c.ABS_R[0], c.R[0], c.R[1]), (c.ABS_R[1], c.R[2], c.R[3]))):
        b.emit(2 + off, (one - m(c.SIGN_R)) * (m(absc) - (m(lo) + m(hi) * two16)))
    for off, (absc, lo, hi) in enumerate(((c.ABS_D[0], c.D[0], c.D[1]), (c.ABS_D[1], c.D[2], c.D[3]))):
        b.emit(4 + off, (one - m(c.SIGN_D)) * (m(absc) - (m(lo) + m(hi) * two16)))
    b.emit(6, m(c.SIGNED) * (one - m(c.OVERFLOW)) - m(c.SIGN_Q))
    for i in range(4):
        ci = carry(i)
        b.emit(7 + i, ci * (one - ci))
    _emit_is_bit(b, 11, c.SIGN_N_SUB_R)
    for off, col in enumerate((c.SIGN_N, c.SIGN_R, c.SIGN_D)):
        b.emit(12 + off, (one - m(c.SIGNED)) * m(col))
    fill = b.const_base(DVRM_SIGN_FILL)
    for i, col in enumerate(c.Q):
        b.emit(15 + i, m(c.DIV_BY_ZERO) * (m(col) - fill))
    its = dvrm_bus_interactions()
    n_total = emit_logup_constraints(b, its, 19)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("dvrm", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3 if pairs else 0, 1 + absorbed, 2))


# KECCAK_RND (prover/src/tables/keccak_rnd.rs): the Keccak-f[1600] round chip,
# one row per round (24 per permutation), every intermediate state as bytes.
# XOR/AND go to BYTE_ALU lookups; the theta rotate-by-1 and the rho rotations
# are halfword shifts pinned by mu-gated linear identities over (left, right)
# byte pairs, the carries of theta's single-bit shifts as bit columns.
KECCAK_RHO = ((0, 36, 3, 41, 18), (1, 44, 10, 45, 2), (62, 6, 43, 15, 61), (28, 55, 25, 21, 56),
              (27, 20, 39, 8, 14))  # [x][y], executor/src/vm/instruction/execution.rs:1029-1035


class KeccakRndCols:
    """keccak_rnd.rs:28-90: timestamp (2), round, start (200), Cxz chain (160),
    Cxz_left (40), Cxz_right carries (20), Dxz (40), theta (200), rot_left and
    rot_right (200 each), chi_ands (200), chi (200), rc (8), iota (8), mu."""
    TIMESTAMP_0, TIMESTAMP_1, ROUND = 0, 1, 2
    START, CXZ, CXZ_LEFT, CXZ_RIGHT, DXZ = 3, 203, 363, 403, 423
    THETA, ROT_LEFT, ROT_RIGHT, CHI_ANDS, CHI = 463, 663, 863, 1063, 1263
    RC, IOTA, MU = 1463, 1471, 1479
    NUM_COLUMNS = 1480

    @staticmethod
    def lane(base, x, y, byte):
        """Byte `byte` of lane (x, y) of a 200-byte state group."""
        return base + (x + 5 * y) * 8 + byte

    @staticmethod
    def cxz(x, stage, byte):
        return KeccakRndCols.CXZ + (x * 4 + stage) * 8 + byte

    @staticmethod
    def cxz_left(x, byte):
        return KeccakRndCols.CXZ_LEFT + x * 8 + byte

    @staticmethod
    def cxz_right(x, hw):
        return KeccakRndCols.CXZ_RIGHT + x * 4 + hw

    @staticmethod
    def dxz(x, byte):
        return KeccakRndCols.DXZ + x * 8 + byte

    @staticmethod
    def pi_src(x, y, z):
        """keccak_rnd.rs:165-179: the (rot_left, rot_right) cells whose sum is
        byte z of pi lane (x, y) = rho lane ((x + 3y) mod 5, x): a rotation by
        16·rbc + rnc takes byte z from the left part of byte z - 2 rbc and the
        right part of byte z - 2 rbc - 2."""
        c = KeccakRndCols
        sx, sy = (x + 3 * y) % 5, x
        rbc = KECCAK_RHO[sx][sy] // 16
        return (c.lane(c.ROT_LEFT, sx, sy, (z - 2 * rbc) % 8), c.lane(c.ROT_RIGHT, sx, sy, (z - 2 * rbc - 2) % 8))


@dataclass(frozen=True)
class KeccakRoundOperation:
    """keccak_rnd.rs:233-238: one permutation (24 rows); lanes indexed x + 5y."""
    timestamp: int
    input: tuple
    output: tuple


def _hwsl(h: int, shift: int):
    """(h << shift) mod 2^16 and h >> (16 - shift), (h, 0) for shift 0."""
    return (h, 0) if shift == 0 else ((h << shift) & 0xFFFF, h >> (16 - shift))


def _rotl64(v: int, r: int) -> int:
    r %= 64
    return ((v << r) | (v >> (64 - r))) & (2**64 - 1) if r else v


def generate_keccak_rnd_trace(ops) -> np.ndarray:
    """keccak_rnd.rs:262-448: per operation 24 rows, each round's values at
    byte granularity; zero rows to a power of two (at least 4)."""
    c = KeccakRndCols
    n = max(4, 1 << max(0, (24 * len(ops) - 1).bit_length()))
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)
    by = lambda v: [(v >> (8 * i)) & 0xFF for i in range(8)]  # noqa: E731
    lane_of = lambda bs: sum(b << (8 * i) for i, b in enumerate(bs))  # noqa: E731
    for k, op in enumerate(ops):
        state = list(op.input)
        for rnd in range(24):
            row = t[24 * k + rnd]
            row[c.TIMESTAMP_0], row[c.TIMESTAMP_1], row[c.ROUND] = op.timestamp & 0xFFFFFFFF, op.timestamp >> 32, rnd
            for x in range(5):
                for y in range(5):
                    row[c.lane(c.START, x, y, 0):c.lane(c.START, x, y, 8)] = by(state[x + 5 * y])
            # theta: parities as a chain of XORs, then C rotated by one bit
            cpar = []
            for x in range(5):
                acc = by(state[x])
                for stage in range(4):
                    acc = [a ^ b for a, b in zip(acc, by(state[x + 5 * (stage + 1)]))]
                    row[c.cxz(x, stage, 0):c.cxz(x, stage, 8)] = acc
                cpar.append(acc)
            rot_c = []
            for x in range(5):
                left, carries = [0] * 8, [0] * 4
                for hw in range(4):
                    sh, carry = _hwsl(cpar[x][2 * hw] | cpar[x][2 * hw + 1] << 8, 1)
                    left[2 * hw], left[2 * hw + 1], carries[hw] = sh & 0xFF, sh >> 8, carry
                row[c.cxz_left(x, 0):c.cxz_left(x, 8)] = left
                row[c.cxz_right(x, 0):c.cxz_right(x, 4)] = carries
                # the carry of halfword hw - 1 lands in the low byte of halfword hw
                rot_c.append([left[b] + (carries[(b // 2 + 3) % 4] if b % 2 == 0 else 0) for b in range(8)])
            dd = []
            for x in range(5):
                d = [a ^ b for a, b in zip(cpar[(x + 4) % 5], rot_c[(x + 1) % 5])]
                row[c.dxz(x, 0):c.dxz(x, 8)] = d
                dd.append(lane_of(d))
            theta = [state[i] ^ dd[i % 5] for i in range(25)]
            for x in range(5):
                for y in range(5):
                    row[c.lane(c.THETA, x, y, 0):c.lane(c.THETA, x, y, 8)] = by(theta[x + 5 * y])
                    left, right = [0] * 8, [0] * 8
                    for hw in range(4):
                        sh, carry = _hwsl((theta[x + 5 * y] >> (16 * hw)) & 0xFFFF, KECCAK_RHO[x][y] % 16)
                        left[2 * hw], left[2 * hw + 1] = sh & 0xFF, sh >> 8
                        right[2 * hw], right[2 * hw + 1] = carry & 0xFF, carry >> 8
                    row[c.lane(c.ROT_LEFT, x, y, 0):c.lane(c.ROT_LEFT, x, y, 8)] = left
                    row[c.lane(c.ROT_RIGHT, x, y, 0):c.lane(c.ROT_RIGHT, x, y, 8)] = right
            pi = [0] * 25
            for x in range(5):
                for y in range(5):
                    pi[y + 5 * ((2 * x + 3 * y) % 5)] = _rotl64(theta[x + 5 * y], KECCAK_RHO[x][y])
            chi = [0] * 25
            for x in range(5):
                for y in range(5):
                    a = ~pi[(x + 1) % 5 + 5 * y] & pi[(x + 2) % 5 + 5 * y] & (2**64 - 1)
                    chi[x + 5 * y] = pi[x + 5 * y] ^ a
                    row[c.lane(c.CHI_ANDS, x, y, 0):c.lane(c.CHI_ANDS, x, y, 8)] = by(a)
                    row[c.lane(c.CHI, x, y, 0):c.lane(c.CHI, x, y, 8)] = by(chi[x + 5 * y])
            rc = KECCAK_ROUND_CONSTANTS[rnd]
            row[c.RC:c.RC + 8] = by(rc)
            chi[0] ^= rc
            row[c.IOTA:c.IOTA + 8] = by(chi[0])
            row[c.MU] = 1
            state = chi
    return t


def keccak_rnd_bus_interactions():
    """keccak_rnd.rs:461-860, in its order (1031): the Keccak receiver [ts,
    round, start bytes] and sender [ts, round + 1, round output bytes] (x-major,
    iota in lane (0, 0)), the KeccakRc lookup [round, rc bytes] (all mu); the
    BYTE_ALU XORs of the parity chain (160), ARE_BYTES of the Cxz_left pairs
    (20, multiplicity one), the XORs of D (40) — the rotated parity byte as
    Cxz_left plus, on even bytes, a carry bit — and of theta (200), ARE_BYTES of
    the (rot_left, rot_right) pairs (200, one), chi's ANDs of 255 - pi and pi
    (200) and XORs (200) with pi as rot_left + rot_right cells, iota's XORs
    (8)."""
    c = KeccakRndCols
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    k = lambda v: BusValue.of_linear([LinearTerm(v)])  # noqa: E731
    lin = lambda *cols: BusValue.of_linear([LinearTerm(1, col) for col in cols])  # noqa: E731
    mu = Multiplicity("Column", (c.MU,))
    xyz = [(x, y, b) for x in range(5) for y in range(5) for b in range(8)]
    ts = [d(c.TIMESTAMP_0), d(c.TIMESTAMP_1)]
    out_col = lambda x, y, b: c.IOTA + b if (x, y) == (0, 0) else c.lane(c.CHI, x, y, b)  # noqa: E731
    its = [BusInteraction.receiver(BUS_KECCAK, mu, ts + [d(c.ROUND)] + [d(c.lane(c.START, *p)) for p in xyz]),
           BusInteraction.sender(BUS_KECCAK, mu, ts + [BusValue.of_linear([LinearTerm(1, c.ROUND), LinearTerm(1)])]
                                 + [d(out_col(*p)) for p in xyz]),
           BusInteraction.sender(BUS_KECCAK_RC, mu, [d(c.ROUND)] + [d(c.RC + b) for b in range(8)])]
    xor = lambda a, b_, out: BusInteraction.sender(BUS_BYTE_ALU, mu, [k(ALU_OP_XOR), a, b_, d(out)])  # noqa: E731
    its += [xor(d(c.lane(c.START, x, 0, b)), d(c.lane(c.START, x, 1, b)), c.cxz(x, 0, b))
            for x in range(5) for b in range(8)]
    its += [xor(d(c.cxz(x, s - 1, b)), d(c.lane(c.START, x, s + 1, b)), c.cxz(x, s, b))
            for x in range(5) for s in range(1, 4) for b in range(8)]
    its += [BusInteraction.sender(BUS_ARE_BYTES, ONE, [d(c.cxz_left(x, 2 * i)), d(c.cxz_left(x, 2 * i + 1))])
            for x in range(5) for i in range(4)]
    for x in range(5):
        for b in range(8):
            xr = (x + 1) % 5
            rot = lin(c.cxz_left(xr, b), c.cxz_right(xr, (b // 2 + 3) % 4)) if b % 2 == 0 else lin(c.cxz_left(xr, b))
            its.append(xor(d(c.cxz((x + 4) % 5, 3, b)), rot, c.dxz(x, b)))
    its += [xor(d(c.lane(c.START, x, y, b)), d(c.dxz(x, b)), c.lane(c.THETA, x, y, b)) for x, y, b in xyz]
    its += [BusInteraction.sender(BUS_ARE_BYTES, ONE, [d(c.lane(c.ROT_LEFT, x, y, b)), d(c.lane(c.ROT_RIGHT, x, y, b))])
            for x, y, b in xyz]
    for x, y, b in xyz:
        p1, p2 = c.pi_src((x + 1) % 5, y, b), c.pi_src((x + 2) % 5, y, b)
        its.append(BusInteraction.sender(BUS_BYTE_ALU, mu, [
            k(ALU_OP_AND), BusValue.of_linear([LinearTerm(255), LinearTerm(-1, p1[0]), LinearTerm(-1, p1[1])]),
            lin(*p2), d(c.lane(c.CHI_ANDS, x, y, b))]))
    its += [xor(lin(*c.pi_src(x, y, b)), d(c.lane(c.CHI_ANDS, x, y, b)), c.lane(c.CHI, x, y, b)) for x, y, b in xyz]
    its += [xor(d(c.lane(c.CHI, 0, 0, b)), d(c.RC + b), c.IOTA + b) for b in range(8)]
    return its


def keccak_rnd_table_air() -> AirProgram:
    """KECCAK_RND as an AIR (keccak_rnd.rs:862-951, 141 constraints of degree
    2): 0-19 the theta carries are bits; 20-39 mu·(C_hw·2 - carry·2^16 -
    left_hw) per parity halfword; 40-139 mu·(theta_hw·2^rnc - right_hw·2^16 -
    left_hw) per lane and halfword (rnc = rho mod 16); 140 mu a bit; then the
    LogUp constraints of its 1031 interactions."""
    c = KeccakRndCols
    b = ir.IrBuilder()
    m = lambda col: b.main(0, col)  # noqa: E731
    two16, c256, one = b.const_base(1 << 16), b.const_base(256), b.const_base(1)
    half = lambda lo: m(lo) + m(lo + 1) * c256  # noqa: E731
    idx = 0
    for x in range(5):
        for hw in range(4):
            _emit_is_bit(b, idx, c.cxz_right(x, hw))
            idx += 1
    for x in range(5):
        for hw in range(4):
            ident = half(c.cxz(x, 3, 2 * hw)) * b.const_base(2) - m(c.cxz_right(x, hw)) * two16 - half(c.cxz_left(x, 2 * hw))
            b.emit(idx, m(c.MU) * ident)
            idx += 1
    for x in range(5):
        for y in range(5):
            pw = b.const_base(1 << (KECCAK_RHO[x][y] % 16))
            for hw in range(4):
                ident = (half(c.lane(c.THETA, x, y, 2 * hw)) * pw - half(c.lane(c.ROT_RIGHT, x, y, 2 * hw)) * two16
                         - half(c.lane(c.ROT_LEFT, x, y, 2 * hw)))
                b.emit(idx, m(c.MU) * ident)
                idx += 1
    b.emit(idx, m(c.MU) * (one - m(c.MU)))
    idx += 1
    its = keccak_rnd_bus_interactions()
    n_total = emit_logup_constraints(b, its, idx)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("keccak_rnd", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3 if pairs else 0, 1 + absorbed, 2))


# CPU32 (prover/src/tables/cpu32.rs): the 32-bit word instructions (*W) the
# CPU delegates; decodes through DECODE, reads rs1/rs2, extends the low words
# (sign or zero by the signed bit of alu_flags), runs the ALU or the ADD/SUB
# fast path and writes rd with the sign-extended low word of the result.
BUS_DECODE, BUS_CPU32 = 18, 27      # types.rs:309, :343
ALU_OP_SHIFTW = 6                   # types.rs:568
# packed_decode bit positions (types.rs:586-604)
PD_READ_REG1, PD_READ_REG2, PD_WRITE_REG, PD_WORD_INSTR, PD_ALU, PD_ADD, PD_SUB = 0, 1, 2, 3, 4, 5, 6
PD_MEMORY, PD_BRANCH, PD_ECALL, PD_RS1, PD_RS2, PD_RD = 7, 8, 9, 10, 18, 26
PD_HALF_INSTRUCTION_LENGTH, PD_ALU_FLAGS, PD_MEM_FLAGS = 34, 42, 50
ALU_FLAGS_SIGNED = 5
CPU32_HI_FILL = 0xFFFFFFFF

# (alu op, signed, signed2 | invert, muldiv, add, sub) of each RISC-V arithmetic
# op (types.rs:870-895); "shift" stands for SHIFT, or SHIFTW on word forms
ARITH_OP_FLAGS = {
    "add": (0, False, False, False, True, False), "sub": (0, False, False, False, False, True),
    "and": (ALU_OP_AND, False, False, False, False, False), "or": (ALU_OP_OR, False, False, False, False, False),
    "xor": (ALU_OP_XOR, False, False, False, False, False), "sll": ("shift", False, False, False, False, False),
    "srl": ("shift", False, True, False, False, False), "sra": ("shift", True, True, False, False, False),
    "slt": (ALU_OP_LT, True, False, False, False, False), "sltu": (ALU_OP_LT, False, False, False, False, False),
    "mul": (ALU_OP_MUL, True, True, False, False, False), "mulh": (ALU_OP_MUL, True, True, True, False, False),
    "mulhsu": (ALU_OP_MUL, True, False, True, False, False), "mulhu": (ALU_OP_MUL, False, False, True, False, False),
    "div": (ALU_OP_DIVREM, True, False, False, False, False), "divu": (ALU_OP_DIVREM, False, False, False, False, False),
    "rem": (ALU_OP_DIVREM, True, False, True, False, False), "remu": (ALU_OP_DIVREM, False, False, True, False, False),
}


def arith_decode(op: str, word_instr: bool):
    """alu_flags byte and the alu/add/sub flags of an arithmetic op
    (types.rs:864-900, build_alu_flags: op + 32 signed + 64 signed2|invert +
    128 muldiv)."""
    alu, signed, s2, muldiv, add, sub = ARITH_OP_FLAGS[op]
    if alu == "shift":
        alu = ALU_OP_SHIFTW if word_instr else ALU_OP_SHIFT
    return alu + 32 * signed + 64 * s2 + 128 * muldiv, not (add or sub), add, sub


class Cpu32Cols:
    """cpu32.rs:37-98."""
    TIMESTAMP_0, TIMESTAMP_1, PC_0, PC_1 = 0, 1, 2, 3
    RS1, READ_REGISTER1, RV1 = 4, 5, (6, 7, 8)
    RV1_SIGN, ARG1_0, ARG1_1 = 9, 10, 11
    RS2, READ_REGISTER2, RV2 = 12, 13, (14, 15, 16)
    RV2_SIGN, IMM_0, IMM_1, ARG2_0, ARG2_1 = 17, 18, 19, 20, 21
    RES = (22, 23, 24, 25)
    RES_SIGN, RD, WRITE_REGISTER, RVD_0, RVD_1 = 26, 27, 28, 29, 30
    ALU, ALU_FLAGS, ADD, SUB, HALF_INSTRUCTION_LENGTH, SIGNED, MU = 31, 32, 33, 34, 35, 36, 37
    NUM_COLUMNS = 38


@dataclass(frozen=True)
class Cpu32Operation:
    """cpu32.rs:108-190: res is the raw 64-bit ALU (or ADD/SUB) result."""
    timestamp: int
    pc: int
    rs1: int
    read_register1: bool
    rv1: int
    rs2: int
    read_register2: bool
    rv2: int
    imm: int
    res: int
    rd: int
    write_register: bool
    alu: bool
    alu_flags: int
    add: bool
    sub: bool
    half_instruction_length: int

    def signed(self) -> bool:
        return bool(self.alu_flags >> ALU_FLAGS_SIGNED & 1)

    def aux(self):
        """cpu32.rs:150-188: rv1/rv2 signs (gated by signed), arg1 = ext(rv1
        low word), arg2 = ext(rv2 low word) + imm per word, res sign, rvd =
        the sign-extended low word of res."""
        s = self.signed()
        rv1_sign = s and (self.rv1 >> 31) & 1 == 1
        rv2_sign = s and (self.rv2 >> 31) & 1 == 1
        res_sign = (self.res >> 31) & 1 == 1
        arg1 = (self.rv1 & 0xFFFFFFFF) | (CPU32_HI_FILL << 32 if rv1_sign else 0)
        lo = ((self.rv2 & 0xFFFFFFFF) + (self.imm & 0xFFFFFFFF)) & 0xFFFFFFFF
        hi = ((CPU32_HI_FILL if rv2_sign else 0) + (self.imm >> 32)) & 0xFFFFFFFF
        rvd = (self.res & 0xFFFFFFFF) | (CPU32_HI_FILL << 32 if res_sign else 0)
        return s, rv1_sign, arg1, rv2_sign, lo | hi << 32, res_sign, rvd


def generate_cpu32_trace(ops) -> np.ndarray:
    """cpu32.rs:192-247: one row per operation (mu = 1), zero rows to a power
    of two (at least 4)."""
    c = Cpu32Cols
    n = max(4, 1 << max(0, (len(ops) - 1).bit_length()))
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)
    whh = lambda v: (v & 0xFFFF, (v >> 16) & 0xFFFF, v >> 32)  # noqa: E731
    for i, op in enumerate(ops):
        signed, rv1_sign, arg1, rv2_sign, arg2, res_sign, rvd = op.aux()
        r = t[i]
        r[c.TIMESTAMP_0], r[c.TIMESTAMP_1] = op.timestamp & 0xFFFFFFFF, op.timestamp >> 32
        r[c.PC_0], r[c.PC_1] = op.pc & 0xFFFFFFFF, op.pc >> 32
        r[c.RS1], r[c.READ_REGISTER1], r[c.RV1_SIGN] = op.rs1, int(op.read_register1), int(rv1_sign)
        r[c.RV1[0]:c.RV1[2] + 1] = whh(op.rv1)
        r[c.ARG1_0], r[c.ARG1_1] = arg1 & 0xFFFFFFFF, arg1 >> 32
        r[c.RS2], r[c.READ_REGISTER2], r[c.RV2_SIGN] = op.rs2, int(op.read_register2), int(rv2_sign)
        r[c.RV2[0]:c.RV2[2] + 1] = whh(op.rv2)
        r[c.IMM_0], r[c.IMM_1] = op.imm & 0xFFFFFFFF, op.imm >> 32
        r[c.ARG2_0], r[c.ARG2_1] = arg2 & 0xFFFFFFFF, arg2 >> 32
        r[c.RES[0]:c.RES[3] + 1] = [(op.res >> (16 * k)) & 0xFFFF for k in range(4)]
        r[c.RES_SIGN], r[c.RD], r[c.WRITE_REGISTER] = int(res_sign), op.rd, int(op.write_register)
        r[c.RVD_0], r[c.RVD_1] = rvd & 0xFFFFFFFF, rvd >> 32
        r[c.ALU], r[c.ALU_FLAGS], r[c.ADD], r[c.SUB] = int(op.alu), op.alu_flags, int(op.add), int(op.sub)
        r[c.HALF_INSTRUCTION_LENGTH], r[c.SIGNED], r[c.MU] = op.half_instruction_length, int(signed), 1
    return t


def _register_dword(lo0, lo1, hi):
    """A register value as the 8-element MEMW value: [lo0 + 2^16 lo1, hi, 0 x 6]
    (cpu32.rs:255-274)."""
    return [BusValue.of_linear([LinearTerm(1, lo0), LinearTerm(1 << 16, lo1)]), BusValue.packed(Packing.DIRECT, hi)] \
        + [BusValue.of_linear([LinearTerm(0)])] * 6


def _cpu32_ts(offset):
    return [BusValue.of_linear([LinearTerm(1, Cpu32Cols.TIMESTAMP_0), LinearTerm(offset)]),
            BusValue.packed(Packing.DIRECT, Cpu32Cols.TIMESTAMP_1)]


def cpu32_bus_interactions():
    """cpu32.rs:359-560, in its order (23): DECODE[pc, imm, packed_decode] (mu;
    word_instr the constant bit); ARE_BYTES[x, 0] of half_instruction_length,
    alu_flags, rs1, rs2, rd (mu); IS_HALFWORD of the rv1/rv2 low halves and the
    four res halves (mu); the MEMW register reads of rs1 at ts and rs2 at ts + 1
    (24 elements, with read_register1/2) and the write of rd at ts + 2 (16, with
    write_register); ALU[arg1, arg2, alu_flags, res] (alu); BYTE_ALU[AND, 32,
    alu_flags, 32 signed] (mu); MSB16 of rv1[1], rv2[1] (signed) and res[1]
    (mu); the CPU32 receiver [ts, pc, half_instruction_length] (mu)."""
    c = Cpu32Cols
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    k = lambda v: BusValue.of_linear([LinearTerm(v)])  # noqa: E731
    col = lambda cc: Multiplicity("Column", (cc,))  # noqa: E731
    mu = col(c.MU)
    packed = [LinearTerm(1 << PD_READ_REG1, c.READ_REGISTER1), LinearTerm(1 << PD_READ_REG2, c.READ_REGISTER2),
              LinearTerm(1 << PD_WRITE_REG, c.WRITE_REGISTER), LinearTerm(1 << PD_WORD_INSTR),
              LinearTerm(1 << PD_ALU, c.ALU), LinearTerm(1 << PD_ADD, c.ADD), LinearTerm(1 << PD_SUB, c.SUB),
              LinearTerm(1 << PD_RS1, c.RS1), LinearTerm(1 << PD_RS2, c.RS2), LinearTerm(1 << PD_RD, c.RD),
              LinearTerm(1 << PD_HALF_INSTRUCTION_LENGTH, c.HALF_INSTRUCTION_LENGTH),
              LinearTerm(1 << PD_ALU_FLAGS, c.ALU_FLAGS)]
    its = [BusInteraction.sender(BUS_DECODE, mu, [BusValue.packed(Packing.DWORD_WL, c.PC_0),
                                                  BusValue.packed(Packing.DWORD_WL, c.IMM_0),
                                                  BusValue.of_linear(packed)])]
    its += [BusInteraction.sender(BUS_ARE_BYTES, mu, [d(cc), k(0)])
            for cc in (c.HALF_INSTRUCTION_LENGTH, c.ALU_FLAGS, c.RS1, c.RS2, c.RD)]
    its += [BusInteraction.sender(BUS_IS_HALFWORD, mu, [d(cc)])
            for cc in (c.RV1[0], c.RV1[1], c.RV2[0], c.RV2[1]) + c.RES]
    for rs, rv, off, flag in ((c.RS1, c.RV1, 0, c.READ_REGISTER1), (c.RS2, c.RV2, 1, c.READ_REGISTER2)):
        its.append(BusInteraction.sender(BUS_MEMW, col(flag), _register_dword(*rv) + [
            k(1), BusValue.of_linear([LinearTerm(2, rs)]), k(0)] + _register_dword(*rv) + _cpu32_ts(off)
            + [k(1), k(0), k(0)]))
    its.append(BusInteraction.sender(BUS_MEMW, col(c.WRITE_REGISTER), [
        k(1), BusValue.of_linear([LinearTerm(2, c.RD)]), k(0), d(c.RVD_0), d(c.RVD_1)] + [k(0)] * 6
        + _cpu32_ts(2) + [k(1), k(0), k(0)]))
    its.append(BusInteraction.sender(BUS_ALU, col(c.ALU), [
        BusValue.packed(Packing.DWORD_WL, c.ARG1_0), BusValue.packed(Packing.DWORD_WL, c.ARG2_0), d(c.ALU_FLAGS),
        BusValue.packed(Packing.DWORD_HL, c.RES[0])]))
    its.append(BusInteraction.sender(BUS_BYTE_ALU, mu, [
        k(ALU_OP_AND), k(1 << ALU_FLAGS_SIGNED), d(c.ALU_FLAGS),
        BusValue.of_linear([LinearTerm(1 << ALU_FLAGS_SIGNED, c.SIGNED)])]))
    its += [BusInteraction.sender(BUS_MSB16, col(gate), [d(half), d(sign)])
            for half, sign, gate in ((c.RV1[1], c.RV1_SIGN, c.SIGNED), (c.RV2[1], c.RV2_SIGN, c.SIGNED),
                                     (c.RES[1], c.RES_SIGN, c.MU))]
    its.append(BusInteraction.receiver(BUS_CPU32, mu, [
        BusValue.packed(Packing.DWORD_WL, c.TIMESTAMP_0), BusValue.packed(Packing.DWORD_WL, c.PC_0),
        d(c.HALF_INSTRUCTION_LENGTH)]))
    return its


def cpu32_table_air() -> AirProgram:
    """CPU32 as an AIR (cpu32.rs:600-724, 32 constraints, degree 3): 0-6 the
    flags and mu are bits; 7-8 the ADD carry pair of arg1 + arg2 = res (gated
    by add) and 9-10 of arg2 + res = arg1 (gated by sub); 11-16 unread register
    limbs are zero; 17-22 arg1, arg2 and rvd are the extensions of rv1, rv2
    (+ imm) and res; 23-24 no sign on unsigned rows; 25-26 no immediate when
    rs2 is read; 27-31 no register flag, signed or res sign on padding rows;
    then the LogUp constraints of its 23 interactions."""
    c = Cpu32Cols
    b = ir.IrBuilder()
    m = lambda col: b.main(0, col)  # noqa: E731
    one, inv, s16, fill = b.const_base(1), b.const_base(INV_SHIFT_32), b.const_base(1 << 16), b.const_base(CPU32_HI_FILL)
    for i, col in enumerate((c.READ_REGISTER1, c.READ_REGISTER2, c.WRITE_REGISTER, c.ALU, c.ADD, c.SUB, c.MU)):
        _emit_is_bit(b, i, col)
    res = (m(c.RES[0]) + m(c.RES[1]) * s16, m(c.RES[2]) + m(c.RES[3]) * s16)
    arg1, arg2 = (m(c.ARG1_0), m(c.ARG1_1)), (m(c.ARG2_0), m(c.ARG2_1))
    for idx, cond, lhs, rhs, out in ((7, c.ADD, arg1, arg2, res), (9, c.SUB, arg2, res, arg1)):
        carry0 = (lhs[0] + rhs[0] - out[0]) * inv
        carry1 = (lhs[1] + rhs[1] + carry0 - out[1]) * inv
        b.emit(idx, m(cond) * carry0 * (one - carry0))
        b.emit(idx + 1, m(cond) * carry1 * (one - carry1))
    idx = 11
    for read, limbs in ((c.READ_REGISTER1, c.RV1), (c.READ_REGISTER2, c.RV2)):
        for limb in limbs:
            b.emit(idx, (one - m(read)) * m(limb))
            idx += 1
    b.emit(17, m(c.ARG1_0) - m(c.RV1[0]) - s16 * m(c.RV1[1]))
    b.emit(18, m(c.ARG1_1) - fill * m(c.RV1_SIGN))
    b.emit(19, m(c.ARG2_0) - m(c.RV2[0]) - s16 * m(c.RV2[1]) - m(c.IMM_0))
    b.emit(20, m(c.ARG2_1) - fill * m(c.RV2_SIGN) - m(c.IMM_1))
    b.emit(21, m(c.RVD_0) - m(c.RES[0]) - s16 * m(c.RES[1]))
    b.emit(22, m(c.RVD_1) - fill * m(c.RES_SIGN))
    for i, sign in enumerate((c.RV1_SIGN, c.RV2_SIGN)):
        b.emit(23 + i, (one - m(c.SIGNED)) * m(sign))
    for i, imm in enumerate((c.IMM_0, c.IMM_1)):
        b.emit(25 + i, m(c.READ_REGISTER2) * m(imm))
    for i, flag in enumerate((c.READ_REGISTER1, c.READ_REGISTER2, c.WRITE_REGISTER, c.SIGNED, c.RES_SIGN)):
        b.emit(27 + i, (one - m(c.MU)) * m(flag))
    its = cpu32_bus_interactions()
    n_total = emit_logup_constraints(b, its, 32)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("cpu32", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3, 1 + absorbed))


# DECODE (prover/src/tables/decode.rs): the program's decoded instructions as
# fixed columns (pc, packed_decode, imm) with a multiplicity, one receiver of
# the CPU/CPU32 DECODE lookups. The RV64IM decoder (executor/src/vm/
# instruction/decoding.rs) and the decode-word layout (types.rs:580-1100) are
# restated with it.
@dataclass(frozen=True)
class Instruction:
    """One decoded instruction (decoding.rs:172-240). kind: arith, arith_imm,
    arith_w, arith_imm_w, jal, jalr, store, load, branch, lui, auipc, csr,
    ecall, fence; op an ARITH_OP_FLAGS key, width b/h/w/d/bu/hu/wu, cond
    eq/ne/lt/ge/ltu/geu."""
    kind: str
    dst: int = 0
    src1: int = 0
    src2: int = 0
    imm: int = 0
    op: str = ""
    width: str = ""
    cond: str = ""
    csr: int = 0


_R_OPS = {(0, 0x00): "add", (0, 0x20): "sub", (4, 0): "xor", (6, 0): "or", (7, 0): "and", (1, 0): "sll",
          (5, 0): "srl", (5, 0x20): "sra", (2, 0): "slt", (3, 0): "sltu", (0, 1): "mul", (1, 1): "mulh",
          (2, 1): "mulhsu", (3, 1): "mulhu", (4, 1): "div", (5, 1): "divu", (6, 1): "rem", (7, 1): "remu"}
_W_OPS = {"add", "sub", "sll", "srl", "sra", "mul", "div", "divu", "rem", "remu"}
_LOAD_WIDTH = {0: "b", 1: "h", 2: "w", 3: "d", 4: "bu", 5: "hu", 6: "wu"}
_BRANCH_COND = {0: "eq", 1: "ne", 4: "lt", 5: "ge", 6: "ltu", 7: "geu"}


def decode_instruction(word: int) -> Instruction:
    """decoding.rs:255-600: the RV64IM decoder (I-type immediates sign-extended
    from 12 bits, shift immediates 6 bits — 5 for the word forms — with their
    function bits checked; ECALL/EBREAK one instruction; CSRRW/CSRRS/CSRRCI
    accepted, the other CSR forms rejected). Raises ValueError on an encoding
    the reference rejects."""
    word &= 0xFFFFFFFF
    opc, rd, f3 = word & 0x7F, (word >> 7) & 0x1F, (word >> 12) & 7
    rs1, rs2, f7 = (word >> 15) & 0x1F, (word >> 20) & 0x1F, word >> 25
    neg = word >> 31 == 1
    if opc in (0b0110011, 0b0111011):  # R: OP, OP-32
        op = _R_OPS.get((f3, f7))
        if op is None:
            raise ValueError(f"unknown func3/func7 {f3:#x}/{f7:#x}")
        if opc == 0b0110011:
            return Instruction("arith", rd, rs1, rs2, op=op)
        if op not in _W_OPS:
            raise ValueError("invalid word instruction")
        return Instruction("arith_w", rd, rs1, rs2, op=op)
    if opc in (0b0010011, 0b0011011, 0b0000011, 0b1100111, 0b1110011, 0b0001111):  # I
        csr = word >> 20
        imm = csr - (1 << 12) if neg else csr
        if opc == 0b0010011:
            ops = {0: "add", 4: "xor", 6: "or", 7: "and", 2: "slt", 3: "sltu"}
            if f3 == 1:
                if imm >> 6 != 0:
                    raise ValueError("unknown shift-left variant")
                op, imm = "sll", imm & 0x3F
            elif f3 == 5:
                fid, imm = imm >> 6, imm & 0x3F
                op = {0: "srl", 0x10: "sra"}.get(fid)
                if op is None:
                    raise ValueError("unknown shift-right variant")
            else:
                op = ops[f3]
            return Instruction("arith_imm", rd, rs1, imm=imm, op=op)
        if opc == 0b0011011:
            if f3 == 0:
                op = "add"
            elif f3 == 1:
                if imm >> 5 != 0:
                    raise ValueError("unknown shift-left variant")
                op, imm = "sll", imm & 0x1F
            elif f3 == 5:
                fid, imm = imm >> 5, imm & 0x1F
                op = {0: "srl", 0x20: "sra"}.get(fid)
                if op is None:
                    raise ValueError("unknown shift-right variant")
            else:
                raise ValueError("invalid word instruction")
            return Instruction("arith_imm_w", rd, rs1, imm=imm, op=op)
        if opc == 0b1100111:
            if f3 != 0:
                raise ValueError("invalid JALR")
            return Instruction("jalr", rd, rs1, imm=imm)
        if opc == 0b0000011:
            if f3 not in _LOAD_WIDTH:
                raise ValueError("invalid load width")
            return Instruction("load", rd, rs1, imm=imm, width=_LOAD_WIDTH[f3])
        if opc == 0b1110011:
            if f3 == 0:
                return Instruction("ecall")
            if f3 in (1, 2, 7):
                return Instruction("csr", rd, rs1, op={1: "csrrw", 2: "csrrs", 7: "csrrci"}[f3], csr=csr)
            raise ValueError("invalid system instruction")
        return Instruction("fence")
    if opc == 0b0100011:  # S
        imm = ((word >> 25) << 5) | rd
        imm = imm - (1 << 12) if neg else imm
        if f3 > 3:
            raise ValueError("invalid store width")
        return Instruction("store", 0, rs1, rs2, imm=imm, width=_LOAD_WIDTH[f3])
    if opc == 0b1100011:  # B
        imm = ((word >> 20) & 0x7E0) | ((word >> 7) & 0x1E) | ((word & 0x80) << 4)
        imm = imm - (1 << 12) if neg else imm
        if f3 not in _BRANCH_COND:
            raise ValueError("unknown branch condition")
        return Instruction("branch", 0, rs1, rs2, imm=imm, cond=_BRANCH_COND[f3])
    if opc == 0b1101111:  # J
        imm = (word & 0xFF000) | ((word & 0x100000) >> 9) | ((word >> 20) & 0x7FE)
        return Instruction("jal", rd, imm=imm - (1 << 20) if neg else imm)
    if opc in (0b0110111, 0b0010111):  # U
        return Instruction("lui" if opc == 0b0110111 else "auipc", rd, imm=word & 0xFFFFF000)
    raise ValueError(f"unknown opcode {opc:#x}")


@dataclass(frozen=True)
class ShrunkDecode:
    """types.rs:655-700: the decode fields packed into one field element."""
    read_register1: bool = False
    read_register2: bool = False
    write_register: bool = False
    word_instr: bool = False
    alu: bool = False
    add: bool = False
    sub: bool = False
    memory: bool = False
    branch: bool = False
    ecall: bool = False
    rs1: int = 0
    rs2: int = 0
    rd: int = 0
    half_instruction_length: int = 0
    alu_flags: int = 0
    mem_flags: int = 0

    def pack(self) -> int:
        """types.rs:702-720 (58 bits)."""
        bits = ((self.read_register1, PD_READ_REG1), (self.read_register2, PD_READ_REG2),
                (self.write_register, PD_WRITE_REG), (self.word_instr, PD_WORD_INSTR), (self.alu, PD_ALU),
                (self.add, PD_ADD), (self.sub, PD_SUB), (self.memory, PD_MEMORY), (self.branch, PD_BRANCH),
                (self.ecall, PD_ECALL), (self.rs1, PD_RS1), (self.rs2, PD_RS2), (self.rd, PD_RD),
                (self.half_instruction_length, PD_HALF_INSTRUCTION_LENGTH), (self.alu_flags, PD_ALU_FLAGS),
                (self.mem_flags, PD_MEM_FLAGS))
        return sum(int(v) << pos for v, pos in bits)


def build_mem_flags(jalr_or_op, signed, m2, m4, m8) -> int:
    """types.rs:635-648: JALR|memory_op, signed, 2-, 4-, 8-byte bits."""
    return int(jalr_or_op) | int(signed) << 1 | int(m2) << 2 | int(m4) << 3 | int(m8) << 4


_STORE_WIDTH_BITS = {"b": (0, 0, 0), "bu": (0, 0, 0), "h": (1, 0, 0), "hu": (1, 0, 0), "w": (0, 1, 0),
                     "wu": (0, 1, 0), "d": (0, 0, 1)}
_LOAD_WIDTH_BITS = {"b": (0, 0, 0, 1), "bu": (0, 0, 0, 0), "h": (1, 0, 0, 1), "hu": (1, 0, 0, 0),
                    "w": (0, 1, 0, 1), "wu": (0, 1, 0, 0), "d": (0, 0, 1, 0)}
# branch condition -> (alu op, signed, invert) (types.rs:1025-1034)
_BRANCH_FLAGS = {"eq": (ALU_OP_EQ, 0, 0), "ne": (ALU_OP_EQ, 0, 1), "lt": (ALU_OP_LT, 1, 0), "ltu": (ALU_OP_LT, 0, 0),
                 "ge": (ALU_OP_LT, 1, 1), "geu": (ALU_OP_LT, 0, 1)}


def shrunk_decode(ins: Instruction, instruction_length: int = 4) -> ShrunkDecode:
    """types.rs:734-860: the decode fields of an instruction (JAL as JALR rd,
    x255; AUIPC as ADDI rd, x255; LUI as ADDI rd, x0; FENCE and CSR as ADDI x0,
    x0, 0; ECALL reads a7)."""
    k, f = ins.kind, dict(half_instruction_length=instruction_length // 2)
    if k in ("arith", "arith_w", "arith_imm", "arith_imm_w"):
        word = k in ("arith_w", "arith_imm_w")
        flags, alu, add, sub = arith_decode(ins.op, word)
        f.update(rd=ins.dst, rs1=ins.src1, read_register1=ins.src1 != 0, write_register=ins.dst != 0,
                 word_instr=word, alu=alu, add=add, sub=sub, alu_flags=flags)
        if k in ("arith", "arith_w"):
            f.update(rs2=ins.src2, read_register2=ins.src2 != 0)
    elif k == "jal":
        f.update(rd=ins.dst, rs1=255, read_register1=True, write_register=ins.dst != 0, branch=True,
                 mem_flags=build_mem_flags(1, 0, 0, 0, 0))
    elif k == "jalr":
        f.update(rd=ins.dst, rs1=ins.src1, read_register1=ins.src1 != 0, write_register=ins.dst != 0, branch=True,
                 mem_flags=build_mem_flags(1, 0, 0, 0, 0))
    elif k == "store":
        f.update(rs1=ins.src1, rs2=ins.src2, read_register1=ins.src1 != 0, read_register2=ins.src2 != 0, add=True,
                 memory=True, mem_flags=build_mem_flags(1, 0, *_STORE_WIDTH_BITS[ins.width]))
    elif k == "load":
        m2, m4, m8, s = _LOAD_WIDTH_BITS[ins.width]
        f.update(rd=ins.dst, rs1=ins.src1, read_register1=ins.src1 != 0, write_register=ins.dst != 0, add=True,
                 memory=True, mem_flags=build_mem_flags(0, s, m2, m4, m8))
    elif k == "branch":
        op, s, inv = _BRANCH_FLAGS[ins.cond]
        f.update(rs1=ins.src1, rs2=ins.src2, read_register1=ins.src1 != 0, read_register2=ins.src2 != 0,
                 branch=True, alu=True, alu_flags=op + 32 * s + 64 * inv)
    elif k == "lui":
        f.update(rd=ins.dst, write_register=ins.dst != 0, add=True)
    elif k == "auipc":
        f.update(rd=ins.dst, rs1=255, read_register1=True, write_register=ins.dst != 0, add=True)
    elif k == "ecall":
        f.update(rs1=17, read_register1=True, ecall=True)
    else:  # fence, csr
        f.update(add=True)
    return ShrunkDecode(**f)


def instruction_imm(ins: Instruction) -> int:
    """types.rs:1096-1110: the immediate sign-extended to 64 bits (0 for R-type,
    ECALL, FENCE, CSR)."""
    if ins.kind in ("arith_imm", "arith_imm_w", "jal", "jalr", "store", "load", "branch"):
        return ins.imm & (2**64 - 1)
    if ins.kind in ("lui", "auipc"):
        v = ins.imm & 0xFFFFFFFF
        return (v - (1 << 32) if v >> 31 else v) & (2**64 - 1)
    return 0


class DecodeCols:
    """decode.rs:48-78: pc (DWordWL), packed_decode, imm (DWordWL) fixed; mu."""
    PC_0, PC_1, PACKED_DECODE, IMM_0, IMM_1, MU = range(6)
    NUM_COLUMNS = 6
    NUM_PRECOMPUTED = 5


def decode_entries_from_words(segments) -> list:
    """decode.rs:356-377 tables_from_elf: (pc, imm, packed) of every 4-byte
    word of the executable segments [(base_address, words)], in segment order."""
    out = []
    for base, words in segments:
        for i, w in enumerate(words):
            ins = decode_instruction(w)
            out.append((base + 4 * i, instruction_imm(ins), shrunk_decode(ins, 4).pack()))
    return out


def decode_precomputed(entries) -> np.ndarray:
    """decode.rs:379-429: the entries, then the CPU padding entry (pc =
    CPU_PADDING_PC, zero decode), then padding entries (pc 1, zero decode) to a
    power of two rows (at least 2)."""
    rows = max(2, 1 << len(entries).bit_length())  # next power of two of len + 1
    t = np.zeros((rows, DecodeCols.NUM_PRECOMPUTED), dtype=np.uint64)
    t[:, DecodeCols.PC_0] = 1  # padding_entry pc (types.rs:1073-1079)
    t[len(entries), DecodeCols.PC_0] = CPU_PADDING_PC
    for r, (pc, imm, packed) in enumerate(entries):
        t[r] = (pc & 0xFFFFFFFF, pc >> 32, packed, imm & 0xFFFFFFFF, imm >> 32)
    return t


def generate_decode_trace(entries, lookups) -> np.ndarray:
    """The fixed columns and the multiplicity: one count per looked-up pc
    (decode.rs:140-155 update_multiplicities; CPU_PADDING_PC counts on the CPU
    padding row, pcs outside the program are not counted)."""
    pre = decode_precomputed(entries)
    t = np.zeros((pre.shape[0], DecodeCols.NUM_COLUMNS), dtype=np.uint64)
    t[:, :DecodeCols.NUM_PRECOMPUTED] = pre
    row_of = {pc: r for r, (pc, _, _) in enumerate(entries)}
    row_of[CPU_PADDING_PC] = len(entries)
    for pc in lookups:
        if pc in row_of:
            t[row_of[pc], DecodeCols.MU] += 1
    return t


def decode_bus_interactions():
    """decode.rs:157-180: DECODE[pc, imm, packed_decode] received with mu."""
    c = DecodeCols
    return [BusInteraction.receiver(BUS_DECODE, Multiplicity("Column", (c.MU,)), [
        BusValue.packed(Packing.DWORD_WL, c.PC_0), BusValue.packed(Packing.DWORD_WL, c.IMM_0),
        BusValue.packed(Packing.DIRECT, c.PACKED_DECODE)])]


def decode_table_air(entries) -> AirProgram:
    """DECODE as an AIR: the LogUp constraints of its one receiver; the first
    five columns fixed (they depend on the program, so no static commitment of
    the reference pins them)."""
    c = DecodeCols
    its = decode_bus_interactions()
    pre = decode_precomputed(entries)
    return _bus_only_air("decode", c.NUM_COLUMNS, its, c.NUM_PRECOMPUTED, lambda: pre)


# CPU (prover/src/tables/cpu.rs, constraints/cpu.rs): the central execution
# table, one row per executed instruction. It fetches through DECODE, reads and
# writes registers through MEMW, dispatches to the ALU chips (ALU bus), to
# LOAD/STORE (MEMORY op bus), to BRANCH and ECALL, delegates word instructions
# to CPU32, and carries the pc as inline Memory-bus tokens at address 510/511.
BUS_ECALL_CPU = BUS_ECALL


class CpuCols:
    """cpu.rs:218-307 (TIMESTAMP fixed)."""
    TIMESTAMP, PC_0, PC_1, RS1, RS2, RD = 0, 1, 2, 3, 4, 5
    READ_REGISTER1, READ_REGISTER2, WRITE_REGISTER, IMM_0, IMM_1 = 6, 7, 8, 9, 10
    HALF_INSTRUCTION_LENGTH, WORD_INSTR, ALU, ALU_FLAGS, ADD, SUB = 11, 12, 13, 14, 15, 16
    MEMORY, MEM_FLAGS, BRANCH, ECALL, NEXT_PC_0, NEXT_PC_1, RVD_0, RVD_1 = 17, 18, 19, 20, 21, 22, 23, 24
    PC_DOUBLE_READ, RV1_0, RV1_1, RV2_0, RV2_1, ARG2_0, ARG2_1 = 25, 26, 27, 28, 29, 30, 31
    RES = (32, 33, 34, 35)
    BRANCH_COND = 36
    NUM_COLUMNS = 37
    NUM_PRECOMPUTED = 1


@dataclass(frozen=True)
class CpuOperation:
    """cpu.rs:318-350: the decoded instruction (pc, imm, fields) and the row's
    run-time values."""
    pc: int
    imm: int
    fields: ShrunkDecode
    timestamp: int = 0
    next_pc: int = 0
    rvd: int = 0
    rv1: int = 0
    rv2: int = 0
    arg2: int = 0
    res: int = 0
    branch_cond: bool = False


def cpu_operation(pc, ins: Instruction, timestamp, src1_val, src2_val, dst_val, next_pc) -> CpuOperation:
    """cpu.rs:381-500 from_log: the executor's log of one step (register values
    read, the value written, the next pc) and the decoded instruction. rv1 is
    the pc for the x255 reads (JAL, AUIPC); arg2 = imm (memory), rv2 (branch)
    or rv2 + imm; res = rv1 ± arg2 on the ADD/SUB paths, the branch decision
    on branches, the logged result otherwise; rvd = the loaded value (0 for
    stores), pc + length on branches, res otherwise; word rows only carry
    their register values for CPU32."""
    f = shrunk_decode(ins, 4)
    imm = instruction_imm(ins)
    M = 2**64 - 1
    length = 2 * f.half_instruction_length
    if f.word_instr:
        return CpuOperation(pc, imm, f, timestamp, (pc + length) & M, dst_val, src1_val,
                            src2_val if f.read_register2 else 0)
    rv1 = pc if f.rs1 == 255 else (src1_val if f.read_register1 else 0)
    rv2 = src2_val if f.read_register2 else 0
    jalr = f.mem_flags & 1 == 1
    arg2 = imm if f.memory else rv2 if f.branch else (rv2 + imm) & M
    cond = False
    if f.branch:
        if jalr:
            cond = True
        else:
            op, signed, invert = f.alu_flags & 0x1F, f.alu_flags >> 5 & 1, f.alu_flags >> 6 & 1
            s = lambda v: v - 2**64 if v >> 63 else v  # noqa: E731
            cmp = rv1 == rv2 if op == ALU_OP_EQ else (s(rv1) < s(rv2) if signed else rv1 < rv2) if op == ALU_OP_LT \
                else False
            cond = cmp != bool(invert)
    if f.add:
        res = (rv1 + arg2) & M
    elif f.sub:
        res = (rv1 - arg2) & M
    elif f.alu:
        res = int(cond) if f.branch else dst_val
    else:
        res = 0
    if f.memory:
        rvd = 0 if jalr else dst_val
    elif f.branch:
        rvd = (pc + length) & M
    else:
        rvd = res
    nxt = next_pc if (cond and not f.ecall) else (pc + length) & M
    return CpuOperation(pc, imm, f, timestamp, nxt, rvd, rv1, rv2, arg2, res, cond)


def generate_cpu_trace(ops) -> np.ndarray:
    """cpu.rs:624-700: one row per operation (TIMESTAMP = 4 (row + 1)); a word
    row keeps only pc, next pc, the half length and WORD_INSTR; padding rows
    hold pc = next pc = CPU_PADDING_PC; zero rows to a power of two (at least
    4)."""
    c = CpuCols
    n = max(4, 1 << max(0, (len(ops) - 1).bit_length()))
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)
    t[:, c.TIMESTAMP] = CPU_TIMESTAMP_STRIDE * np.arange(1, n + 1, dtype=np.uint64)
    t[len(ops):, c.PC_0] = CPU_PADDING_PC
    t[len(ops):, c.NEXT_PC_0] = CPU_PADDING_PC
    w = lambda r, col, v: (r.__setitem__(col, v & 0xFFFFFFFF), r.__setitem__(col + 1, (v >> 32) & 0xFFFFFFFF))  # noqa
    for i, op in enumerate(ops):
        f, r = op.fields, t[i]
        word = f.word_instr
        eff = lambda flag: int(flag and not word)  # noqa: E731
        w(r, c.PC_0, op.pc)
        if not word:
            r[c.RS1], r[c.RS2], r[c.RD] = f.rs1, f.rs2, f.rd
            w(r, c.IMM_0, op.imm)
            w(r, c.RVD_0, op.rvd)
            w(r, c.RV1_0, op.rv1)
            w(r, c.RV2_0, op.rv2)
            w(r, c.ARG2_0, op.arg2)
            r[c.RES[0]:c.RES[3] + 1] = [(op.res >> (16 * k)) & 0xFFFF for k in range(4)]
            r[c.ALU_FLAGS], r[c.MEM_FLAGS] = f.alu_flags, f.mem_flags
        r[c.READ_REGISTER1] = eff(f.read_register1 and f.rs1 != 0)
        r[c.READ_REGISTER2] = eff(f.read_register2 and f.rs2 != 0)
        r[c.WRITE_REGISTER] = eff(f.write_register and f.rd != 0)
        r[c.HALF_INSTRUCTION_LENGTH], r[c.WORD_INSTR] = f.half_instruction_length, int(word)
        r[c.ALU], r[c.ADD], r[c.SUB] = eff(f.alu), eff(f.add), eff(f.sub)
        r[c.MEMORY], r[c.BRANCH], r[c.ECALL] = eff(f.memory), eff(f.branch), eff(f.ecall)
        w(r, c.NEXT_PC_0, op.next_pc)
        r[c.BRANCH_COND] = int(op.branch_cond)
        r[c.PC_DOUBLE_READ] = int(not word and f.read_register1 and f.rs1 == 255)
    return t


def cpu_bus_interactions(ts_base: int = 0):
    """cpu.rs:832-1160, in its order (20): DECODE[pc, imm, packed decode] (1 -
    word_instr); ALU[rv1, arg2, alu_flags, res] (ALU); CPU32[ts, 0, pc,
    half length] (word_instr); MEMW reads of rs1 at ts and rs2 at ts + 1 and
    the write of rd at ts + 2 (24/24/16 elements, their flags); the MEMORY
    op [ts, 0, res, rv2, mem_flags, rvd] (MEMORY); per pc word the Memory
    token [1, 510 + i, 0, ts - 3 + 3 pc_double_read, 0, pc_i] sent and [1, 510 +
    i, 0, ts + 1, 0, next_pc_i] received (multiplicity one); BRANCH[next pc,
    pc, imm, rv1, mem_flags] (branch_cond); ARE_BYTES of (rs1, rs2), (rd, half
    length), (alu_flags, mem_flags) and IS_HALFWORD of the res halves (one);
    ECALL[ts, 0, rv1] (ECALL). Timestamps are the fixed column plus ts_base."""
    c = CpuCols
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    wl = lambda col: BusValue.packed(Packing.DWORD_WL, col)  # noqa: E731
    k = lambda v: BusValue.of_linear([LinearTerm(v)])  # noqa: E731
    col = lambda cc: Multiplicity("Column", (cc,))  # noqa: E731
    ts = lambda off: BusValue.of_linear([LinearTerm(1, c.TIMESTAMP), LinearTerm(ts_base + off)])  # noqa: E731
    res_wl = BusValue.packed(Packing.DWORD_HL, c.RES[0])
    packed = [LinearTerm(1 << pos, cc) for pos, cc in (
        (PD_READ_REG1, c.READ_REGISTER1), (PD_READ_REG2, c.READ_REGISTER2), (PD_WRITE_REG, c.WRITE_REGISTER),
        (PD_WORD_INSTR, c.WORD_INSTR), (PD_ALU, c.ALU), (PD_ADD, c.ADD), (PD_SUB, c.SUB), (PD_MEMORY, c.MEMORY),
        (PD_BRANCH, c.BRANCH), (PD_ECALL, c.ECALL), (PD_RS1, c.RS1), (PD_RS2, c.RS2), (PD_RD, c.RD),
        (PD_HALF_INSTRUCTION_LENGTH, c.HALF_INSTRUCTION_LENGTH), (PD_ALU_FLAGS, c.ALU_FLAGS),
        (PD_MEM_FLAGS, c.MEM_FLAGS))]
    zeros6 = [k(0)] * 6
    its = [BusInteraction.sender(BUS_DECODE, Multiplicity("Negated", (c.WORD_INSTR,)),
                                 [wl(c.PC_0), wl(c.IMM_0), BusValue.of_linear(packed)]),
           BusInteraction.sender(BUS_ALU, col(c.ALU), [wl(c.RV1_0), wl(c.ARG2_0), d(c.ALU_FLAGS), res_wl]),
           BusInteraction.sender(BUS_CPU32, col(c.WORD_INSTR), [ts(0), k(0), wl(c.PC_0), d(c.HALF_INSTRUCTION_LENGTH)])]
    for flag, rs, lo, off in ((c.READ_REGISTER1, c.RS1, c.RV1_0, 0), (c.READ_REGISTER2, c.RS2, c.RV2_0, 1)):
        val = [d(lo), d(lo + 1)] + zeros6
        its.append(BusInteraction.sender(BUS_MEMW, col(flag), val + [k(1), BusValue.of_linear([LinearTerm(2, rs)]),
                                                                      k(0)] + val + [ts(off), k(0), k(1), k(0), k(0)]))
    its.append(BusInteraction.sender(BUS_MEMW, col(c.WRITE_REGISTER), [
        k(1), BusValue.of_linear([LinearTerm(2, c.RD)]), k(0), d(c.RVD_0), d(c.RVD_1)] + zeros6
        + [ts(2), k(0), k(1), k(0), k(0)]))
    its.append(BusInteraction.sender(BUS_MEMORY_OP, col(c.MEMORY), [ts(0), k(0), res_wl, wl(c.RV2_0), d(c.MEM_FLAGS),
                                                                    wl(c.RVD_0)]))
    prev = BusValue.of_linear([LinearTerm(1, c.TIMESTAMP), LinearTerm(ts_base - 3), LinearTerm(3, c.PC_DOUBLE_READ)])
    for i in range(2):
        its.append(BusInteraction.sender(BUS_MEMORY, ONE, [k(1), k(510 + i), k(0), prev, k(0), d(c.PC_0 + i)]))
        its.append(BusInteraction.receiver(BUS_MEMORY, ONE, [k(1), k(510 + i), k(0), ts(1), k(0), d(c.NEXT_PC_0 + i)]))
    its.append(BusInteraction.sender(BUS_BRANCH, col(c.BRANCH_COND), [d(cc) for cc in (
        c.NEXT_PC_0, c.NEXT_PC_1, c.PC_0, c.PC_1, c.IMM_0, c.IMM_1, c.RV1_0, c.RV1_1, c.MEM_FLAGS)]))
    its += [BusInteraction.sender(BUS_ARE_BYTES, ONE, [d(a), d(b)]) for a, b in (
        (c.RS1, c.RS2), (c.RD, c.HALF_INSTRUCTION_LENGTH), (c.ALU_FLAGS, c.MEM_FLAGS))]
    its += [BusInteraction.sender(BUS_IS_HALFWORD, ONE, [d(cc)]) for cc in c.RES]
    its.append(BusInteraction.sender(BUS_ECALL, col(c.ECALL), [ts(0), k(0), wl(c.RV1_0)]))
    return its


def cpu_table_air(height: int, ts_base: int = 0) -> AirProgram:
    """CPU as an AIR (constraints/cpu.rs:278-379, 38 constraints): 0-10 the
    flags and pc_double_read are bits; 11-12 the ADD carry pair of rv1 + arg2 =
    res (gated by ADD), 13-14 of arg2 + res = rv1 (gated by SUB); 15-20 a word
    row sets no MEMORY, BRANCH, ECALL, register write or reads; 21-22 arg2 is
    imm (MEMORY), rv2 (BRANCH) or rv2 + imm; 23-26 unread registers are zero;
    27-28 rvd = res off MEMORY and BRANCH; 29-30 BRANCH: the carry pair of
    rvd = pc + 2 half length; 31 branch_cond = BRANCH (JALR + (1 - JALR)
    res[0]); 32-33 without a taken branch the carry pair of next pc = pc +
    length; 34 MEMORY·BRANCH = 0; 35-36 no immediate beside rs2 on ALU rows; 37
    mem_flags a bit off MEMORY; then the LogUp constraints of its 20
    interactions. The TIMESTAMP column is fixed."""
    c = CpuCols
    b = ir.IrBuilder()
    m = lambda col: b.main(0, col)  # noqa: E731
    one, inv, s16 = b.const_base(1), b.const_base(INV_SHIFT_32), b.const_base(1 << 16)
    flags = (c.READ_REGISTER1, c.READ_REGISTER2, c.WRITE_REGISTER, c.WORD_INSTR, c.ALU, c.ADD, c.SUB, c.MEMORY,
             c.BRANCH, c.ECALL, c.PC_DOUBLE_READ)
    for i, col in enumerate(flags):
        _emit_is_bit(b, i, col)
    res = (m(c.RES[0]) + m(c.RES[1]) * s16, m(c.RES[2]) + m(c.RES[3]) * s16)
    rv1, arg2 = (m(c.RV1_0), m(c.RV1_1)), (m(c.ARG2_0), m(c.ARG2_1))

    def carry_pair(idx, gate, lhs, rhs, out):
        c0 = (lhs[0] + rhs[0] - out[0]) * inv
        c1 = (lhs[1] + rhs[1] + c0 - out[1]) * inv
        b.emit(idx, gate * c0 * (one - c0))
        b.emit(idx + 1, gate * c1 * (one - c1))

    carry_pair(11, m(c.ADD), rv1, arg2, res)
    carry_pair(13, m(c.SUB), arg2, res, rv1)
    for i, col in enumerate((c.MEMORY, c.BRANCH, c.ECALL, c.WRITE_REGISTER, c.READ_REGISTER1, c.READ_REGISTER2)):
        b.emit(15 + i, m(c.WORD_INSTR) * m(col))
    mem, br = m(c.MEMORY), m(c.BRANCH)
    for i, (a2, imm, rv2) in enumerate(((c.ARG2_0, c.IMM_0, c.RV2_0), (c.ARG2_1, c.IMM_1, c.RV2_1))):
        b.emit(21 + i, m(a2) - (mem * m(imm) + br * m(rv2) + (one - mem - br) * (m(rv2) + m(imm))))
    for i, (flag, v) in enumerate(((c.READ_REGISTER1, c.RV1_0), (c.READ_REGISTER1, c.RV1_1),
                                   (c.READ_REGISTER2, c.RV2_0), (c.READ_REGISTER2, c.RV2_1))):
        b.emit(23 + i, (one - m(flag)) * m(v))
    for i in range(2):
        b.emit(27 + i, (one - mem - br) * (m(c.RVD_0 + i) - res[i]))
    half = m(c.HALF_INSTRUCTION_LENGTH)
    pc_len = (m(c.PC_0) + half + half, m(c.PC_1))
    zero = b.const_base(0)
    carry_pair(29, br, pc_len, (zero, zero), (m(c.RVD_0), m(c.RVD_1)))
    jalr = m(c.MEM_FLAGS)
    b.emit(31, m(c.BRANCH_COND) - (br * jalr + br * (one - jalr) * m(c.RES[0])))
    carry_pair(32, one - m(c.BRANCH_COND), pc_len, (zero, zero), (m(c.NEXT_PC_0), m(c.NEXT_PC_1)))
    b.emit(34, mem * br)
    for i, imm in enumerate((c.IMM_0, c.IMM_1)):
        b.emit(35 + i, (one - mem - br) * m(c.READ_REGISTER2) * m(imm))
    b.emit(37, (one - mem) * jalr * (one - jalr))
    its = cpu_bus_interactions(ts_base)
    n_total = emit_logup_constraints(b, its, 38)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("cpu", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3, 1 + absorbed), num_precomputed=c.NUM_PRECOMPUTED,
                      precomputed=lambda: cpu_timestamp_precomputed(height))


# ECDAS (prover/src/tables/ecdas.rs) with the secp256k1 double-and-add witness
# of crypto/ecsm (curve.rs, witness.rs): one row per double (op 0) or add (op
# 1) step of k·G. λ, xR and yR are pinned by three byte-limb convolution
# relations with 33-byte quotients, the offset r = 3p and 64 signed carries
# each (range-checked through IS_HALFWORD with per-relation offsets); the
# accumulator travels on the self-referential Ecdas bus, the scalar bits of
# the adds on the Bit bus.
BUS_ECDAS, BUS_BIT = 28, 30     # types.rs:350, :354
SECP_P = 2**256 - 2**32 - 977   # crypto/ecsm/src/lib.rs P_BYTES
SECP_N = 0xFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFEBAAEDCE6AF48A03BBFD25E8CD0364141  # N_BYTES
SECP_B = 7
SECP_R = 3 * SECP_P             # R_BYTES (33 bytes)
ECDAS_CARRY_OFFSETS = (32636, 8161, 16320)  # λ, xR, yR (ecdas.rs:24-26)


def secp_recover_y(x: int):
    """The even square root of x^3 + 7 mod p (curve.rs:26-33), None for a
    non-residue; p = 3 mod 4, so the root is a power."""
    rhs = (pow(x, 3, SECP_P) + SECP_B) % SECP_P
    y = pow(rhs, (SECP_P + 1) // 4, SECP_P)
    if y * y % SECP_P != rhs:
        return None
    return y if y % 2 == 0 else SECP_P - y


def ecsm_schedule(k: int):
    """curve.rs:105-125: (round, op, next_op) per step; from bit msb - 1 down,
    a double at every round and an add at the rounds whose bit is set."""
    sched, rnd, op = [], k.bit_length() - 2, 0
    while rnd >= 0:
        nxt = ((k >> rnd) & 1) if op == 0 else 0
        sched.append((rnd, op, nxt))
        sent = rnd - (1 - nxt)
        if sent < 0:
            break
        rnd, op = sent, nxt
    return sched


def ecsm_steps(k: int, gx: int, gy: int):
    """curve.rs:223-300: the affine steps (A, round, op, next_op, R, λ) of k·G
    (λ = 3xA²/2yA for a double, (yG - yA)/(xG - xA) for an add)."""
    out, ax, ay = [], gx, gy
    for rnd, op, nxt in ecsm_schedule(k):
        if op == 0:
            lam = 3 * ax * ax * pow(2 * ay, -1, SECP_P) % SECP_P
            rx = (lam * lam - 2 * ax) % SECP_P
        else:
            lam = (gy - ay) * pow(gx - ax, -1, SECP_P) % SECP_P
            rx = (lam * lam - ax - gx) % SECP_P
        ry = (lam * (ax - rx) - ay) % SECP_P
        out.append(((ax, ay), rnd, op, nxt, (rx, ry), lam))
        ax, ay = rx, ry
    return out, (ax, ay)


def _le(v: int, n: int) -> list:
    return [(v >> (8 * i)) & 0xFF for i in range(n)]


def _limb_carries(terms) -> list:
    """witness.rs:88-106: 2^8 c_i = c_{i-1} + terms_i, exact at every limb,
    closing carry 0."""
    c, carry = [], 0
    for t in terms:
        s = carry + t
        if s & 0xFF:
            raise ValueError("convolution limb not divisible by 256")
        carry = s >> 8
        c.append(carry)
    if c[63] != 0:
        raise ValueError("closing carry must be 0")
    return c


def _conv(a, b, i):
    return sum(a[j] * b[i - j] for j in range(i + 1))


@dataclass(frozen=True)
class EcdasStep:
    """witness.rs:60-85 and build_step (:381-470): one step's coordinates and λ
    as 32 little-endian bytes, quotients as 33 bytes, 64 carries per
    relation."""
    xa: int
    ya: int
    xg: int
    yg: int
    round: int
    op: int
    next_op: int
    lam: int
    xr: int
    yr: int

    def witness(self):
        p = SECP_P
        L, XA, YA, XG, YG, XR, YR = (_le(v, 32) + [0] * 32 for v in (self.lam, self.xa, self.ya, self.xg, self.yg,
                                                                     self.xr, self.yr))
        PP, RR = _le(p, 32) + [0] * 32, _le(SECP_R, 33) + [0] * 31
        num0 = (self.xg - self.xa) * self.lam - self.yg + self.ya if self.op \
            else 2 * self.lam * self.ya - 3 * self.xa * self.xa
        num1 = self.lam * self.lam - self.xa - self.xg - self.xr + (0 if self.op else self.xg - self.xa)
        num2 = self.lam * (self.xa - self.xr) - self.ya - self.yr
        qs = []
        for num in (num0, num1, num2):
            if num % p:
                raise ValueError("relation numerator not divisible by p")
            q = SECP_R + num // p
            if q < 0 or q >> 264:
                raise ValueError("quotient out of range")
            qs.append(q)
        Q = [_le(q, 33) + [0] * 31 for q in qs]
        rp = [_conv(RR, PP, i) for i in range(64)]
        t0 = [(YA[i] - YG[i] + sum(L[j] * (XG[i - j] - XA[i - j]) for j in range(i + 1))) if self.op else
              sum(2 * L[j] * YA[i - j] - 3 * XA[j] * XA[i - j] for j in range(i + 1)) for i in range(64)]
        t1 = [_conv(L, L, i) - XA[i] - XG[i] - XR[i] - (0 if self.op else XA[i] - XG[i]) for i in range(64)]
        t2 = [sum(L[j] * (XA[i - j] - XR[i - j]) for j in range(i + 1)) - YA[i] - YR[i] for i in range(64)]
        cs = [_limb_carries([t[i] + rp[i] - _conv(q, PP, i) for i in range(64)]) for t, q in zip((t0, t1, t2), Q)]
        return [_le(q, 33) for q in qs], cs


def ecdas_steps(k: int, gx: int):
    """The ECDAS steps of k·G for the base point with x-coordinate gx (its
    even y), k in [1, N) (witness.rs:281-330)."""
    if not 0 < k < SECP_N:
        raise ValueError("scalar out of range")
    if not 0 <= gx < SECP_P:
        raise ValueError("coordinate out of range")
    gy = secp_recover_y(gx)
    if gy is None:
        raise ValueError("not on the curve")
    steps, res = ecsm_steps(k, gx, gy)
    return [EcdasStep(a[0], a[1], gx, gy, rnd, op, nxt, lam, r[0], r[1]) for a, rnd, op, nxt, r, lam in steps], res


class EcdasCols:
    """ecdas.rs:30-58."""
    TIMESTAMP_0, TIMESTAMP_1 = 0, 1
    XG, YG, XA, YA = 2, 34, 66, 98
    ROUND, OP = 130, 131
    XR, YR, LAMBDA = 132, 164, 196
    Q0, C0, Q1, C1, Q2, C2 = 228, 261, 325, 358, 422, 455
    NEXT_OP, MU = 519, 520
    NUM_COLUMNS = 521


def generate_ecdas_trace(ops) -> np.ndarray:
    """ecdas.rs:80-130: ops are (timestamp, EcdasStep); one row each (mu = 1),
    carries as field elements; padding rows carry only op = 1; zero rows to a
    power of two (at least 4)."""
    c = EcdasCols
    n = max(4, 1 << max(0, (len(ops) - 1).bit_length()))
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)
    for i, (ts, s) in enumerate(ops):
        qs, cs = s.witness()
        r = t[i]
        r[c.TIMESTAMP_0], r[c.TIMESTAMP_1] = ts & 0xFFFFFFFF, ts >> 32
        for base, v in ((c.XG, s.xg), (c.YG, s.yg), (c.XA, s.xa), (c.YA, s.ya), (c.XR, s.xr), (c.YR, s.yr),
                        (c.LAMBDA, s.lam)):
            r[base:base + 32] = _le(v, 32)
        r[c.ROUND], r[c.OP], r[c.NEXT_OP], r[c.MU] = s.round, s.op, s.next_op, 1
        for qb, cb, q, cc in ((c.Q0, c.C0, qs[0], cs[0]), (c.Q1, c.C1, qs[1], cs[1]), (c.Q2, c.C2, qs[2], cs[2])):
            r[qb:qb + 33] = q
            r[cb:cb + 64] = [v % P for v in cc]
    t[len(ops):, c.OP] = 1
    return t


def ecdas_tuple(acc_x, acc_y, gen_x, gen_y, rnd, op, ts_lo, ts_hi):
    """ecsm.rs:585-606: [0 (the curve id), ts_lo, ts_hi, the 4 × 32 coordinate
    bytes, round, op]."""
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    v = [BusValue.of_linear([LinearTerm(0)]), ts_lo, ts_hi]
    for base in (acc_x, acc_y, gen_x, gen_y):
        v += [d(base + i) for i in range(32)]
    return v + [rnd, op]


def ecdas_bus_interactions():
    """ecdas.rs:152-245, in its order (388): the Ecdas receiver of the
    accumulator [A, G, round, op] (mu); ARE_BYTES[x, 0] of round and of the λ,
    q0, xR, q1, yR, q2 bytes (mu); IS_HALFWORD[c_i + offset] of carries 0..62
    of each relation (mu); Bit[ts, round] (next_op); the Ecdas sender of [R,
    G, round + next_op - 1, next_op] (mu)."""
    c = EcdasCols
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    mu = Multiplicity("Column", (c.MU,))
    ts = (d(c.TIMESTAMP_0), d(c.TIMESTAMP_1))
    its = [BusInteraction.receiver(BUS_ECDAS, mu, ecdas_tuple(c.XA, c.YA, c.XG, c.YG, d(c.ROUND), d(c.OP), *ts))]
    for base, n in ((c.ROUND, 1), (c.LAMBDA, 32), (c.Q0, 33), (c.XR, 32), (c.Q1, 33), (c.YR, 32), (c.Q2, 33)):
        its += [BusInteraction.sender(BUS_ARE_BYTES, mu, [d(base + i), BusValue.of_linear([LinearTerm(0)])])
                for i in range(n)]
    for base, off in zip((c.C0, c.C1, c.C2), ECDAS_CARRY_OFFSETS):
        its += [BusInteraction.sender(BUS_IS_HALFWORD, mu, [BusValue.of_linear([LinearTerm(1, base + i),
                                                                                LinearTerm(off)])])
                for i in range(63)]
    its.append(BusInteraction.sender(BUS_BIT, Multiplicity("Column", (c.NEXT_OP,)), list(ts) + [d(c.ROUND)]))
    its.append(BusInteraction.sender(BUS_ECDAS, mu, ecdas_tuple(
        c.XR, c.YR, c.XG, c.YG, BusValue.of_linear([LinearTerm(1, c.ROUND), LinearTerm(1, c.NEXT_OP), LinearTerm(-1)]),
        d(c.NEXT_OP), *ts)))
    return its


def ecdas_table_air() -> AirProgram:
    """ECDAS as an AIR (ecdas.rs:270-453, 200 constraints, degree 3): 0-2 mu,
    op, next_op bits; 3 op·next_op = 0; 4 next_op implies mu; then per
    relation (λ, xR, yR) 64 carry identities 256 c_i - c_{i-1} - S_i and c_63 =
    0, where S_i is limb i of the relation (op-selected for λ and xR) plus
    mu·(r·p)_i - (q·p)_i; then the LogUp constraints of its 388
    interactions."""
    c = EcdasCols
    b = ir.IrBuilder()
    m = lambda col: b.main(0, col)  # noqa: E731
    one, zero = b.const_base(1), b.const_base(0)
    pb, rb = _le(SECP_P, 32), _le(SECP_R, 33)
    at = lambda base, n, j: m(base + j) if j < n else zero  # noqa: E731
    for i, col in enumerate((c.MU, c.OP, c.NEXT_OP)):
        _emit_is_bit(b, i, col)
    b.emit(3, m(c.OP) * m(c.NEXT_OP))
    b.emit(4, m(c.NEXT_OP) * (one - m(c.MU)))

    def rq(i, qbase):
        rp = sum(rb[j] * pb[i - j] for j in range(i + 1) if j < 33 and i - j < 32)
        qp = None
        for j in range(i + 1):
            if j < 33 and i - j < 32 and pb[i - j]:
                t = m(qbase + j) * b.const_base(pb[i - j])
                qp = t if qp is None else qp + t
        out = m(c.MU) * b.const_base(rp)
        return out - qp if qp is not None else out

    lam = lambda j: at(c.LAMBDA, 32, j)  # noqa: E731
    xg, xa, ya = (lambda j: at(c.XG, 32, j)), (lambda j: at(c.XA, 32, j)), (lambda j: at(c.YA, 32, j))
    yg, xr, yr = (lambda j: at(c.YG, 32, j)), (lambda j: at(c.XR, 32, j)), (lambda j: at(c.YR, 32, j))
    op = m(c.OP)

    def s_i(rel, i):
        if rel == 0:
            ob = ya(i) - yg(i)
            nb = None
            for j in range(i + 1):
                if j < 32 and i - j < 32:
                    ob = ob + lam(j) * (xg(i - j) - xa(i - j))
                    t = b.const_base(2) * lam(j) * ya(i - j) - b.const_base(3) * xa(j) * xa(i - j)
                    nb = t if nb is None else nb + t
            nb = nb if nb is not None else zero
            return op * ob + (one - op) * nb + rq(i, c.Q0)
        if rel == 1:
            s = zero
            for j in range(i + 1):
                if j < 32 and i - j < 32:
                    s = s + lam(j) * lam(i - j)
            return s - xa(i) - xg(i) - xr(i) - (one - op) * (xa(i) - xg(i)) + rq(i, c.Q1)
        s = zero
        for j in range(i + 1):
            if j < 32 and i - j < 32:
                s = s + lam(j) * (xa(i - j) - xr(i - j))
        return s - ya(i) - yr(i) + rq(i, c.Q2)

    idx = 5
    for rel, cb in enumerate((c.C0, c.C1, c.C2)):
        for i in range(64):
            prev = m(cb + i - 1) if i else zero
            b.emit(idx, b.const_base(256) * m(cb + i) - prev - s_i(rel, i))
            idx += 1
        b.emit(idx, m(cb + 63))
        idx += 1
    its = ecdas_bus_interactions()
    n_total = emit_logup_constraints(b, its, idx)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("ecdas", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3, 1 + absorbed))


# ECSM (prover/src/tables/ecsm.rs): one row per ECALL(-11). Reads xG and k
# from memory through the x11/x12 pointers, witnesses yG and proves yG^2 =
# xG^3 + 7 through two convolution relations (x2 = xG^2 mod p with quotient q0,
# then yG^2 + p^2 - xG·x2 - 7 - q1·p = 0), bounds 0 < k < N, xG < p and xR < p
# by overflowing additions, serves the scalar's bits on the Bit bus, seeds and
# drains the ECDAS chain and writes xR back through the x10 pointer.
ECSM_SYSCALL_NUMBER = 2**64 - 1 - 10    # executor/src/vm/instruction/execution.rs:61
ECSM_CARRY_OFFSETS = (8160, 16319)      # x2, yG (ecsm.rs:27-28)


class EcsmCols:
    """ecsm.rs:33-98."""
    TIMESTAMP_0, TIMESTAMP_1 = 0, 1
    ADDR_XG_0, ADDR_XG_1, ADDR_K_0, ADDR_K_1, ADDR_XR_0, ADDR_XR_1 = 2, 3, 4, 5, 6, 7
    XR, YR, K, LEN_K = 8, 40, 72, 328
    XG, YG, X2, Q0, C0, Q1, C1 = 329, 361, 393, 425, 457, 521, 554
    XG_SUB_P, K_SUB_N, XR_SUB_P = 618, 634, 650
    MU = 666
    NUM_COLUMNS = 667


@dataclass(frozen=True)
class EcsmOperation:
    """ecsm.rs:119-130: the ecall's timestamp and the three memory addresses
    (xG, k, xR); the scalar and xG as integers."""
    timestamp: int
    addr_xg: int
    addr_k: int
    addr_xr: int
    k: int
    xg: int

    def witness(self):
        """witness.rs:281-345: yG, x2 = xG^2 mod p with q0, q1 = p + (yG^2 - xG·x2
        - 7)/p (33 bytes), their carries, the three (v - bound) mod 2^256
        addends, the top bit of k and the result R = k·G."""
        p = SECP_P
        steps, (rx, ry) = ecdas_steps(self.k, self.xg)
        yg = secp_recover_y(self.xg)
        x2 = self.xg * self.xg % p
        q0 = (self.xg * self.xg - x2) // p
        num = yg * yg - self.xg * x2 - SECP_B
        if num % p:
            raise ValueError("yG relation not divisible by p")
        q1 = p + num // p
        XG, YG, X2, Q0 = (_le(v, 32) + [0] * 32 for v in (self.xg, yg, x2, q0))
        Q1 = _le(q1, 33) + [0] * 31
        PP, BB = _le(p, 32) + [0] * 32, [SECP_B] + [0] * 63
        c0 = _limb_carries([_conv(XG, XG, i) - X2[i] - _conv(Q0, PP, i) for i in range(64)])
        c1 = _limb_carries([_conv(YG, YG, i) + _conv(PP, PP, i) - _conv(X2, XG, i) - _conv(Q1, PP, i) - BB[i]
                            for i in range(64)])
        sub = [(1 << 256) + v - bound for v, bound in ((self.xg, p), (self.k, SECP_N), (rx, p))]
        return dict(yg=yg, x2=x2, q0=q0, q1=q1, c0=c0, c1=c1, sub=sub, len_k=self.k.bit_length() - 1, xr=rx, yr=ry,
                    steps=steps)


def generate_ecsm_trace(ops) -> np.ndarray:
    """ecsm.rs:149-195: one row per ecall (mu = 1), all-zero padding rows to a
    power of two (at least 4)."""
    c = EcsmCols
    n = max(4, 1 << max(0, (len(ops) - 1).bit_length()))
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)
    for i, op in enumerate(ops):
        w = op.witness()
        r = t[i]
        for col, v in ((c.TIMESTAMP_0, op.timestamp), (c.ADDR_XG_0, op.addr_xg), (c.ADDR_K_0, op.addr_k),
                       (c.ADDR_XR_0, op.addr_xr)):
            r[col], r[col + 1] = v & 0xFFFFFFFF, v >> 32
        for base, v, nb in ((c.XR, w["xr"], 32), (c.YR, w["yr"], 32), (c.XG, op.xg, 32), (c.YG, w["yg"], 32),
                            (c.X2, w["x2"], 32), (c.Q0, w["q0"], 32), (c.Q1, w["q1"], 33)):
            r[base:base + nb] = _le(v, nb)
        r[c.K:c.K + 256] = [(op.k >> b) & 1 for b in range(256)]
        r[c.LEN_K] = w["len_k"]
        for base, v in zip((c.XG_SUB_P, c.K_SUB_N, c.XR_SUB_P), w["sub"]):
            r[base:base + 16] = [(v >> (16 * j)) & 0xFFFF for j in range(16)]
        r[c.C0:c.C0 + 64] = [v % P for v in w["c0"]]
        r[c.C1:c.C1 + 64] = [v % P for v in w["c1"]]
        r[c.MU] = 1
    return t


def _memw_read_tuple(value, is_register, base_lo, base_hi, ts_lo, ts_hi, w2, w8):
    k = lambda v: BusValue.of_linear([LinearTerm(v)])  # noqa: E731
    return list(value) + [k(is_register), base_lo, base_hi] + list(value) + [ts_lo, ts_hi, k(w2), k(0), k(w8)]


def ecsm_bus_interactions():
    """ecsm.rs:300-583, in its order (579): the Ecall receiver [ts, -11]; per
    operand the register read of its pointer (x11 xG at ts, x12 k at ts + 1,
    x10 xR at ts + 2) and four 8-byte MEMW accesses (reads of xG and of k —
    whose bytes are sums of its bit columns — and writes of xR); ARE_BYTES[x,
    0] of the x2, q0, yG, q1 bytes; IS_HALFWORD of carries 0..62 (offset) of
    both relations and of the 48 halfwords of the three overflow addends;
    ZERO[sum of k's bytes, 0]; Bit[ts, i] received with multiplicity k_i for
    every bit, Bit[ts, len_k] sent; the Ecdas seed [G, G, len_k - 1, 0] sent and
    result [R, G, -1, 0] received (all mu unless noted)."""
    c = EcsmCols
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    k = lambda v: BusValue.of_linear([LinearTerm(v)])  # noqa: E731
    mu = Multiplicity("Column", (c.MU,))
    ts_lo, ts_hi = d(c.TIMESTAMP_0), d(c.TIMESTAMP_1)
    ts_plus = lambda off: BusValue.of_linear([LinearTerm(1, c.TIMESTAMP_0), LinearTerm(off)])  # noqa: E731
    reg = lambda lo: [d(lo), d(lo + 1)] + [k(0)] * 6  # noqa: E731
    its = [BusInteraction.receiver(BUS_ECALL, mu, [ts_lo, ts_hi, k(ECSM_SYSCALL_NUMBER & 0xFFFFFFFF),
                                                  k(ECSM_SYSCALL_NUMBER >> 32)])]
    kbyte = lambda i: BusValue.of_linear([LinearTerm(1 << j, c.K + 8 * i + j) for j in range(8)])  # noqa: E731
    for addr, regno, off, value_of, write in ((c.ADDR_XG_0, 11, 0, lambda i: [d(c.XG + 8 * i + b) for b in range(8)],
                                               False),
                                              (c.ADDR_K_0, 12, 1, lambda i: [kbyte(8 * i + b) for b in range(8)], False),
                                              (c.ADDR_XR_0, 10, 2, lambda i: [d(c.XR + 8 * i + b) for b in range(8)],
                                               True)):
        ts = ts_plus(off) if off else ts_lo
        its.append(BusInteraction.sender(BUS_MEMW, mu, _memw_read_tuple(reg(addr), 1, k(2 * regno), k(0), ts, ts_hi,
                                                                         1, 0)))
        for i in range(4):
            base = BusValue.of_linear([LinearTerm(1, addr), LinearTerm(8 * i)])
            if write:
                its.append(BusInteraction.sender(BUS_MEMW, mu, [k(0), base, d(addr + 1)] + value_of(i)
                                                 + [ts, ts_hi, k(0), k(0), k(1)]))
            else:
                its.append(BusInteraction.sender(BUS_MEMW, mu, _memw_read_tuple(value_of(i), 0, base, d(addr + 1), ts,
                                                                                 ts_hi, 0, 1)))
    for base, n in ((c.X2, 32), (c.Q0, 32), (c.YG, 32), (c.Q1, 33)):
        its += [BusInteraction.sender(BUS_ARE_BYTES, mu, [d(base + i), k(0)]) for i in range(n)]
    for base, off in zip((c.C0, c.C1), ECSM_CARRY_OFFSETS):
        its += [BusInteraction.sender(BUS_IS_HALFWORD, mu, [BusValue.of_linear([LinearTerm(1, base + i),
                                                                                LinearTerm(off)])])
                for i in range(63)]
    for base in (c.XG_SUB_P, c.K_SUB_N, c.XR_SUB_P):
        its += [BusInteraction.sender(BUS_IS_HALFWORD, mu, [d(base + i)]) for i in range(16)]
    its.append(BusInteraction.sender(BUS_ZERO, mu, [
        BusValue.of_linear([LinearTerm(1 << (b % 8), c.K + b) for b in range(256)]), k(0)]))
    its += [BusInteraction.receiver(BUS_BIT, Multiplicity("Column", (c.K + i,)), [ts_lo, ts_hi, k(i)])
            for i in range(256)]
    its.append(BusInteraction.sender(BUS_BIT, mu, [ts_lo, ts_hi, d(c.LEN_K)]))
    its.append(BusInteraction.sender(BUS_ECDAS, mu, ecdas_tuple(
        c.XG, c.YG, c.XG, c.YG, BusValue.of_linear([LinearTerm(1, c.LEN_K), LinearTerm(-1)]), k(0), ts_lo, ts_hi)))
    its.append(BusInteraction.receiver(BUS_ECDAS, mu, ecdas_tuple(
        c.XR, c.YR, c.XG, c.YG, BusValue.of_linear([LinearTerm(-1)]), k(0), ts_lo, ts_hi)))
    return its


def ecsm_table_air() -> AirProgram:
    """ECSM as an AIR (ecsm.rs:700-902, 413 constraints): 0 mu a bit; 1-256 the
    scalar bits; 257 no bits on padding rows; 258-322 the x2 relation's carry
    identities and closing carry, 323-387 the yG relation's (with mu·p^2 and
    mu·7); 388 q1's top byte a bit; then per overflow addition (p + (xG - p)
    mod 2^256 = xG + 2^256, likewise k against N and xR against p) its seven
    inner word carries mu-gated bits and the carry-out mu-gated 1; then the
    LogUp constraints of its 579 interactions."""
    c = EcsmCols
    b = ir.IrBuilder()
    m = lambda col: b.main(0, col)  # noqa: E731
    one, zero = b.const_base(1), b.const_base(0)
    pb = _le(SECP_P, 32)
    at = lambda base, n, j: m(base + j) if j < n else zero  # noqa: E731
    _emit_is_bit(b, 0, c.MU)
    for i in range(256):
        _emit_is_bit(b, 1 + i, c.K + i)
    ksum = m(c.K)
    for i in range(1, 256):
        ksum = ksum + m(c.K + i)
    b.emit(257, ksum * (one - m(c.MU)))
    idx = 258
    for rel, cb in ((0, c.C0), (1, c.C1)):
        for i in range(64):
            s = zero
            p2 = 0
            for j in range(i + 1):
                if rel == 0:
                    if j < 32 and i - j < 32:
                        s = s + at(c.XG, 32, j) * at(c.XG, 32, i - j)
                    if j < 32 and i - j < 32 and pb[i - j]:
                        s = s - at(c.Q0, 32, j) * b.const_base(pb[i - j])
                else:
                    if j < 32 and i - j < 32:
                        s = s + at(c.YG, 32, j) * at(c.YG, 32, i - j) - at(c.X2, 32, j) * at(c.XG, 32, i - j)
                        p2 += pb[j] * pb[i - j]
                    if j < 33 and i - j < 32 and pb[i - j]:
                        s = s - at(c.Q1, 33, j) * b.const_base(pb[i - j])
            if rel == 0:
                s = s - at(c.X2, 32, i)
            else:
                s = s + m(c.MU) * b.const_base(p2)
                if i == 0:
                    s = s - m(c.MU) * b.const_base(SECP_B)
            prev = m(cb + i - 1) if i else zero
            b.emit(idx, b.const_base(256) * m(cb + i) - prev - s)
            idx += 1
        b.emit(idx, m(cb + 63))
        idx += 1
    _emit_is_bit(b, idx, c.Q1 + 32)
    idx += 1
    inv = b.const_base(INV_SHIFT_32)
    for bound, hl, base, bits in ((SECP_P, c.XG_SUB_P, c.XG, False), (SECP_N, c.K_SUB_N, c.K, True),
                                  (SECP_P, c.XR_SUB_P, c.XR, False)):
        prev = zero
        carries = []
        for i in range(8):
            addend1 = m(hl + 2 * i) + m(hl + 2 * i + 1) * b.const_base(1 << 16)
            if bits:
                s = m(base + 32 * i)
                for bit in range(1, 32):
                    s = s + m(base + 32 * i + bit) * b.const_base(1 << bit)
            else:
                s = m(base + 4 * i)
                for byte in range(1, 4):
                    s = s + m(base + 4 * i + byte) * b.const_base(1 << (8 * byte))
            ci = (b.const_base((bound >> (32 * i)) & 0xFFFFFFFF) + addend1 + prev - s) * inv
            carries.append(ci)
            prev = ci
        for ci in carries[:7]:
            b.emit(idx, m(c.MU) * ci * (one - ci))
            idx += 1
        b.emit(idx, m(c.MU) * (one - carries[7]))
        idx += 1
    its = ecsm_bus_interactions()
    n_total = emit_logup_constraints(b, its, idx)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("ecsm", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3, 1 + absorbed))


# LOCAL_TO_GLOBAL (prover/src/tables/local_to_global.rs, continuation.rs:
# 150-200): per epoch, one row per memory cell the epoch touched, with its
# init claim (the value when first touched, the epoch that last wrote it) and
# fini claim (the value at the epoch's end, its last access timestamp). Inside
# the epoch proof the table bookends the cell on the Memory bus like PAGE; in
# the global proof its init/fini claims chain the epochs on the GlobalMemory
# bus between GLOBAL_MEMORY's genesis and final tokens. Epoch labels are 1-based
# (genesis 0); init_epoch < fini_epoch is an IS_B20 check of their gap.
MAX_EPOCHS = 1 << 20


class L2gCols:
    """local_to_global.rs:210-250."""
    ADDRESS_LO, ADDRESS_HI, INIT_VALUE, INIT_EPOCH_0, INIT_EPOCH_1 = 0, 1, 2, 3, 4
    FINI_VALUE, FINI_TIMESTAMP_LO, FINI_TIMESTAMP_HI, MU = 5, 6, 7, 8
    NUM_COLUMNS = 9


@dataclass(frozen=True)
class CellBoundary:
    """local_to_global.rs:90-130: a cell's init claim (value, originating epoch,
    its timestamp) and fini claim (value, this epoch's label, last access
    timestamp)."""
    address: int
    init_value: int
    init_epoch: int
    init_timestamp: int
    fini_value: int
    fini_epoch: int
    fini_timestamp: int


def epoch_boundaries(initial_memory: dict, epochs) -> list:
    """local_to_global.rs:140-205: epochs[e] lists (address, end value, end
    timestamp) of the cells epoch e touched; each cell's init comes from its
    last writer (genesis: epoch 0, timestamp 0, the program-start value or 0),
    its fini names epoch label e + 1."""
    prov = {a: (GENESIS_EPOCH, v, 0) for a, v in initial_memory.items()}
    out = []
    for e, touched in enumerate(epochs):
        label = e + 1
        rows = []
        for addr, val, ts in touched:
            oe, ov, ot = prov.get(addr, (GENESIS_EPOCH, 0, 0))
            rows.append(CellBoundary(addr, ov, oe, ot, val, label, ts))
            prov[addr] = (label, val, ts)
        out.append(rows)
    return out


def generate_l2g_trace(boundaries) -> np.ndarray:
    """local_to_global.rs:267-292: one row per boundary (mu = 1), zero rows to
    a power of two (at least 1)."""
    c = L2gCols
    n = 1 << max(0, (len(boundaries) - 1).bit_length()) if boundaries else 1
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)
    for i, bd in enumerate(boundaries):
        if bd.init_epoch >= 1 << 32:
            raise ValueError("epoch label exceeds 32 bits")
        t[i] = (bd.address & 0xFFFFFFFF, bd.address >> 32, bd.init_value & 0xFF, bd.init_epoch & 0xFFFF,
                (bd.init_epoch >> 16) & 0xFFFF, bd.fini_value & 0xFF, bd.fini_timestamp & 0xFFFFFFFF,
                bd.fini_timestamp >> 32, 1)
    return t


def l2g_global_bus_interactions(epoch_label: int):
    """local_to_global.rs:339-365: GlobalMemory[address, init value, init epoch]
    received and [address, fini value, epoch_label] sent (mu)."""
    c = L2gCols
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    mu = Multiplicity("Column", (c.MU,))
    word = BusValue.of_linear([LinearTerm(1, c.INIT_EPOCH_0), LinearTerm(1 << 16, c.INIT_EPOCH_1)])
    return [BusInteraction.receiver(BUS_GLOBAL_MEMORY, mu, [d(c.ADDRESS_LO), d(c.ADDRESS_HI), d(c.INIT_VALUE), word]),
            BusInteraction.sender(BUS_GLOBAL_MEMORY, mu, [d(c.ADDRESS_LO), d(c.ADDRESS_HI), d(c.FINI_VALUE),
                                                          BusValue.of_linear([LinearTerm(epoch_label)])])]


def l2g_memory_bus_interactions(epoch_label: int):
    """local_to_global.rs:367-440 (continuation.rs:186-200 order): the Memory
    token [0, address, 0, 0, init value] received and [0, address, fini
    timestamp, fini value] sent; ARE_BYTES[init value, fini value], IS_HALFWORD
    of both init-epoch halves and IS_B20[epoch_label - 1 - init_epoch] (all mu)."""
    if epoch_label < 1:
        raise ValueError("epoch_label must be a 1-based fini epoch")
    c = L2gCols
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    k = lambda v: BusValue.of_linear([LinearTerm(v)])  # noqa: E731
    mu = Multiplicity("Column", (c.MU,))
    return [
        BusInteraction.receiver(BUS_MEMORY, mu, [k(0), d(c.ADDRESS_LO), d(c.ADDRESS_HI), k(0), k(0), d(c.INIT_VALUE)]),
        BusInteraction.sender(BUS_MEMORY, mu, [k(0), d(c.ADDRESS_LO), d(c.ADDRESS_HI), d(c.FINI_TIMESTAMP_LO),
                                               d(c.FINI_TIMESTAMP_HI), d(c.FINI_VALUE)]),
        BusInteraction.sender(BUS_ARE_BYTES, mu, [d(c.INIT_VALUE), d(c.FINI_VALUE)]),
        BusInteraction.sender(BUS_IS_HALFWORD, mu, [d(c.INIT_EPOCH_0)]),
        BusInteraction.sender(BUS_IS_HALFWORD, mu, [d(c.INIT_EPOCH_1)]),
        BusInteraction.sender(BUS_IS_B20, mu, [BusValue.of_linear([
            LinearTerm(epoch_label - 1), LinearTerm(-1, c.INIT_EPOCH_0), LinearTerm(-(1 << 16), c.INIT_EPOCH_1)])]),
    ]


def l2g_memory_air(epoch_label: int) -> AirProgram:
    """The epoch-local L2G AIR (continuation.rs:150-200): mu a bit, then the
    LogUp constraints of its six interactions."""
    c = L2gCols
    b = ir.IrBuilder()
    _emit_is_bit(b, 0, c.MU)
    its = l2g_memory_bus_interactions(epoch_label)
    n_total = emit_logup_constraints(b, its, 1)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("l2g_memory", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3 if pairs else 0, 1 + absorbed, 2))


def l2g_global_air(epoch_label: int) -> AirProgram:
    """The global-proof L2G AIR (continuation.rs:159-183): no constraints of
    its own, the LogUp constraints of its two GlobalMemory interactions."""
    return _bus_only_air("l2g_global", L2gCols.NUM_COLUMNS, l2g_global_bus_interactions(epoch_label), 0, None)


# COPY (prover/src/tables/copy.rs): the streaming copy behind memcpy/memmove,
# memset and write(2). A call is a chain of rows of 8 bytes (then 1-byte rows
# for the last seven) linked on the CopyNext bus; every row of a call shares
# the call's timestamp T, reads at T + is_set and writes at T + 1 - is_set (so
# copies are snapshots and memset's self-overlapping copy propagates its
# seed), and the read/write multiplicity mu - borrow is a linear form of the
# count columns (the borrow out of count - step marks the zero-length call).
BUS_COMMIT, BUS_COPY_NEXT = 21, 33          # types.rs:316, :388
WRITE_SYSCALL_NUMBER = 64
DMA_MEMCPY_SYSCALL_NUMBER = 2**64 - 1 - 29  # execution.rs:72
DMA_MEMSET_SYSCALL_NUMBER = 2**64 - 1 - 30  # execution.rs:143
DMA_MEMSET_GAP = 8                          # execution.rs:159


class CopyCols:
    """copy.rs:150-213."""
    TIMESTAMP_0, TIMESTAMP_1 = 0, 1
    VALUE = tuple(range(2, 10))
    SRC_0, SRC_1, SRC_INCR = 10, 11, 12
    DST_0, DST_1, DST_INCR = 16, 17, 18
    COUNT_0, COUNT_1, COUNT_DECR = 22, 23, 24
    DST_RES = 28
    FIRST, END, SINGLE, IS_SET, IS_WRITE, FIRST_WRITE, MU, MU_COM, MU_COM_WIDE = range(32, 41)
    NUM_COLUMNS = 41


@dataclass(frozen=True)
class CopyOperation:
    """copy.rs:226-260: one row (functionality: copy, set or write)."""
    timestamp: int
    src: int
    dst: int
    count: int
    width: int
    first: bool
    end: bool
    functionality: str
    value: tuple

    def dst_res(self) -> int:
        return (self.dst + self.count) & (2**64 - 1) if self.functionality == "write" else self.dst


def copy_call_rows(functionality: str, timestamp: int, src: int, dst: int, count: int, memory: dict) -> list:
    """trace_builder.rs:1155-1255 collect_copy_ops over a byte memory (address
    -> byte, updated): 8-byte rows while 8 or more bytes remain, then 1-byte
    rows (execution.rs:101-107); a copy reads every row before its writes land
    (snapshot), a set writes each row before the next reads (propagation), a
    write leaves memory unchanged; a zero-length call is one row with count 0,
    width 1. The last row ends the chain."""
    M = 2**64 - 1
    rows, writes, off, rem = [], [], 0, count
    while rem:
        width = 1 if rem < 8 else 8
        s, d = (src + off) & M, (dst + off) & M
        vals = [memory.get((s + i) & M, 0) for i in range(width)]
        if functionality == "set":
            for i, v in enumerate(vals):
                memory[(d + i) & M] = v
        elif functionality == "copy":
            writes.append((d, vals))
        rows.append(CopyOperation(timestamp, s, d, rem, width, off == 0, False, functionality,
                                  tuple(vals + [0] * (8 - width))))
        off += width
        rem -= width
    for d, vals in writes:
        for i, v in enumerate(vals):
            memory[(d + i) & M] = v
    if rows:
        last = rows[-1]
        rows[-1] = CopyOperation(last.timestamp, last.src, last.dst, last.count, last.width, last.first, True,
                                 last.functionality, last.value)
    else:
        rows.append(CopyOperation(timestamp, src, dst, 0, 1, True, True, functionality, (0,) * 8))
    return rows


def generate_copy_trace(ops) -> np.ndarray:
    """copy.rs:262-320: one row per operation; padding rows carry count =
    src_incr = dst_incr = 8 (every relation and borrow at zero); zero rows to a
    power of two (at least 4)."""
    c = CopyCols
    M = 2**64 - 1
    n = max(4, 1 << max(0, (len(ops) - 1).bit_length()))
    t = np.zeros((n, c.NUM_COLUMNS), dtype=np.uint64)
    wl = lambda r, col, v: r.__setitem__(slice(col, col + 2), [v & 0xFFFFFFFF, (v >> 32) & 0xFFFFFFFF])  # noqa
    hl = lambda r, col, v: r.__setitem__(slice(col, col + 4), [(v >> (16 * k)) & 0xFFFF for k in range(4)])  # noqa
    for i, op in enumerate(ops):
        r = t[i]
        wl(r, c.TIMESTAMP_0, op.timestamp)
        r[c.VALUE[0]:c.VALUE[-1] + 1] = op.value
        wl(r, c.SRC_0, op.src)
        hl(r, c.SRC_INCR, (op.src + op.width) & M)
        wl(r, c.DST_0, op.dst)
        hl(r, c.DST_INCR, (op.dst + op.width) & M)
        wl(r, c.COUNT_0, op.count)
        hl(r, c.COUNT_DECR, (op.count - op.width) & M)
        hl(r, c.DST_RES, op.dst_res())
        is_set, is_write = op.functionality == "set", op.functionality == "write"
        r[c.FIRST], r[c.END], r[c.SINGLE], r[c.MU] = int(op.first), int(op.end), int(op.width == 1), 1
        r[c.IS_SET], r[c.IS_WRITE], r[c.FIRST_WRITE] = int(is_set), int(is_write), int(op.first and is_write)
        r[c.MU_COM] = int(op.count != 0 and is_write)
        r[c.MU_COM_WIDE] = int(op.count != 0 and is_write and op.width == 8)
    t[len(ops):, c.COUNT_0] = 8
    t[len(ops):, c.SRC_INCR] = 8
    t[len(ops):, c.DST_INCR] = 8
    return t


def _copy_borrow_terms():
    """copy.rs:355-400: borrow_1 of count - step as a linear form:
    2^-32 decr2 + 2^-16 decr3 - 2^-32 count1 + 2^-64 decr0 + 2^-48 decr1 -
    2^-64 count0 - 7·2^-64 single + 8·2^-64."""
    c = CopyCols
    inv = lambda k: pow(2, P - 1 - k, P)  # noqa: E731
    return [LinearTerm(inv(32), c.COUNT_DECR + 2, True), LinearTerm(inv(16), c.COUNT_DECR + 3, True),
            LinearTerm(P - inv(32), c.COUNT_1, True), LinearTerm(inv(64), c.COUNT_DECR, True),
            LinearTerm(inv(48), c.COUNT_DECR + 1, True), LinearTerm(P - inv(64), c.COUNT_0, True),
            LinearTerm(P - 7 * inv(64) % P, c.SINGLE, True), LinearTerm(-(1 << 35))]


def _neg_terms(terms):
    return [LinearTerm(-t.coefficient, t.column) if t.column is None or not t.unsigned
            else LinearTerm((P - t.coefficient) % P, t.column, True) for t in terms]


def copy_bus_interactions():
    """copy.rs:505-775, in its order (33): Ecall [ts, number] with the number a
    linear function of is_set/is_write (first); the register reads of x11
    (src) and x12 (count), the read-write of x10 (x254 for write) from dst to
    dst_res (first), and for write the x10 read-write from the fd 1 to count
    (first_write); CopyNext [ts, src + step, dst + step, count - step, is_set,
    is_write] sent (mu - end) and [ts, src, dst, count, is_set, is_write]
    received (mu - first); IS_HALFWORD of the increments (mu) and of dst_res
    (is_write); the 8-byte MEMW read of src (at ts + is_set) and write of dst
    (at ts + 1 - is_set) with multiplicities mu - borrow and mu - borrow -
    mu_com; the Commit pairs [dst + lane, value_lane] (mu_com for lane 0,
    mu_com_wide for lanes 1-7)."""
    c = CopyCols
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    k = lambda v: BusValue.of_linear([LinearTerm(v)])  # noqa: E731
    col = lambda cc: Multiplicity("Column", (cc,))  # noqa: E731
    ts = [d(c.TIMESTAMP_0), d(c.TIMESTAMP_1)]
    lo = lambda v: v & 0xFFFFFFFF  # noqa: E731
    hi = lambda v: v >> 32  # noqa: E731
    num = lambda f: BusValue.of_linear([  # noqa: E731
        LinearTerm(f(DMA_MEMCPY_SYSCALL_NUMBER)),
        LinearTerm(f(DMA_MEMSET_SYSCALL_NUMBER) - f(DMA_MEMCPY_SYSCALL_NUMBER), c.IS_SET),
        LinearTerm(f(WRITE_SYSCALL_NUMBER) - f(DMA_MEMCPY_SYSCALL_NUMBER), c.IS_WRITE)])

    def reg(address, old, value):
        return old + [k(0)] * 6 + [k(1), address, k(0)] + value + [k(0)] * 6 + ts + [k(1), k(0), k(0)]

    w8 = BusValue.of_linear([LinearTerm(1), LinearTerm(-1, c.SINGLE)])
    w2l = lambda s: [BusValue.packed(Packing.WORD2L, s), BusValue.packed(Packing.WORD2L, s + 2)]  # noqa: E731
    its = [BusInteraction.receiver(BUS_ECALL, col(c.FIRST), ts + [num(lo), num(hi)]),
           BusInteraction.sender(BUS_MEMW, col(c.FIRST), reg(k(22), [d(c.SRC_0), d(c.SRC_1)], [d(c.SRC_0), d(c.SRC_1)])),
           BusInteraction.sender(BUS_MEMW, col(c.FIRST), reg(k(24), [d(c.COUNT_0), d(c.COUNT_1)],
                                                             [d(c.COUNT_0), d(c.COUNT_1)])),
           BusInteraction.sender(BUS_MEMW, col(c.FIRST), reg(BusValue.of_linear([LinearTerm(20), LinearTerm(488, c.IS_WRITE)]),
                                                             [d(c.DST_0), d(c.DST_1)], w2l(c.DST_RES))),
           BusInteraction.sender(BUS_MEMW, col(c.FIRST_WRITE), reg(k(20), [k(1), k(0)], [d(c.COUNT_0), d(c.COUNT_1)])),
           BusInteraction.sender(BUS_COPY_NEXT, Multiplicity("Diff", (c.MU, c.END)), ts + [
               BusValue.packed(Packing.DWORD_HL, c.SRC_INCR), BusValue.packed(Packing.DWORD_HL, c.DST_INCR),
               BusValue.packed(Packing.DWORD_HL, c.COUNT_DECR), d(c.IS_SET), d(c.IS_WRITE)]),
           BusInteraction.receiver(BUS_COPY_NEXT, Multiplicity("Diff", (c.MU, c.FIRST)), ts + [
               BusValue.packed(Packing.DWORD_WL, c.SRC_0), BusValue.packed(Packing.DWORD_WL, c.DST_0),
               BusValue.packed(Packing.DWORD_WL, c.COUNT_0), d(c.IS_SET), d(c.IS_WRITE)])]
    for start in (c.SRC_INCR, c.DST_INCR, c.COUNT_DECR):
        its += [BusInteraction.sender(BUS_IS_HALFWORD, col(c.MU), [d(start + j)]) for j in range(4)]
    its += [BusInteraction.sender(BUS_IS_HALFWORD, col(c.IS_WRITE), [d(c.DST_RES + j)]) for j in range(4)]
    active = [LinearTerm(1, c.MU)] + _neg_terms(_copy_borrow_terms())
    vals = [d(v) for v in c.VALUE]
    tso = lambda off, s: BusValue.of_linear([LinearTerm(1, c.TIMESTAMP_0), LinearTerm(s, c.IS_SET), LinearTerm(off)])  # noqa
    its.append(BusInteraction.sender(BUS_MEMW, Multiplicity("Linear", linear=tuple(active)),
                                     vals + [k(0), d(c.SRC_0), d(c.SRC_1)] + vals
                                     + [tso(0, 1), d(c.TIMESTAMP_1), k(0), k(0), w8]))
    its.append(BusInteraction.sender(BUS_MEMW, Multiplicity("Linear", linear=tuple(active + [LinearTerm(-1, c.MU_COM)])),
                                     [k(0), d(c.DST_0), d(c.DST_1)] + vals
                                     + [tso(1, -1), d(c.TIMESTAMP_1), k(0), k(0), w8]))
    for lane in range(8):
        its.append(BusInteraction.sender(BUS_COMMIT, col(c.MU_COM if lane == 0 else c.MU_COM_WIDE), [
            BusValue.of_linear([LinearTerm(1, c.DST_0), LinearTerm(lane)]), d(c.VALUE[lane])]))
    return its


def copy_table_air() -> AirProgram:
    """COPY as an AIR (copy.rs:780-1060, 40 constraints, degree 3): 0-5 bits;
    6 first/end only on real rows; 7 not both set and write; 8 a functionality
    only on real rows; 9 first_write = first·is_write; 10-15 the carry pairs of
    src + step = src_incr, dst + step = dst_incr, count_decr + step = count
    (step = 8 - 7 single); 16-17 of dst + count = dst_res on write rows; 18-19
    dst_res = dst otherwise; 20-21 of src + 8 = dst on set rows; 22-28 a
    single-byte row carries only value[0]; 29-34 the borrow of count - step
    only on a first row with count 0, and end exactly on the row moving the
    last bytes (count = step) or on the zero-length row; 35-36 mu_com and
    mu_com_wide; 37-39 no carry out of src + step, dst + step (while moving
    bytes) and src + 8 (set rows); then the LogUp constraints of its 33
    interactions."""
    c = CopyCols
    b = ir.IrBuilder()
    m = lambda col: b.main(0, col)  # noqa: E731
    one, zero, inv, s16 = b.const_base(1), b.const_base(0), b.const_base(INV_SHIFT_32), b.const_base(1 << 16)
    for i, col in enumerate((c.FIRST, c.END, c.SINGLE, c.MU, c.IS_SET, c.IS_WRITE)):
        _emit_is_bit(b, i, col)
    first, end, mu, single, is_set, is_write = (m(x) for x in (c.FIRST, c.END, c.MU, c.SINGLE, c.IS_SET, c.IS_WRITE))
    b.emit(6, (first + end) * (one - mu))
    b.emit(7, is_set * is_write)
    b.emit(8, (is_set + is_write) * (one - mu))
    b.emit(9, m(c.FIRST_WRITE) - first * is_write)
    step = (b.const_base(8) - b.const_base(7) * single, zero)
    wl = lambda col: (m(col), m(col + 1))  # noqa: E731
    hl = lambda col: (m(col) + m(col + 1) * s16, m(col + 2) + m(col + 3) * s16)  # noqa: E731

    def carries(lhs, rhs, out):
        c0 = (lhs[0] + rhs[0] - out[0]) * inv
        return c0, (lhs[1] + rhs[1] + c0 - out[1]) * inv

    def pair(idx, gate, lhs, rhs, out):
        c0, c1 = carries(lhs, rhs, out)
        b.emit(idx, c0 * (one - c0) if gate is None else gate * c0 * (one - c0))
        b.emit(idx + 1, c1 * (one - c1) if gate is None else gate * c1 * (one - c1))

    pair(10, None, wl(c.SRC_0), step, hl(c.SRC_INCR))
    pair(12, None, wl(c.DST_0), step, hl(c.DST_INCR))
    pair(14, None, hl(c.COUNT_DECR), step, wl(c.COUNT_0))
    pair(16, is_write, wl(c.DST_0), wl(c.COUNT_0), hl(c.DST_RES))
    dres = hl(c.DST_RES)
    b.emit(18, (one - is_write) * (dres[0] - m(c.DST_0)))
    b.emit(19, (one - is_write) * (dres[1] - m(c.DST_1)))
    gap = (b.const_base(DMA_MEMSET_GAP), zero)
    pair(20, is_set, wl(c.SRC_0), gap, wl(c.DST_0))
    for i in range(1, 8):
        b.emit(21 + i, single * m(c.VALUE[i]))
    bt = _copy_borrow_terms()
    borrow = b.const_base(bt[-1].coefficient % P)
    for t in bt[:-1]:
        borrow = borrow + b.const_base(t.coefficient) * m(t.column)
    not_borrow = one - borrow
    b.emit(29, borrow * (one - first))
    b.emit(30, borrow * m(c.COUNT_0))
    b.emit(31, borrow * m(c.COUNT_1))
    b.emit(32, end * not_borrow * m(c.COUNT_1))
    b.emit(33, end * not_borrow * (m(c.COUNT_0) - step[0]))
    b.emit(34, borrow * (one - end))
    b.emit(35, m(c.MU_COM) - (mu - borrow) * is_write)
    b.emit(36, m(c.MU_COM_WIDE) - m(c.MU_COM) * (one - single))
    active = mu - borrow
    b.emit(37, active * carries(wl(c.SRC_0), step, hl(c.SRC_INCR))[1])
    b.emit(38, active * carries(wl(c.DST_0), step, hl(c.DST_INCR))[1])
    b.emit(39, is_set * carries(wl(c.SRC_0), gap, wl(c.DST_0))[1])
    its = copy_bus_interactions()
    n_total = emit_logup_constraints(b, its, 40)
    prog = b.finish(c.NUM_COLUMNS)
    pairs, absorbed = split_interactions(len(its))
    return AirProgram("copy", c.NUM_COLUMNS, pairs + 1, prog, ir.lower(prog, pin_roots=False), [0] * n_total,
                      boundary=[Boundary(pairs, True, 0, (0, 0, 0))], interactions=its,
                      max_degree=max(3, 1 + absorbed))


# HALT (prover/src/tables/halt.rs): the one-row table of the sys_exit ecall.
# It receives the ecall, finalizes the general registers at timestamp 2^32 - 1
# (x1-x9 and x11-x31 written 0, x10 read with old value 0: a nonzero exit code
# leaves the MEMW bus unbalanced) and swaps the pc token the halting
# instruction wrote for the padding pc 1.
HALT_SYSCALL_NUMBER = 93
HALT_FINAL_TIMESTAMP = 2**32 - 1


class HaltCols:
    """halt.rs:50-61."""
    TIMESTAMP_0, TIMESTAMP_1, PC_0, PC_1 = 0, 1, 2, 3
    NUM_COLUMNS = 4


def generate_halt_trace(timestamp: int, next_pc: int) -> np.ndarray:
    """halt.rs:74-92: the single row [timestamp, next pc] (DWordWL each)."""
    if timestamp >= 1 << 32:
        raise ValueError("HALT timestamp exceeds 32 bits")
    return np.array([[timestamp & 0xFFFFFFFF, timestamp >> 32, next_pc & 0xFFFFFFFF, next_pc >> 32]], dtype=np.uint64)


def halt_bus_interactions():
    """halt.rs:120-273, in its order (36, all multiplicity one): the Ecall
    receiver [ts, 93, 0]; MEMW register writes of 0 to x1-x9, the read of x10
    with old value 0, writes of 0 to x11-x31, all at 2^32 - 1 (register-width
    accesses); per pc word the Memory token [1, 510 + i, 0, ts + 1, next_pc_i]
    sent and [1, 510 + i, 0, ts + 1, (1, 0)_i] received."""
    c = HaltCols
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    k = lambda v: BusValue.of_linear([LinearTerm(v)])  # noqa: E731
    ts = [k(HALT_FINAL_TIMESTAMP), k(0)]
    zero8 = [k(0)] * 8
    its = [BusInteraction.receiver(BUS_ECALL, ONE, [d(c.TIMESTAMP_0), d(c.TIMESTAMP_1), k(HALT_SYSCALL_NUMBER), k(0)])]
    for r in range(1, 32):
        head = [k(1), k(2 * r), k(0)]
        if r == 10:
            its.append(BusInteraction.sender(BUS_MEMW, ONE, zero8 + head + zero8 + ts + [k(1), k(0), k(0)]))
        else:
            its.append(BusInteraction.sender(BUS_MEMW, ONE, head + zero8 + ts + [k(1), k(0), k(0)]))
    nxt = BusValue.of_linear([LinearTerm(1, c.TIMESTAMP_0), LinearTerm(1)])
    for i in range(2):
        its.append(BusInteraction.sender(BUS_MEMORY, ONE, [k(1), k(510 + i), k(0), nxt, d(c.TIMESTAMP_1), d(c.PC_0 + i)]))
    for i, v in enumerate((CPU_PADDING_PC, 0)):
        its.append(BusInteraction.receiver(BUS_MEMORY, ONE, [k(1), k(510 + i), k(0), nxt, d(c.TIMESTAMP_1), k(v)]))
    return its


def halt_table_air() -> AirProgram:
    """HALT as an AIR: no constraints of its own, the LogUp constraints of its
    36 interactions."""
    return _bus_only_air("halt", HaltCols.NUM_COLUMNS, halt_bus_interactions(), 0, None)
