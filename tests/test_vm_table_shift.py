"""The reference VM's SHIFT table (prover/src/tables/shift.rs): 29 columns, 22
constraints (degree 3), 17 interactions.

Pins independent of the restatement: the out words equal Python's shifts of
the 64-bit value (SLL/SRL/SRA, shift amounts with high bits set) and, for word
shifts of 32-bit inputs extended as the CPU extends them, the low word equals
the 32-bit result. Constraints vanish on generated traces (every direction,
signedness and width, shifts by multiples of 16 included) and break for a
wrong intra-limb value, a wrong output or a sign fill on an unsigned shift.
Cross-table: every BITWISE-bound send of SHIFT (MSB16, BYTE_ALU AND, ZERO,
HWSL, ARE_BYTES, IS_HALFWORD, padding rows' ARE_BYTES[0, 0] included)
balances in one transcript against the BITWISE table's receivers fed the
lookups shift.rs collect_bitwise_from_shift lists — a check of the HWSL pairs
of both restatements against each other."""
import numpy as np
import pytest

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from oracle.air_proof import replay_air_tables, verify_air_tables
from test_vm_table_eq import _constraint_values

P = 0xFFFFFFFF00000001
M64 = 2**64 - 1


def shift_ops(k, seed):
    rng = np.random.default_rng(seed)
    ops = []
    for i in range(k):
        word = i % 5 == 4
        direction, signed = bool(i & 1), bool(i & 2) and bool(i & 1)
        v = int(rng.integers(0, 2**64, dtype=np.uint64))
        if word:
            lo = v & 0xFFFFFFFF
            v = ((lo - (1 << 32) if lo >> 31 else lo) & M64) if signed or not direction else lo
        amount = int(rng.integers(0, 2**64, dtype=np.uint64)) if i % 7 == 0 else int(rng.integers(0, 256))
        if i % 6 == 3:
            amount = (amount & ~0xF) | 0  # a multiple of 16: bit_shift = 0
        ops.append(V.ShiftOperation(v, amount, direction, signed, word))
    return ops


def test_shift_results():
    ops = shift_ops(300, 1)
    tr = V.generate_shift_trace(ops)
    c = V.ShiftCols
    for r, op in enumerate(ops):
        out = int(tr[r, c.OUT_0]) | int(tr[r, c.OUT_1]) << 32
        v, s = op.value & M64, op.shift_amount & 0xFF
        if not op.word_instr:
            sh = s % 64
            want = (v << sh) & M64 if not op.direction else \
                (v >> sh if not op.signed else ((v - 2**64 if v >> 63 else v) >> sh) & M64)
            assert out == want == op.result(), (r, op)
        else:
            assert out & 0xFFFFFFFF == op.result() & 0xFFFFFFFF, (r, op)
    assert any(tr[r, c.ZBS] and r < len(ops) for r in range(tr.shape[0]))


def test_shift_constraints(ora):
    air = V.shift_table_air()
    ops = shift_ops(120, 2)
    tr = V.generate_shift_trace(ops)
    assert (air.num_main, len(air.interactions)) == (29, 17)
    assert not np.any(_constraint_values(ora, air, tr) % np.uint64(P))
    c = V.ShiftCols
    r = next(i for i, o in enumerate(ops) if not o.direction and tr[i, c.ZBS] == 0)
    for col in (c.OUT_0, c.X[0]):  # X[0] reaches the output of every left shift
        bad = tr.copy()
        bad[r, col] += 1
        ev = _constraint_values(ora, air, bad) % np.uint64(P)
        assert np.any(ev[14:16, r])
    u = next(i for i, o in enumerate(ops) if o.direction and not o.signed)
    bad = tr.copy()
    bad[u, c.IS_NEGATIVE] = 1
    assert np.any(_constraint_values(ora, air, bad)[21, u] % np.uint64(P))


def bitwise_lookups(ops, rows):
    """shift.rs:951-1035 collect_bitwise_from_shift, plus the padding rows'
    ARE_BYTES[0, 0]."""
    out = []
    for op in ops:
        is_neg, bit_shift, zbs, x, y, ls, _ = op.aux()
        h = op.halves()
        s = op.shift_amount & 0xFF
        if op.signed:
            out.append(("msb16", h[3] & 0xFF, h[3] >> 8, 0))
        if not op.direction:
            out.append(("and", s, 15, 0))
        else:
            out.append(("and", 256 - (16 if zbs else 0) - s, 15, 0))
        out.append(("zero", bit_shift, 0, 0))
        if not zbs:
            for i in range(4):
                out.append(("hwsl", h[i] & 0xFF, h[i] >> 8, bit_shift))
            ext = 0xFFFF if is_neg else 0
            out.append(("hwsl", ext & 0xFF, ext >> 8, bit_shift))
        out.append(("and", s, 16 if op.word_instr else 48, 0))
        out.append(("are_bytes", s, (op.shift_amount >> 8) & 0xFF, 0))
        half = (op.shift_amount >> 16) & 0xFFFF
        out.append(("is_half", half & 0xFF, half >> 8, 0))
        for i in range(4):
            out.append(("is_half", h[i] & 0xFF, h[i] >> 8, 0))
    out += [("are_bytes", 0, 0, 0)] * (rows - len(ops))
    return out


@pytest.mark.slow
def test_shift_lookups_balance_against_bitwise():
    ops = shift_ops(40, 3)
    tr = V.generate_shift_trace(ops)
    its = V.shift_bus_interactions()
    sends = [it for it in its if it.bus_id != V.BUS_ALU]
    shift_air = A.logup_air(V.ShiftCols.NUM_COLUMNS, sends)
    shift_air.bus_descriptor = A.descriptor(shift_air)
    bw = A.logup_air(V.BitwiseCols.NUM_COLUMNS, V.bitwise_bus_interactions())
    bw.bus_descriptor = A.descriptor(bw)
    P_ = P

    def total(lookups):
        out = replay_air_tables([(shift_air, tr), (bw, V.generate_bitwise_trace(lookups))], 2, 3, 4, 2, 0, b"shift-bw")
        tot = [0, 0, 0]
        for o in out:
            tot = [(a + int(b)) % P_ for a, b in zip(tot, o["table_contribution"])]
        return tuple(tot)

    good = bitwise_lookups(ops, tr.shape[0])
    assert total(good) == (0, 0, 0)
    i = next(k for k, l in enumerate(good) if l[0] == "hwsl" and l[3] > 0)
    bad = list(good)
    bad[i] = ("hwsl", bad[i][1], bad[i][2], (bad[i][3] % 15) + 1)
    assert total(bad) != (0, 0, 0)


def test_shift_proof_verifies():
    air = V.shift_table_air()
    air.bus_descriptor = A.descriptor(air)
    tables = [(air, V.generate_shift_trace(shift_ops(50, 4)))]
    proofs = replay_air_tables(tables, 4, 3, 2, 6, 0, b"shift")
    assert verify_air_tables([air], proofs, 4, 3, 2, 6, 0, b"shift")[0]
