"""The reference VM's MUL table (prover/src/tables/mul.rs): 64 x 64 -> 128-bit
products, 26 columns, 8 constraints, 25 interactions.

Pins independent of the restatement: the lo/hi result halves equal Python's
big-integer product of the sign-interpreted operands (signed, unsigned and
mixed, with extreme operands), and the four virtual carries that the IS_B20
sends range-check — evaluated from the bus values' linear forms over the trace
— are the integer carries of raw_product into the result words (mul.rs
compute_carries), below 2^20. Constraints vanish on generated traces and
break for a wrong raw product, a non-bit sign flag or a negative flag on an
unsigned operand; deduplication counts lo and hi requests per operation; a
proof replays and verifies."""
import numpy as np
import pytest

from lambda_vm_amd import vm_tables as V
from oracle.air_proof import replay_air_tables, verify_air_tables
from test_vm_table_eq import _constraint_values
from lambda_vm_amd import air as A

P = 0xFFFFFFFF00000001
EXTREMES = [0, 1, 2**63 - 1, 2**63, 2**64 - 1, 0xFFFF, 0x1_0000, 0x8000_0000_0000_0001]


def mul_ops(k, seed):
    rng = np.random.default_rng(seed)
    ops = []
    for i in range(k):
        a = EXTREMES[i % len(EXTREMES)] if i % 3 == 0 else int(rng.integers(0, 2**64, dtype=np.uint64))
        b = EXTREMES[(i // 3) % len(EXTREMES)] if i % 4 == 0 else int(rng.integers(0, 2**64, dtype=np.uint64))
        ops.append((V.MulOperation(a, bool(i & 1), b, bool(i & 2)), bool(i % 5 == 0)))
    return ops


def _signed(v, s):
    return v - 2**64 if s and v >> 63 else v


def test_products_and_carries():
    ops = mul_ops(60, 1)
    tr = V.generate_mul_trace(ops)
    c = V.MulCols
    rows = V.dedup_mul_rows(ops)
    its = V.mul_bus_interactions()
    b20 = [it for it in its if it.bus_id == V.BUS_IS_B20]
    assert len(its) == 25 and len(b20) == 4
    for r, (op, mu_lo, mu_hi) in enumerate(rows):
        want = (_signed(op.lhs, op.lhs_signed) * _signed(op.rhs, op.rhs_signed)) % 2**128
        lo = sum(int(tr[r, col]) << (16 * i) for i, col in enumerate(c.LO))
        hi = sum(int(tr[r, col]) << (16 * i) for i, col in enumerate(c.HI))
        assert lo | hi << 64 == want
        # virtual carries from the IS_B20 linear forms
        res = [want >> (32 * w) & 0xFFFFFFFF for w in range(4)]
        carry = 0
        for i, it in enumerate(b20):
            v = sum(t.coefficient * int(tr[r, t.column]) for t in it.values[0].linear) % P
            carry = (int(tr[r, c.RAW_PRODUCT[i]]) + carry - res[i]) >> 32
            assert v == carry and v < 1 << 20, (r, i)
    assert sum(m for _, m, _ in rows) == sum(1 for _, h in ops if not h)
    assert sum(m for _, _, m in rows) == sum(1 for _, h in ops if h)


def test_mul_constraints(ora):
    air = V.mul_table_air()
    tr = V.generate_mul_trace(mul_ops(40, 2))
    assert not np.any(_constraint_values(ora, air, tr) % np.uint64(P))
    c = V.MulCols
    for col, delta, idx in ((c.RAW_PRODUCT[2], 1, 6), (c.LHS_SIGNED, 2, 0)):
        bad = tr.copy()
        bad[3, col] += delta
        assert np.any(_constraint_values(ora, air, bad)[idx, 3] % np.uint64(P))
    bad = tr.copy()
    row = next(r for r in range(tr.shape[0]) if tr[r, c.LHS_SIGNED] == 0 and tr[r, c.MU_LO] + tr[r, c.MU_HI])
    bad[row, c.LHS_IS_NEGATIVE] = 1
    assert np.any(_constraint_values(ora, air, bad)[2, row] % np.uint64(P))


def test_mul_proof_verifies():
    air = V.mul_table_air()
    air.bus_descriptor = A.descriptor(air)
    tables = [(air, V.generate_mul_trace(mul_ops(50, 3)))]
    proofs = replay_air_tables(tables, 4, 3, 2, 6, 0, b"mul")
    assert verify_air_tables([air], proofs, 4, 3, 2, 6, 0, b"mul")[0]
