"""The reference VM's LT table (prover/src/tables/lt.rs) captured as an AIR:
its six constraints (the carry pair of rhs + (lhs - rhs) = lhs, the signed /
unsigned less-than from the borrow and the sign bits, out = lt XOR invert,
two bit checks) and the LogUp constraints of its ten bus interactions vanish
on every row of a generated trace (signed and unsigned, equal, extreme and
sign-crossing operands, inverted outputs); a trace whose lt or out bit is
flipped, or whose sign bit is wrong, breaks one of the table's own
constraints; and the oracle's replay of a full proof is accepted."""
import numpy as np
import pytest

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from oracle.air_proof import replay_air_tables, verify_air_tables
from test_vm_table_eq import _constraint_values

P = 0xFFFFFFFF00000001
M = 2**64 - 1


def _ops(k, seed):
    rng = np.random.default_rng(seed)
    ops = []
    for i in range(k):
        a, b = int(rng.integers(0, 2**63)) * 2 + int(rng.integers(0, 2)), int(rng.integers(0, 2**63)) * 2 + 1
        ops.append(V.LtOperation(a, b if i % 4 else a, bool(i % 2), bool(i % 3 == 0)))
    ops += [V.LtOperation(0, M, True), V.LtOperation(0, M, False), V.LtOperation(M, 0, True),
            V.LtOperation(1 << 63, (1 << 63) - 1, True), V.LtOperation(1 << 63, (1 << 63) - 1, False, True),
            V.LtOperation(5, 5, False)]
    return ops


def test_semantics_of_the_operations():
    assert V.LtOperation(M, 0, True).compute_lt()          # -1 < 0 signed
    assert not V.LtOperation(M, 0, False).compute_lt()     # 2^64-1 > 0 unsigned
    assert V.LtOperation(3, 9, False, True).compute_out() is False


def test_constraints_vanish_on_generated_trace(ora):
    air = V.lt_table_air()
    tr = V.generate_lt_trace(_ops(23, 5))
    ev = _constraint_values(ora, air, tr)
    assert not np.any(ev % np.uint64(P))


@pytest.mark.parametrize("col", [V.LtCols.LT, V.LtCols.OUT, V.LtCols.LHS_MSB])
def test_wrong_bit_breaks_constraints(ora, col):
    air = V.lt_table_air()
    ops = _ops(23, 5)
    tr = V.generate_lt_trace(ops)
    # a row whose flipped bit matters: signed, sign bits differ
    row = next(i for i, op in enumerate(V.dedup_rows(ops)) if op[0].signed and (op[0].lhs >> 63) != (op[0].rhs >> 63))
    tr[row, col] ^= np.uint64(1)
    ev = _constraint_values(ora, air, tr)
    assert np.any(ev[:6, row] % np.uint64(P))


def test_lt_proof_verifies():
    air = V.lt_table_air()
    air.bus_descriptor = A.descriptor(air)
    proofs = replay_air_tables([(air, V.generate_lt_trace(_ops(40, 9)))], 4, 3, 2, 6, 0, b"lt")
    ok, why = verify_air_tables([air], proofs, 4, 3, 2, 6, 0, b"lt")
    assert ok, why
