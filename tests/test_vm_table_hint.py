"""The reference VM's HINT table (prover/src/tables/hint.rs): 41 columns, one
constraint (mu a bit), 27 interactions.

Cross-table checks against other captured tables in one transcript: HINT's
three ALU LT range checks balance against the LT table's receiver fed the same
comparisons, and stop balancing for a selector the executor rejects (the LT row
then says `false` where HINT asserts `true`); HINT's three MEMW register reads
(x12, x10, x11 at the ecall timestamp) balance against the MEMW_REGISTER
table's read receiver for the same accesses, and stop balancing when one
register value differs. A proof of HINT beside LT replays and verifies."""
import numpy as np
import pytest

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from oracle.air_proof import replay_air_tables, verify_air_tables
from test_vm_table_eq import _constraint_values

P = 0xFFFFFFFF00000001


def hint_ops(k, seed, selector=None):
    rng = np.random.default_rng(seed)
    ops = []
    for i in range(k):
        ops.append(V.HintOperation(timestamp=int(rng.integers(1 << 20, 1 << 29)) * 4 + 2,
                                   out_addr=int(rng.integers(0, 2**40)) & ~7,
                                   out_bytes=bytes(rng.integers(0, 256, 32, dtype=np.uint8)),
                                   hint_id=int(rng.integers(0, 3)) if selector is None else selector,
                                   in_addr=int(rng.integers(0, 2**44))))
    return ops


def _desc(a):
    a.bus_descriptor = A.descriptor(a)
    return a


def _total(tables, label):
    out = replay_air_tables(tables, 2, 3, 2, 4, 0, label)
    tot = [0, 0, 0]
    for o in out:
        tot = [(x + int(y)) % P for x, y in zip(tot, o["table_contribution"])]
    return tuple(tot)


def test_hint_rows_and_constraint(ora):
    ops = hint_ops(5, 1)
    tr = V.generate_hint_trace(ops)
    c = V.HintCols
    assert tr.shape == (8, 41)
    op = ops[2]
    assert (int(tr[2, c.TIMESTAMP_0]), int(tr[2, c.TIMESTAMP_1])) == (op.timestamp, 0)
    assert int(tr[2, c.ADDR_OUT_0]) | int(tr[2, c.ADDR_OUT_1]) << 32 == op.out_addr
    assert bytes(tr[2, c.OUT[0]:c.OUT[-1] + 1].astype(np.uint8)) == op.out_bytes
    assert list(tr[:, c.MU]) == [1] * 5 + [0] * 3
    air = V.hint_table_air()
    assert len(air.interactions) == 27 and air.num_main == 41
    assert not np.any(_constraint_values(ora, air, tr) % np.uint64(P))
    bad = tr.copy()
    bad[1, c.MU] = 2
    ev = _constraint_values(ora, air, bad) % np.uint64(P)  # (constraint, row, Fp3 coefficient)
    assert np.any(ev[0, 1])
    with pytest.raises(ValueError):
        V.generate_hint_trace([V.HintOperation(1 << 32, 0, bytes(32), 0, 0)])


def _lt_ops(ops):
    out = []
    for op in ops:
        out += [V.LtOperation(op.hint_id, V.HINT_SELECTOR_BOUND, False, False),
                V.LtOperation(op.in_addr & 0xFFFFFFFF, V.HINT_ADDR_LIMB_BOUND, False, False),
                V.LtOperation(op.out_addr & 0xFFFFFFFF, V.HINT_ADDR_LIMB_BOUND, False, False)]
    return out


def test_hint_range_checks_balance_against_lt():
    ops = hint_ops(6, 2)
    hint_lt = _desc(A.logup_air(V.HintCols.NUM_COLUMNS, V.hint_bus_interactions()[4:7]))
    lt_recv = _desc(A.logup_air(V.LtCols.NUM_COLUMNS, V.lt_bus_interactions()[-1:]))
    assert _total([(hint_lt, V.generate_hint_trace(ops)), (lt_recv, V.generate_lt_trace(_lt_ops(ops)))],
                  b"hint-lt") == (0, 0, 0)
    bad = hint_ops(6, 2, selector=3)  # the executor rejects selector 3: LT says 3 < 3 is false
    assert _total([(hint_lt, V.generate_hint_trace(bad)), (lt_recv, V.generate_lt_trace(_lt_ops(bad)))],
                  b"hint-lt") != (0, 0, 0)


def _register_reads(ops, tamper=None):
    acc = []
    for i, op in enumerate(ops):
        for reg, val in ((12, op.out_addr), (10, op.hint_id), (11, op.in_addr)):
            if tamper == (i, reg):
                val ^= 1 << 33
            acc.append(V.RegisterAccess(2 * reg, op.timestamp, val, val, op.timestamp - 3, True))
    return acc


def test_hint_register_reads_balance_against_memw_register():
    ops = hint_ops(5, 3)
    hint_reads = _desc(A.logup_air(V.HintCols.NUM_COLUMNS, V.hint_bus_interactions()[1:4]))
    memw_read = _desc(A.logup_air(V.MemwRegisterCols.NUM_COLUMNS, V.memw_register_bus_interactions()[5:6]))
    good = V.generate_memw_register_trace(_register_reads(ops))
    assert _total([(hint_reads, V.generate_hint_trace(ops)), (memw_read, good)], b"hint-memw") == (0, 0, 0)
    bad = V.generate_memw_register_trace(_register_reads(ops, tamper=(2, 11)))
    assert _total([(hint_reads, V.generate_hint_trace(ops)), (memw_read, bad)], b"hint-memw") != (0, 0, 0)


def test_hint_proof_verifies():
    ops = hint_ops(7, 4)
    hint, lt = _desc(V.hint_table_air()), _desc(V.lt_table_air())
    tables = [(hint, V.generate_hint_trace(ops)), (lt, V.generate_lt_trace(_lt_ops(ops)))]
    proofs = replay_air_tables(tables, 4, 3, 2, 6, 0, b"hint")
    assert verify_air_tables([hint, lt], proofs, 4, 3, 2, 6, 0, b"hint")[0]
