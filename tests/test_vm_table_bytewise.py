"""The reference VM's BYTEWISE table (prover/src/tables/bytewise.rs) captured
as an AIR: trace generation merges duplicate operations with their count as
multiplicity and splits at 255, the table's LogUp constraints vanish on every
row of a generated trace, a wrong result byte changes the table's bus
contribution, and the oracle's replay of a full proof is accepted by the
verifier restatement, alone and beside the EQ table in one transcript."""
import numpy as np
import pytest

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from oracle.air_proof import replay_air_tables, verify_air_tables
from test_vm_table_eq import _constraint_values, _ops as _eq_ops

P = 0xFFFFFFFF00000001


def _ops(k, seed):
    rng = np.random.default_rng(seed)
    ops = [V.BytewiseOperation(int(a), int(b), int(o))
           for a, b, o in zip(rng.integers(0, 2**63, k), rng.integers(0, 2**63, k), rng.integers(0, 3, k))]
    ops += [V.BytewiseOperation(2**64 - 1, 0x0123456789ABCDEF, V.ALU_OP_AND)] * 3
    return ops


def test_trace_generation_dedups_and_splits():
    op = V.BytewiseOperation(0xFF00FF00FF00FF00, 0x0F0F0F0F0F0F0F0F, V.ALU_OP_XOR)
    other = V.BytewiseOperation(1, 2, V.ALU_OP_OR)
    tr = V.generate_bytewise_trace([op] * 300 + [other, op])
    c = V.BytewiseCols
    assert tr.shape == (4, c.NUM_COLUMNS)
    assert tr[:, c.MU].tolist() == [255, 46, 1, 0]
    res = op.a ^ op.b
    assert [int(x) for x in tr[0, list(c.RES)]] == [(res >> (8 * k)) & 0xFF for k in range(8)]
    assert int(tr[2, c.OP]) == V.ALU_OP_OR and int(tr[2, c.RES[0]]) == 3
    with pytest.raises(ValueError):
        V.BytewiseOperation(1, 2, V.ALU_OP_EQ).compute_res()


def test_constraints_vanish_on_generated_trace(ora):
    air = V.bytewise_table_air()
    tr = V.generate_bytewise_trace(_ops(13, 4))
    ev = _constraint_values(ora, air, tr)
    assert not np.any(ev % np.uint64(P))


def test_wrong_result_byte_changes_bus_contribution(ora):
    air = V.bytewise_table_air()
    air.bus_descriptor = A.descriptor(air)
    tr = V.generate_bytewise_trace(_ops(13, 4))
    good = replay_air_tables([(air, tr)], 4, 3, 2, 4, 0, b"bw")[0]["table_contribution"]
    bad_tr = tr.copy()
    bad_tr[2, V.BytewiseCols.RES[5]] ^= np.uint64(1)
    bad = replay_air_tables([(air, bad_tr)], 4, 3, 2, 4, 0, b"bw")[0]["table_contribution"]
    assert tuple(good) != tuple(bad)


def test_bytewise_and_eq_proof_verifies():
    bw = V.bytewise_table_air()
    bw.bus_descriptor = A.descriptor(bw)
    eq = V.eq_table_air()
    eq.bus_descriptor = A.descriptor(eq)
    tables = [(bw, V.generate_bytewise_trace(_ops(60, 8))), (eq, V.generate_eq_trace(_eq_ops(20, 5)))]
    proofs = replay_air_tables(tables, 4, 3, 2, 6, 0, b"vm-tables")
    ok, why = verify_air_tables([bw, eq], proofs, 4, 3, 2, 6, 0, b"vm-tables")
    assert ok, why
