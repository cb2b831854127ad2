"""The reference VM's IS_WH table (prover/src/tables/is_wh.rs) captured as an
AIR: rows are the distinct (word, half) pairs sorted by key with their send
counts, split at 2^20 - 1; the table's LogUp constraints vanish on every row of
a generated trace; and the oracle's replay of a full proof of IS_WH beside
BYTEWISE and EQ in one transcript is accepted by the verifier restatement."""
import numpy as np

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from oracle.air_proof import replay_air_tables, verify_air_tables
from test_vm_table_bytewise import _ops as _bw_ops
from test_vm_table_eq import _constraint_values, _ops as _eq_ops

P = 0xFFFFFFFF00000001


def _values(k, seed):
    rng = np.random.default_rng(seed)
    base = [int(x) << 16 for x in rng.integers(0, 2**47, k)]
    return base + base[: k // 3]  # repeated pairs


def test_rows_sorted_counted_and_split():
    a, b = 0x0000_0001_0002_0000, 0x0000_0000_0007_0000
    rows = V.dedup_is_wh_rows([V.is_wh_key(v) for v in [a] * (V.IS_WH_MU_MAX + 3) + [b, b]])
    assert rows == [((0, 7), 2), ((1, 2), V.IS_WH_MU_MAX), ((1, 2), 3)]
    tr = V.generate_is_wh_trace(rows)
    c = V.IsWhCols
    assert tr.shape == (4, c.NUM_COLUMNS)
    assert tr[:, c.MU].tolist() == [2, V.IS_WH_MU_MAX, 3, 0]
    assert [int(x) for x in tr[1, :3]] == [2, 1, 0]
    word = 0xDEADBEEF
    key = V.is_wh_key((word << 32) | (0xCAFE << 16) | 0x1234)
    assert key == (word, 0xCAFE)
    row = V.generate_is_wh_trace([(key, 1)])[0]
    assert int(row[c.VALUE_1]) + (int(row[c.VALUE_2]) << 16) == word and int(row[c.VALUE_0]) == 0xCAFE


def test_constraints_vanish_on_generated_trace(ora):
    air = V.is_wh_table_air()
    tr = V.generate_is_wh_trace(V.dedup_is_wh_rows([V.is_wh_key(v) for v in _values(20, 2)]))
    ev = _constraint_values(ora, air, tr)
    assert not np.any(ev % np.uint64(P))


def test_three_vm_tables_proof_verifies():
    airs = [V.is_wh_table_air(), V.bytewise_table_air(), V.eq_table_air()]
    for a in airs:
        a.bus_descriptor = A.descriptor(a)
    trs = [V.generate_is_wh_trace(V.dedup_is_wh_rows([V.is_wh_key(v) for v in _values(40, 3)])),
           V.generate_bytewise_trace(_bw_ops(30, 6)), V.generate_eq_trace(_eq_ops(25, 7))]
    proofs = replay_air_tables(list(zip(airs, trs)), 4, 3, 2, 6, 0, b"vm3")
    ok, why = verify_air_tables(airs, proofs, 4, 3, 2, 6, 0, b"vm3")
    assert ok, why
