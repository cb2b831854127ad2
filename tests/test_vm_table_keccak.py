"""The reference VM's KECCAK core table (prover/src/tables/keccak.rs): 486
columns, 52 constraints (per lane the gated ADD carry pair of addr + 8·lane =
state_ptr[lane], the top lane's no-overflow, mu a bit), 84 interactions.

Constraints vanish on generated traces (addresses whose lane pointers carry
across the 2^16 and 2^32 limb boundaries included) and break for a wrong lane
pointer or a non-bit mu. Cross-table checks in one transcript: the 25 IS_WH
sends per call balance against the IS_WH table built from the same keys, and
the x10 read balances against MEMW_REGISTER's read receiver (both stop
balancing when one value differs). A proof of KECCAK beside IS_WH replays and
verifies."""
import numpy as np
import pytest

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from oracle.air_proof import replay_air_tables, verify_air_tables
from test_vm_table_eq import _constraint_values

P = 0xFFFFFFFF00000001


def keccak_ops(k, seed):
    rng = np.random.default_rng(seed)
    ops = []
    for i in range(k):
        if i % 3 == 0:
            addr = (int(rng.integers(1, 2**31)) << 32) - 8 * int(rng.integers(1, 25))  # lanes cross 2^32
        elif i % 3 == 1:
            addr = (int(rng.integers(0, 2**40)) << 16) - 8 * int(rng.integers(1, 25))  # lanes cross 2^16
        else:
            addr = int(rng.integers(0, 2**60)) & ~7
        ops.append(V.KeccakOperation(timestamp=int(rng.integers(1, 2**31)) * 4 + 2, state_addr=addr,
                                     input=tuple(int(v) for v in rng.integers(0, 2**64, 25, dtype=np.uint64)),
                                     output=tuple(int(v) for v in rng.integers(0, 2**64, 25, dtype=np.uint64))))
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


def test_keccak_rows(ora):
    ops = keccak_ops(5, 1)
    tr = V.generate_keccak_trace(ops)
    c = V.KeccakCols
    assert tr.shape == (8, 486)
    op = ops[3]
    assert int.from_bytes(bytes(tr[3, c.ADDR:c.ADDR + 8].astype(np.uint8)), "little") == op.state_addr
    lane = 7
    assert int.from_bytes(bytes(tr[3, c.INPUT_STATE + 8 * lane:c.INPUT_STATE + 8 * lane + 8].astype(np.uint8)),
                          "little") == op.input[lane]
    ptr = op.state_addr + 8 * lane
    assert [int(tr[3, c.state_ptr(lane, i)]) for i in range(3)] == [ptr & 0xFFFF, (ptr >> 16) & 0xFFFF, ptr >> 32]
    assert [int(tr[6, c.state_ptr(l, 0)]) for l in range(25)] == [8 * l for l in range(25)] and tr[6, c.MU] == 0
    air = V.keccak_table_air()
    assert (air.num_main, len(air.interactions)) == (486, 84)
    assert not np.any(_constraint_values(ora, air, tr) % np.uint64(P))
    bad = tr.copy()
    bad[0, c.state_ptr(4, 1)] += 1  # a lane pointer that is not addr + 32
    ev = _constraint_values(ora, air, bad) % np.uint64(P)
    assert np.any(ev[8:10, 0])
    bad = tr.copy()
    bad[2, c.MU] = 2
    assert np.any(_constraint_values(ora, air, bad)[51, 2] % np.uint64(P))
    with pytest.raises(ValueError):
        V.generate_keccak_trace([V.KeccakOperation(1, 2**64 - 64, (0,) * 25, (0,) * 25)])


def test_keccak_lane_ranges_balance_against_is_wh():
    ops = keccak_ops(6, 2)
    its = V.keccak_bus_interactions()
    kec = _desc(A.logup_air(V.KeccakCols.NUM_COLUMNS, [its[4 + 2 * l] for l in range(25)]))
    iswh = _desc(A.logup_air(V.IsWhCols.NUM_COLUMNS, V.is_wh_bus_interactions()[-1:]))
    good = V.generate_is_wh_trace(V.dedup_is_wh_rows(V.keccak_is_wh_keys(ops)))
    assert _total([(kec, V.generate_keccak_trace(ops)), (iswh, good)], b"kec-iswh") == (0, 0, 0)
    keys = V.keccak_is_wh_keys(ops)
    keys[17] = (keys[17][0], keys[17][1] ^ 1)
    bad = V.generate_is_wh_trace(V.dedup_is_wh_rows(keys))
    assert _total([(kec, V.generate_keccak_trace(ops)), (iswh, bad)], b"kec-iswh") != (0, 0, 0)


def test_keccak_address_read_balances_against_memw_register():
    ops = keccak_ops(5, 3)
    kec = _desc(A.logup_air(V.KeccakCols.NUM_COLUMNS, V.keccak_bus_interactions()[1:2]))
    memw_read = _desc(A.logup_air(V.MemwRegisterCols.NUM_COLUMNS, V.memw_register_bus_interactions()[5:6]))

    def reads(flip=None):
        return V.generate_memw_register_trace(
            [V.RegisterAccess(20, op.timestamp, op.state_addr ^ (8 if i == flip else 0),
                              op.state_addr ^ (8 if i == flip else 0), op.timestamp - 2, True)
             for i, op in enumerate(ops)])

    assert _total([(kec, V.generate_keccak_trace(ops)), (memw_read, reads())], b"kec-x10") == (0, 0, 0)
    assert _total([(kec, V.generate_keccak_trace(ops)), (memw_read, reads(flip=2))], b"kec-x10") != (0, 0, 0)


def test_keccak_proof_verifies():
    ops = keccak_ops(6, 4)
    kec, iswh = _desc(V.keccak_table_air()), _desc(V.is_wh_table_air())
    tables = [(kec, V.generate_keccak_trace(ops)),
              (iswh, V.generate_is_wh_trace(V.dedup_is_wh_rows(V.keccak_is_wh_keys(ops))))]
    proofs = replay_air_tables(tables, 2, 3, 2, 6, 0, b"keccak")
    assert verify_air_tables([kec, iswh], proofs, 2, 3, 2, 6, 0, b"keccak")[0]
