"""The reference VM's STORE table (prover/src/tables/store.rs) captured as an
AIR: its six constraints (four bit checks, the width sum is a bit, a width
implies mu) and the LogUp constraints of its ten bus interactions (a
16-element MEMW write, the MEMORY op, eight byte range checks) vanish on
every row of a generated trace (1-, 2-, 4- and 8-byte stores, extreme
addresses, timestamps and values); two width flags on one row, or a width
flag on a row with mu = 0, break the table's own constraints; a STORE table
and a MEMORY-op sender carrying the same ops balance the MEMORY bus; and
the oracle's replay of a full proof is accepted."""
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
    ops = [V.StoreOperation.of_width(int(a) * 2 + 1, int(t), int(v) * 3, (1, 2, 4, 8)[i % 4])
           for i, (a, t, v) in enumerate(rng.integers(0, 2**62, (k, 3)))]
    ops += [V.StoreOperation.of_width(M, M, M, 8), V.StoreOperation.of_width(0, 1, 0, 1),
            V.StoreOperation.of_width(1 << 63, (1 << 32) - 1, 0xFF, 2)]
    return ops


def test_trace_layout():
    tr = V.generate_store_trace([V.StoreOperation.of_width(0x1122334455667788, 7, 0x0102030405060708, 4)])
    c = V.StoreCols
    assert tr.shape == (4, 16)
    assert tr[0, c.BASE_ADDRESS_0] == 0x55667788 and tr[0, c.BASE_ADDRESS_1] == 0x11223344
    assert [int(x) for x in tr[0, list(c.VALUE)]] == [8, 7, 6, 5, 4, 3, 2, 1]
    assert tr[0, c.WRITE4] == 1 and tr[0, c.WRITE2] == 0 and tr[0, c.WRITE8] == 0 and tr[0, c.MU] == 1
    assert not tr[1:].any()


def test_constraints_vanish_on_generated_trace(ora):
    air = V.store_table_air()
    tr = V.generate_store_trace(_ops(21, 3))
    ev = _constraint_values(ora, air, tr)
    assert not np.any(ev % np.uint64(P))


@pytest.mark.parametrize("case", ["two_widths", "width_without_mu"])
def test_bad_flags_break_constraints(ora, case):
    air = V.store_table_air()
    ops = _ops(9, 4)
    tr = V.generate_store_trace(ops)
    c = V.StoreCols
    if case == "two_widths":
        row = 1  # a 2-byte store
        tr[row, c.WRITE8] = 1
        want = 4
    else:
        row = len(ops)  # a padding row (mu = 0)
        tr[row, c.WRITE2] = 1
        want = 5
    ev = _constraint_values(ora, air, tr)
    assert np.any(ev[want, row] % np.uint64(P))


def test_memory_bus_balances_against_a_sender():
    """In one transcript, the MEMORY receiver of STORE cancels a sender table
    that carries the same [timestamp, address, value, flags, 0, 0] tokens (the
    CPU side of the bus): the two tables' LogUp contributions sum to zero."""
    from lambda_vm_amd.logup import BusInteraction, BusValue, LinearTerm, Multiplicity, Packing
    ops = _ops(13, 6)
    st = V.generate_store_trace(ops)
    c = V.StoreCols
    cpu = np.zeros((st.shape[0], 8), dtype=np.uint64)
    cpu[:, 0:2] = st[:, [c.TIMESTAMP_0, c.TIMESTAMP_1]]
    cpu[:, 2:4] = st[:, [c.BASE_ADDRESS_0, c.BASE_ADDRESS_1]]
    w = np.array([1, 1 << 8, 1 << 16, 1 << 24], dtype=np.uint64)
    cpu[:, 4] = st[:, list(c.VALUE[:4])] @ w
    cpu[:, 5] = st[:, list(c.VALUE[4:])] @ w
    cpu[:, 6] = (1 + 4 * st[:, c.WRITE2] + 8 * st[:, c.WRITE4] + 16 * st[:, c.WRITE8]) * st[:, c.MU]
    cpu[:, 7] = st[:, c.MU]
    zero = BusValue.of_linear([LinearTerm(0)])
    sender = [BusInteraction.sender(V.BUS_MEMORY_OP, Multiplicity("Column", (7,)), [
        BusValue.packed(Packing.DWORD_WL, 0), BusValue.packed(Packing.DWORD_WL, 2), BusValue.packed(Packing.DWORD_WL, 4),
        BusValue.packed(Packing.DIRECT, 6), zero, zero])]
    cpu_air = A.logup_air(8, sender)
    cpu_air.bus_descriptor = A.descriptor(cpu_air)
    # STORE with only its MEMORY receiver, so that no other bus is left open
    st_air = A.logup_air(V.StoreCols.NUM_COLUMNS, [V.store_bus_interactions()[1]])
    st_air.bus_descriptor = A.descriptor(st_air)
    out = replay_air_tables([(st_air, st), (cpu_air, cpu)], 4, 3, 2, 4, 0, b"memory-bus")
    s, r = out[0]["table_contribution"], out[1]["table_contribution"]
    assert tuple((int(x) + int(y)) % P for x, y in zip(s, r)) == (0, 0, 0)
    cpu[0, 4] ^= np.uint64(1)  # one value byte differs: the bus no longer balances
    out = replay_air_tables([(st_air, st), (cpu_air, cpu)], 4, 3, 2, 4, 0, b"memory-bus")
    s, r = out[0]["table_contribution"], out[1]["table_contribution"]
    assert tuple((int(x) + int(y)) % P for x, y in zip(s, r)) != (0, 0, 0)


def test_store_proof_verifies():
    air = V.store_table_air()
    air.bus_descriptor = A.descriptor(air)
    proofs = replay_air_tables([(air, V.generate_store_trace(_ops(30, 8)))], 4, 3, 2, 6, 0, b"store")
    ok, why = verify_air_tables([air], proofs, 4, 3, 2, 6, 0, b"store")
    assert ok, why
