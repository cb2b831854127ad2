"""The reference VM's MEMW_REGISTER table (prover/src/tables/memw_register.rs)
captured as an AIR: its three constraints (both multiplicities and their sum
are bits) and the LogUp constraints of its seven bus interactions (the
timestamp-delta halfword check, four register Memory tokens, the 24-element
MEMW read and the 16-element MEMW write) vanish on every row of a generated
trace of interleaved register reads and writes; a row that is both read and
write breaks the sum constraint; in one transcript the Memory tokens of a
chain of accesses balance against a table that receives each register's
initial token and sends its final one, and stop balancing when one old value
is wrong; and the oracle's replay of a full proof is accepted."""
import numpy as np

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from lambda_vm_amd.logup import BusInteraction, BusValue, LinearTerm, Multiplicity, Packing
from oracle.air_proof import replay_air_tables, verify_air_tables
from test_vm_table_eq import _constraint_values

P = 0xFFFFFFFF00000001
HI = 3 << 32  # every timestamp in one high word, as the shared old-timestamp limb requires


def _chain(k, seed):
    """k accesses to four registers; returns the accesses and, per register,
    (initial timestamp, initial value, final timestamp, final value)."""
    rng = np.random.default_rng(seed)
    regs = [1, 5, 31, 255]
    state = {r: (HI + 0, int(rng.integers(0, 2**63)) * 2) for r in regs}
    init = dict(state)
    ts = HI + 10
    out = []
    for i in range(k):
        r = regs[int(rng.integers(0, len(regs)))]
        ts += int(rng.integers(1, 1000))
        old_ts, old = state[r]
        if ts - old_ts > 1 << 16:  # the table's delta check admits at most 2^16
            ts = old_ts + (1 << 16)
        is_read = bool(rng.integers(0, 2))
        val = old if is_read else int(rng.integers(0, 2**64, dtype=np.uint64))
        out.append(V.RegisterAccess(2 * r, ts, val, old, old_ts, is_read))
        state[r] = (ts, val)
    return out, {r: (init[r][0], init[r][1], state[r][0], state[r][1]) for r in regs}


def test_constraints_vanish_on_generated_trace(ora):
    air = V.memw_register_table_air()
    acc, _ = _chain(45, 1)
    ev = _constraint_values(ora, air, V.generate_memw_register_trace(acc))
    assert not np.any(ev % np.uint64(P))


def test_read_and_write_on_one_row_breaks_constraints(ora):
    air = V.memw_register_table_air()
    acc, _ = _chain(9, 2)
    tr = V.generate_memw_register_trace(acc)
    tr[3, V.MemwRegisterCols.MU_READ] = tr[3, V.MemwRegisterCols.MU_WRITE] = 1
    ev = _constraint_values(ora, air, tr)
    assert np.any(ev[2, 3] % np.uint64(P))


def _bookends(ends):
    """A table receiving each register's initial Memory tokens and sending its
    final ones (the REGISTER table's role on this bus)."""
    rows = []
    for r, (t0, v0, t1, v1) in ends.items():
        rows.append([r, t0 & 0xFFFFFFFF, t0 >> 32, v0 & 0xFFFFFFFF, v0 >> 32, 1, 0])
        rows.append([r, t1 & 0xFFFFFFFF, t1 >> 32, v1 & 0xFFFFFFFF, v1 >> 32, 0, 1])
    n = max(4, 1 << (len(rows) - 1).bit_length())
    tr = np.zeros((n, 7), dtype=np.uint64)
    tr[:len(rows)] = np.array(rows, dtype=np.uint64)
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    k = lambda v: BusValue.of_linear([LinearTerm(v)])  # noqa: E731
    its = []
    for i in range(2):
        tok = [k(1), BusValue.of_linear([LinearTerm(2, 0), LinearTerm(i)]), k(0), d(1), d(2), d(3 + i)]
        its.append(BusInteraction.receiver(V.BUS_MEMORY, Multiplicity("Column", (5,)), tok))
        its.append(BusInteraction.sender(V.BUS_MEMORY, Multiplicity("Column", (6,)), list(tok)))
    air = A.logup_air(7, its)
    air.bus_descriptor = A.descriptor(air)
    return air, tr


def test_memory_tokens_of_a_chain_balance():
    acc, ends = _chain(30, 3)
    reg_air = A.logup_air(V.MemwRegisterCols.NUM_COLUMNS, V.memw_register_bus_interactions()[1:5])
    reg_air.bus_descriptor = A.descriptor(reg_air)
    book_air, book = _bookends(ends)

    def total(accesses):
        out = replay_air_tables([(reg_air, V.generate_memw_register_trace(accesses)), (book_air, book)],
                                4, 3, 2, 4, 0, b"memory")
        a, b = out[0]["table_contribution"], out[1]["table_contribution"]
        return tuple((int(x) + int(y)) % P for x, y in zip(a, b))

    assert total(acc) == (0, 0, 0)
    bad = list(acc)
    x = bad[7]
    bad[7] = V.RegisterAccess(x.base_address, x.timestamp, x.value, x.old ^ 1, x.old_timestamp, x.is_read)
    assert total(bad) != (0, 0, 0)


def test_memw_register_proof_verifies():
    air = V.memw_register_table_air()
    air.bus_descriptor = A.descriptor(air)
    acc, _ = _chain(50, 4)
    proofs = replay_air_tables([(air, V.generate_memw_register_trace(acc))], 4, 3, 2, 6, 0, b"memw-r")
    ok, why = verify_air_tables([air], proofs, 4, 3, 2, 6, 0, b"memw-r")
    assert ok, why


def test_generator_refuses_accesses_the_table_cannot_route():
    """RegRow::from_memw's conditions (prover/src/tables/memw_register.rs:
    130-195): the old access shares the timestamp's high word and lies 1..2^16
    below it in the low word; register indexes fit 16 bits. A violating access
    would give a trace whose Memory/IS_HALFWORD tokens do not balance, so the
    generator refuses it."""
    import pytest
    ok = V.RegisterAccess(2, HI + 100, 7, 5, HI + 99, True)
    V.generate_memw_register_trace([ok])
    bad = [V.RegisterAccess(2, HI + 100, 7, 5, (4 << 32) + 99, True),      # other high word
           V.RegisterAccess(2, HI + 100, 7, 5, HI + 100, True),            # delta 0
           V.RegisterAccess(2, HI + (1 << 17), 7, 5, HI + 1, True),        # delta > 2^16
           V.RegisterAccess(2 << 16, HI + 100, 7, 5, HI + 99, True)]       # index >= 2^16
    for a in bad:
        with pytest.raises(ValueError):
            V.generate_memw_register_trace([a])
    # values wider than 64 bits are taken mod 2^64
    t = V.generate_memw_register_trace([V.RegisterAccess(2, HI + 100, (1 << 64) + 9, -1, HI + 99, False)])
    c = V.MemwRegisterCols
    assert int(t[0, c.VAL_0]) == 9 and int(t[0, c.VAL_1]) == 0
    assert int(t[0, c.OLD_0]) == 0xFFFFFFFF and int(t[0, c.OLD_1]) == 0xFFFFFFFF
