"""The reference VM's REGISTER table (prover/src/tables/register.rs) as a
preprocessed AIR: OFFSET, INIT, FINI and IS_PC fixed, the final timestamp
traced, three Memory-bus interactions. In one transcript the register Memory
tokens of a MEMW_REGISTER chain of accesses (starting from the program-start
image at timestamp 1) balance against REGISTER's initial and final tokens,
and stop balancing when one fixed final value is wrong; a proof of both tables
replays, carries the verifier's own commitment of the fixed columns and is
accepted."""
import numpy as np

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from oracle.air_proof import precomputed_commitment, replay_air_tables, verify_air_tables

P = 0xFFFFFFFF00000001
ENTRY = 0x10_0000


def register_chain(k, seed, regs=(1, 2, 10, 31)):
    """k accesses to `regs` starting from the program-start image at
    timestamp 1; returns the accesses and the final state per word address."""
    rng = np.random.default_rng(seed)
    init = V.register_init_from_entry_point(ENTRY)
    state = {r: (1, init[2 * r] | (init[2 * r + 1] << 32)) for r in regs}
    ts = 1
    out = []
    for _ in range(k):
        r = int(regs[int(rng.integers(0, len(regs)))])
        old_ts, old = state[r]
        ts = max(ts, old_ts) + int(rng.integers(1, 100))
        if ts - old_ts > 1 << 16:
            ts = old_ts + (1 << 16)
        is_read = bool(rng.integers(0, 2))
        val = old if is_read else int(rng.integers(0, 2**64, dtype=np.uint64))
        out.append(V.RegisterAccess(2 * r, ts, val, old, old_ts, is_read))
        state[r] = (ts, val)
    final = {}
    for r, (t, v) in state.items():
        final[2 * r] = (t, v & 0xFFFFFFFF)
        final[2 * r + 1] = (t, v >> 32)
    return out, final, init


def _register_air(final, init):
    fini = V.fini_from_final_state(final, init)
    air = V.register_table_air(init, fini, final_pc_ts=1)
    air.bus_descriptor = A.descriptor(air)
    return air, V.generate_register_trace(final, init)


def test_register_rows_follow_the_reference_layout():
    init = V.register_init_from_entry_point(ENTRY)
    tr = V.generate_register_trace({}, init)
    c = V.RegisterCols
    assert tr.shape == (128, 6)
    assert list(tr[[0, 63, 64, 65, 66, 67], c.OFFSET]) == [0, 63, 508, 509, 510, 511]
    assert (tr[4, c.INIT], tr[5, c.INIT]) == (V.STACK_TOP & 0xFFFFFFFF, V.STACK_TOP >> 32)
    assert (tr[66, c.INIT], tr[67, c.INIT]) == (ENTRY & 0xFFFFFFFF, ENTRY >> 32)
    assert list(tr[:, c.IS_PC].nonzero()[0]) == [66, 67]
    assert np.array_equal(tr[:68, c.FINI], tr[:68, c.INIT]) and set(tr[:, c.TIMESTAMP_LO]) == {1}
    halted = V.halted_fini(12)
    assert halted[64] == 12 and halted[66] == V.CPU_PADDING_PC and sum(halted) == 13


def test_memory_tokens_balance_against_register():
    acc, final, init = register_chain(40, 5)
    mem_air = A.logup_air(V.MemwRegisterCols.NUM_COLUMNS, V.memw_register_bus_interactions()[1:5])
    mem_air.bus_descriptor = A.descriptor(mem_air)

    def total(final_state):
        reg_air, reg_tr = _register_air(final_state, init)
        out = replay_air_tables([(mem_air, V.generate_memw_register_trace(acc)), (reg_air, reg_tr)],
                                4, 3, 2, 4, 0, b"registers")
        a, b = out[0]["table_contribution"], out[1]["table_contribution"]
        return tuple((int(x) + int(y)) % P for x, y in zip(a, b))

    assert total(final) == (0, 0, 0)
    bad = dict(final)
    t, v = bad[20]
    bad[20] = (t, v ^ 1)
    assert total(bad) != (0, 0, 0)


def test_register_proof_verifies_with_its_fixed_commitment():
    acc, final, init = register_chain(20, 6)
    reg_air, reg_tr = _register_air(final, init)
    mem = V.memw_register_table_air()
    mem.bus_descriptor = A.descriptor(mem)
    tables = [(reg_air, reg_tr), (mem, V.generate_memw_register_trace(acc))]
    proofs = replay_air_tables(tables, 4, 3, 2, 6, 0, b"register")
    assert proofs[0]["precomputed_root"] == precomputed_commitment(reg_air, 4, 3)
    ok, why = verify_air_tables([reg_air, mem], proofs, 4, 3, 2, 6, 0, b"register")
    assert ok, why
    # a verifier that expects the halted register file refuses this epoch's FINI
    halted = V.register_table_air(init, V.halted_fini(0), final_pc_ts=1)
    halted.bus_descriptor = A.descriptor(halted)
    assert verify_air_tables([halted, mem], proofs, 4, 3, 2, 6, 0, b"register") == (False, "precomputed commitment")
