"""The reference VM's BRANCH table (prover/src/tables/branch.rs) captured as
an AIR: its five constraints (the carry pair of base + offset = unmasked next
pc on the pc path and on the register path, each weighted by its JALR
selector, and IS_BIT(JALR)) and the LogUp constraints of its six bus
interactions vanish on every row of a generated trace (branches and JALR
jumps, backward offsets that wrap, carries into both words, odd targets whose
low bit is cleared); a wrong target half or a flipped JALR flag breaks one of
the table's own constraints; and the oracle's replay of a full proof is
accepted."""
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
        pc, reg = int(rng.integers(0, 2**62)) * 4, int(rng.integers(0, 2**63)) * 2 + (i & 1)
        off = int(rng.integers(-2**20, 2**20)) & M  # sign-extended immediate
        ops.append(V.BranchOperation(pc, off, reg, bool(i % 3 == 0)))
    ops += [V.BranchOperation(0x1000, (-8) & M, 0, False),            # backward branch
            V.BranchOperation(0xFFFFFFFC, 8, 0, False),               # carry into the high word
            V.BranchOperation(0, 5, M - 2, True),                    # JALR wraps, odd target
            V.BranchOperation(M - 3, 4, 7, False),                   # pc + offset wraps to 0
            V.BranchOperation(0x1000, (-8) & M, 0, False)]           # a repeat: mu = 2
    return ops


def test_semantics_of_the_operations():
    assert V.BranchOperation(0, 5, M - 2, True).next_pc() == 2       # (2^64 - 3 + 5) & ~1
    assert V.BranchOperation(0x100, 0x11, 0, False).next_pc() == 0x110
    assert V.BranchOperation(0x100, 0x11, 0, False).next_pc_unmasked() == 0x111


def test_trace_layout_and_dedup():
    ops = _ops(5, 1)
    tr = V.generate_branch_trace(ops)
    c = V.BranchCols
    rows = V.dedup_rows(ops)
    assert len(rows) == len(ops) - 1 and tr[len(rows):].sum() == 0
    i = next(k for k, (op, _) in enumerate(rows) if op == ops[-1])
    assert tr[i, c.MU] == 2
    nxt = ops[-1].next_pc()
    got = int(tr[i, c.NEXT_PC_LOW_0]) | int(tr[i, c.NEXT_PC_LOW_1]) << 8 | int(tr[i, c.NEXT_PC_HIGH_0]) << 16 \
        | int(tr[i, c.NEXT_PC_HIGH_1]) << 32 | int(tr[i, c.NEXT_PC_HIGH_2]) << 48
    assert got == nxt


def test_constraints_vanish_on_generated_trace(ora):
    air = V.branch_table_air()
    tr = V.generate_branch_trace(_ops(27, 2))
    ev = _constraint_values(ora, air, tr)
    assert not np.any(ev % np.uint64(P))


@pytest.mark.parametrize("case", ["target_half", "jalr_flag"])
def test_wrong_values_break_constraints(ora, case):
    air = V.branch_table_air()
    ops = _ops(11, 3)
    tr = V.generate_branch_trace(ops)
    c = V.BranchCols
    rows = V.dedup_rows(ops)
    # a JALR row whose register differs from its pc
    row = next(k for k, (op, _) in enumerate(rows) if op.jalr and op.register != op.pc)
    if case == "target_half":
        tr[row, c.NEXT_PC_HIGH_2] ^= np.uint64(1)
        want = (2, 3)
    else:
        tr[row, c.JALR] = 0
        want = (0, 1)
    ev = _constraint_values(ora, air, tr)
    assert np.any(ev[list(want), row] % np.uint64(P))


def test_branch_proof_verifies():
    air = V.branch_table_air()
    air.bus_descriptor = A.descriptor(air)
    proofs = replay_air_tables([(air, V.generate_branch_trace(_ops(40, 4)))], 4, 3, 2, 6, 0, b"branch")
    ok, why = verify_air_tables([air], proofs, 4, 3, 2, 6, 0, b"branch")
    assert ok, why
