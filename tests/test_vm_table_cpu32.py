"""The reference VM's CPU32 table (prover/src/tables/cpu32.rs): the RV64 word
instructions (ADDW, SUBW, SLLW, SRLW, SRAW, MULW, DIVW, DIVUW, REMW, REMUW and
the immediate forms ADDIW, SLLIW, SRLIW, SRAIW), 38 columns, 32 constraints,
23 interactions.

Pins independent of the restatement: the register value each row writes
(rvd) equals the RV64 definition of the instruction, computed here on Python
integers (32-bit wrap-around, signed division truncating toward zero,
division by zero and -2^31 / -1 as the ISA defines them), for random and
extreme register values; the alu_flags bytes follow the decode table
(types.rs:864-900). Constraints vanish on generated traces and break for a
wrong ADD result, a wrong sign of the result, a register limb on an unread
register and a flag on a padding row.
Cross-table: the ALU sends balance in one transcript against the SHIFT, MUL
and DVRM chips' receivers fed the operations the trace builder derives from
CPU32 rows (trace_builder.rs:1594-1618 cpu32_chip_op), and stop balancing for
a wrong result; the register reads and writes balance on the MEMW bus against
MEMW_REGISTER's receivers."""
import numpy as np

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from oracle.air_proof import replay_air_tables, verify_air_tables
from test_vm_table_eq import _constraint_values

P = 0xFFFFFFFF00000001
M64, M32 = 2**64 - 1, 2**32 - 1
REG_OPS = ["add", "sub", "sll", "srl", "sra", "mul", "div", "divu", "rem", "remu"]
IMM_OPS = ["add", "sll", "srl", "sra"]
EXTREMES = [0, 1, M64, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF, 0x1_0000_0000, 0xFFFFFFFF_80000000, 5, 2**63]


def _s32(v):
    v &= M32
    return v - 2**32 if v >> 31 else v


def _sext32(v):
    return _s32(v) & M64


def rv64_word(name, a, b):
    """The RV64 *W result written to rd, from the ISA's definition."""
    if name == "add":
        return _sext32(a + b)
    if name == "sub":
        return _sext32(a - b)
    if name == "sll":
        return _sext32((a & M32) << (b & 31))
    if name == "srl":
        return _sext32((a & M32) >> (b & 31))
    if name == "sra":
        return _sext32(_s32(a) >> (b & 31))
    if name == "mul":
        return _sext32(a * b)
    x, y = _s32(a), _s32(b)
    ux, uy = a & M32, b & M32
    if name in ("div", "rem"):
        if y == 0:
            return M64 if name == "div" else _sext32(x)
        if x == -2**31 and y == -1:
            return _sext32(x) if name == "div" else 0
        q = abs(x) // abs(y) * (1 if (x < 0) == (y < 0) else -1)
        return _sext32(q) if name == "div" else _sext32(x - q * y)
    if uy == 0:
        return M64 if name == "divu" else _sext32(ux)
    return _sext32(ux // uy) if name == "divu" else _sext32(ux % uy)


def make_op(name, rv1, operand, imm_form, ts, pc=0x1000, rs1=5, rs2=6, rd=7):
    flags, alu, add, sub = V.arith_decode(name, True)
    rv2, imm = (0, operand & M64) if imm_form else (operand & M64, 0)
    rs2 = 0 if imm_form else rs2
    base = dict(timestamp=ts, pc=pc, rs1=rs1, read_register1=rs1 != 0, rv1=rv1 & M64 if rs1 else 0, rs2=rs2,
                read_register2=rs2 != 0, rv2=rv2 if rs2 else 0, imm=imm, rd=rd, write_register=rd != 0, alu=alu,
                alu_flags=flags, add=add, sub=sub, half_instruction_length=2)
    op = V.Cpu32Operation(res=0, **base)
    _, _, arg1, _, arg2, _, _ = op.aux()
    signed, s2, muldiv = bool(flags >> 5 & 1), bool(flags >> 6 & 1), bool(flags >> 7 & 1)
    code = flags & 0x1F
    if add:
        res = (arg1 + arg2) & M64
    elif sub:
        res = (arg1 - arg2) & M64
    elif code == V.ALU_OP_SHIFTW:
        # the SHIFT chip's raw output words (the word result before CPU32 extends it)
        out = V.ShiftOperation(arg1, arg2, s2, signed, True).aux()[6]
        res = out[0] | out[1] << 32
    elif code == V.ALU_OP_MUL:
        res = V.MulOperation(arg1, signed, arg2, s2).product()[1 if muldiv else 0]
    else:
        d = V.DvrmOperation(arg1, arg2, signed)
        res = d.remainder() if muldiv else d.quotient()
    return V.Cpu32Operation(res=res, **base)


def cpu32_ops(k, seed):
    rng = np.random.default_rng(seed)
    ops, want = [], []
    for i in range(k):
        imm_form = i % 4 == 3
        name = IMM_OPS[i % len(IMM_OPS)] if imm_form else REG_OPS[i % len(REG_OPS)]
        a = EXTREMES[i % len(EXTREMES)] if i % 3 == 0 else int(rng.integers(0, 2**64, dtype=np.uint64))
        if imm_form:
            b = int(rng.integers(0, 32)) if name != "add" else int(rng.integers(-2048, 2048))
        elif i % 5 == 0:
            b = EXTREMES[(i // 5) % len(EXTREMES)]
        else:
            b = int(rng.integers(0, 2**64, dtype=np.uint64))
        ops.append(make_op(name, a, b, imm_form, ts=4 * (i + 1), rs1=(i % 31) + 1, rs2=((i + 7) % 31) + 1,
                           rd=(i % 30) + 1))
        want.append(rv64_word(name, a, b & M64))
    return ops, want


def test_cpu32_results():
    ops, want = cpu32_ops(400, 1)
    tr = V.generate_cpu32_trace(ops)
    c = V.Cpu32Cols
    for r, (op, w) in enumerate(zip(ops, want)):
        assert int(tr[r, c.RVD_0]) | int(tr[r, c.RVD_1]) << 32 == w, (r, op)
    # the decode flags of a few instructions (types.rs:870-895 and build_alu_flags)
    assert V.arith_decode("sra", True) == (6 + 32 + 64, True, False, False)
    assert V.arith_decode("sra", False) == (5 + 32 + 64, True, False, False)
    assert V.arith_decode("remu", True) == (8 + 128, True, False, False)
    assert V.arith_decode("mulhsu", False) == (7 + 32 + 128, True, False, False)
    assert V.arith_decode("sub", True) == (0, False, False, True)


def test_cpu32_constraints(ora):
    air = V.cpu32_table_air()
    ops, _ = cpu32_ops(100, 2)
    tr = V.generate_cpu32_trace(ops)
    assert (air.num_main, len(air.interactions)) == (38, 23)
    assert not np.any(_constraint_values(ora, air, tr) % np.uint64(P))
    c = V.Cpu32Cols
    a = next(i for i, o in enumerate(ops) if o.add)
    bad = tr.copy()
    bad[a, c.RES[0]] ^= 1
    assert np.any(_constraint_values(ora, air, bad)[7:9, a] % np.uint64(P))
    bad = tr.copy()
    bad[3, c.RES_SIGN] ^= 1
    assert np.any(_constraint_values(ora, air, bad)[22, 3] % np.uint64(P))
    u = next(i for i, o in enumerate(ops) if not o.read_register2)
    bad = tr.copy()
    bad[u, c.RV2[1]] = 9
    assert np.any(_constraint_values(ora, air, bad)[15, u] % np.uint64(P))
    bad = tr.copy()
    bad[tr.shape[0] - 1, c.WRITE_REGISTER] = 1
    assert np.any(_constraint_values(ora, air, bad)[29, tr.shape[0] - 1] % np.uint64(P))


def chip_ops(ops):
    """trace_builder.rs:1594-1618: the SHIFT, MUL and DVRM operations of ALU rows."""
    shifts, muls, dvrms = [], [], []
    for op in ops:
        if op.add or op.sub or not op.alu:
            continue
        _, _, arg1, _, arg2, _, _ = op.aux()
        code, signed = op.alu_flags & 0x1F, op.signed()
        s2, muldiv = bool(op.alu_flags >> 6 & 1), bool(op.alu_flags >> 7 & 1)
        if code in (V.ALU_OP_SHIFT, V.ALU_OP_SHIFTW):
            shifts.append(V.ShiftOperation(arg1, arg2, s2, signed, True))
        elif code == V.ALU_OP_MUL:
            muls.append((V.MulOperation(arg1, signed, arg2, s2), muldiv))
        elif code == V.ALU_OP_DIVREM:
            dvrms.append((V.DvrmOperation(arg1, arg2, signed), muldiv))
    return shifts, muls, dvrms


def _alu_only(num_cols, its, senders):
    a = A.logup_air(num_cols, [it for it in its if it.bus_id == V.BUS_ALU and it.is_sender == senders])
    a.bus_descriptor = A.descriptor(a)
    return a


def _total(tables, label):
    tot = [0, 0, 0]
    for o in replay_air_tables(tables, 2, 3, 3, 2, 0, label):
        tot = [(x + int(y)) % P for x, y in zip(tot, o["table_contribution"])]
    return tuple(tot)


def test_cpu32_alu_sends_balance_against_chips():
    ops, _ = cpu32_ops(80, 3)
    cpu = _alu_only(V.Cpu32Cols.NUM_COLUMNS, V.cpu32_bus_interactions(), True)
    sh = _alu_only(V.ShiftCols.NUM_COLUMNS, V.shift_bus_interactions(), False)
    mu = _alu_only(V.MulCols.NUM_COLUMNS, V.mul_bus_interactions(), False)
    dv = _alu_only(V.DvrmCols.NUM_COLUMNS, V.dvrm_bus_interactions(), False)
    shifts, muls, dvrms = chip_ops(ops)
    assert shifts and muls and dvrms

    def total(ops):
        return _total([(cpu, V.generate_cpu32_trace(ops)), (sh, V.generate_shift_trace(shifts)),
                       (mu, V.generate_mul_trace(muls)), (dv, V.generate_dvrm_trace(dvrms))], b"cpu32-alu")

    assert total(ops) == (0, 0, 0)
    i = next(k for k, o in enumerate(ops) if o.alu and (o.alu_flags & 0x1F) == V.ALU_OP_DIVREM)
    bad = list(ops)
    o = bad[i]
    bad[i] = V.Cpu32Operation(**{**o.__dict__, "res": (o.res + 1) & M64})
    assert total(bad) != (0, 0, 0)


def test_cpu32_register_accesses_balance_against_memw_register():
    ops, _ = cpu32_ops(40, 4)
    acc = []
    for op in ops:
        rvd = op.aux()[6]
        if op.read_register1:
            acc.append(V.RegisterAccess(2 * op.rs1, op.timestamp, op.rv1, op.rv1, op.timestamp - 1, True))
        if op.read_register2:
            acc.append(V.RegisterAccess(2 * op.rs2, op.timestamp + 1, op.rv2, op.rv2, op.timestamp, True))
        if op.write_register:
            acc.append(V.RegisterAccess(2 * op.rd, op.timestamp + 2, rvd, 0, op.timestamp + 1, False))
    cpu = A.logup_air(V.Cpu32Cols.NUM_COLUMNS, [it for it in V.cpu32_bus_interactions() if it.bus_id == V.BUS_MEMW])
    reg = A.logup_air(V.MemwRegisterCols.NUM_COLUMNS, [it for it in V.memw_register_bus_interactions()
                                                       if it.bus_id == V.BUS_MEMW])
    for a in (cpu, reg):
        a.bus_descriptor = A.descriptor(a)
    assert _total([(cpu, V.generate_cpu32_trace(ops)), (reg, V.generate_memw_register_trace(acc))],
                  b"cpu32-regs") == (0, 0, 0)
    bad = list(acc)
    w = next(k for k, a in enumerate(bad) if not a.is_read)
    bad[w] = V.RegisterAccess(bad[w].base_address, bad[w].timestamp, bad[w].value ^ 1, 0, bad[w].old_timestamp, False)
    assert _total([(cpu, V.generate_cpu32_trace(ops)), (reg, V.generate_memw_register_trace(bad))],
                  b"cpu32-regs") != (0, 0, 0)


def test_cpu32_proof_verifies():
    air = V.cpu32_table_air()
    air.bus_descriptor = A.descriptor(air)
    tables = [(air, V.generate_cpu32_trace(cpu32_ops(50, 5)[0]))]
    proofs = replay_air_tables(tables, 4, 3, 2, 6, 0, b"cpu32")
    assert verify_air_tables([air], proofs, 4, 3, 2, 6, 0, b"cpu32")[0]
