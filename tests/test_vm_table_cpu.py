"""The reference VM's CPU table (prover/src/tables/cpu.rs, constraints/cpu.rs):
37 columns (TIMESTAMP fixed), 38 constraints, 20 interactions.

The rows come from executing a generated RV64IM program (every arithmetic op
in register, immediate and word form, loads and stores of every width,
taken and not-taken branches of every condition, JAL, JALR, LUI, AUIPC) on a
small interpreter written here from the ISA, whose log of each step feeds
the restated cpu.rs from_log.
Pins independent of the restatement — every one a balance of the CPU's sends
against other restated chips fed the operations the trace builder derives,
in one transcript each:
  - ALU bus: the EQ, LT, BYTEWISE, SHIFT, MUL and DVRM chips (each computes its
    own result, so the interpreter's results and the CPU's res/arg2 wiring
    are checked against independent arithmetic), and stops balancing for one
    wrong result;
  - MEMORY-op bus: the LOAD and STORE chips;
  - BRANCH bus: the BRANCH chip (jump targets, JAL as JALR through x255);
  - DECODE bus: the DECODE table of the program, with the CPU32 rows of the
    delegated word instructions; CPU32 bus: CPU32's receivers;
  - MEMW bus: the register reads and writes of CPU and CPU32 against
    MEMW_REGISTER; Memory bus: the inline pc tokens of every row together
    with MEMW_REGISTER's register tokens close against the REGISTER table's
    initial image (stack top, entry point) and final state.
Constraints vanish on the trace and break for a wrong next pc, a wrong rvd on
an ALU row and a wrong branch decision."""
import numpy as np

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from oracle.air_proof import replay_air_tables, verify_air_tables
from test_vm_table_cpu32 import M64, rv64_word
from test_vm_table_decode import R_FUNCS, enc_b, enc_i, enc_j, enc_r, enc_s
from test_vm_table_eq import _constraint_values

P = 0xFFFFFFFF00000001
ENTRY = 0x1000
DATA = 0x8_0000


def _s64(v):
    return v - 2**64 if v >> 63 else v


def rv64_op(name, a, b):
    """RV64IM register-register semantics from the ISA."""
    sa, sb = _s64(a), _s64(b)
    if name in ("add", "sub", "and", "or", "xor"):
        return {"add": a + b, "sub": a - b, "and": a & b, "or": a | b, "xor": a ^ b}[name] & M64
    if name in ("sll", "srl", "sra"):
        sh = b & 63
        return ((a << sh) if name == "sll" else (a >> sh) if name == "srl" else (sa >> sh)) & M64
    if name in ("slt", "sltu"):
        return int(sa < sb) if name == "slt" else int(a < b)
    if name == "mul":
        return (a * b) & M64
    if name in ("mulh", "mulhsu", "mulhu"):
        x = sa if name in ("mulh", "mulhsu") else a
        y = sb if name == "mulh" else b
        return ((x * y) >> 64) & M64
    if name in ("div", "rem"):
        if b == 0:
            return M64 if name == "div" else a
        if sa == -2**63 and sb == -1:
            return a if name == "div" else 0
        q = abs(sa) // abs(sb) * (1 if (sa < 0) == (sb < 0) else -1)
        return (q if name == "div" else sa - q * sb) & M64
    if b == 0:
        return M64 if name == "divu" else a
    return a // b if name == "divu" else a % b


class Machine:
    """An RV64IM interpreter over a byte memory; step() returns the log the
    CPU rows need: (pc, instruction, x[rs1], x[rs2], value written, next pc)."""

    def __init__(self, words):
        self.mem = {}
        for i, w in enumerate(words):
            self.write(ENTRY + 4 * i, w, 4)
        self.x = [0] * 32
        self.x[2] = V.STACK_TOP
        self.pc = ENTRY

    def read(self, addr, n):
        return sum(self.mem.get((addr + i) & M64, 0) << (8 * i) for i in range(n))

    def write(self, addr, v, n):
        for i in range(n):
            self.mem[(addr + i) & M64] = (v >> (8 * i)) & 0xFF

    def step(self):
        pc = self.pc
        ins = V.decode_instruction(self.read(pc, 4))
        a, b = self.x[ins.src1], self.x[ins.src2]
        imm = V.instruction_imm(ins)
        nxt, dst, write = (pc + 4) & M64, 0, True
        k = ins.kind
        if k == "arith":
            dst = rv64_op(ins.op, a, b)
        elif k == "arith_imm":
            dst = rv64_op(ins.op, a, imm)
        elif k == "arith_w":
            dst = rv64_word(ins.op, a, b)
        elif k == "arith_imm_w":
            dst = rv64_word(ins.op, a, imm)
        elif k == "load":
            n = {"b": 1, "bu": 1, "h": 2, "hu": 2, "w": 4, "wu": 4, "d": 8}[ins.width]
            v = self.read((a + imm) & M64, n)
            if ins.width in ("b", "h", "w") and v >> (8 * n - 1):
                v |= M64 ^ ((1 << (8 * n)) - 1)
            dst = v
        elif k == "store":
            n = {"b": 1, "h": 2, "w": 4, "d": 8}[ins.width]
            self.write((a + imm) & M64, b, n)
            write = False
        elif k == "branch":
            taken = {"eq": a == b, "ne": a != b, "lt": _s64(a) < _s64(b), "ge": _s64(a) >= _s64(b),
                     "ltu": a < b, "geu": a >= b}[ins.cond]
            nxt = (pc + imm) & M64 if taken else nxt
            write = False
        elif k == "jal":
            dst, nxt = (pc + 4) & M64, (pc + imm) & M64
        elif k == "jalr":
            dst, nxt = (pc + 4) & M64, ((a + imm) & M64) & ~1
        elif k == "lui":
            dst = imm
        elif k == "auipc":
            dst = (pc + imm) & M64
        else:
            raise ValueError(k)
        if write and ins.dst:
            self.x[ins.dst] = dst
        self.pc = nxt
        return pc, ins, a, b, dst, nxt


def program(seed, length=400):
    """A program of about `length` words exercising every instruction class
    (straight-line apart from forward jumps, so a run never loops)."""
    rng = np.random.default_rng(seed)
    r = lambda lo=1: int(rng.integers(lo, 16))  # noqa: E731
    words = []
    # registers x1..x15 from LUI + ADDI, x20 the data base
    for i in range(1, 16):
        words += [(int(rng.integers(0, 1 << 20)) << 12) | i << 7 | 0b0110111,
                  enc_i(0b0010011, 0, i, i, int(rng.integers(-2048, 2048)))]
    words.append((DATA >> 12) << 12 | 20 << 7 | 0b0110111)
    words.append(enc_i(0b0010011, 0, 3, 0, -1))                       # x3 = -1
    words.append(enc_r(0b0110011, 0, 0, 4, 0, 0))                    # add x4, x0, x0
    ops = list(R_FUNCS)
    while len(words) < length:
        cls = int(rng.integers(0, 10))
        if cls <= 2:
            op = ops[int(rng.integers(0, len(ops)))]
            f3, f7 = R_FUNCS[op]
            words.append(enc_r(0b0110011, f3, f7, r(), r(0), r(0)))
        elif cls == 3:
            op, f3, hi = [("add", 0, None), ("slt", 2, None), ("sltu", 3, None), ("xor", 4, None), ("or", 6, None),
                          ("and", 7, None), ("sll", 1, 0), ("srl", 5, 0), ("sra", 5, 0x400)][int(rng.integers(0, 9))]
            imm = int(rng.integers(-2048, 2048)) if hi is None else hi | int(rng.integers(0, 64))
            words.append(enc_i(0b0010011, f3, r(), r(0), imm))
        elif cls == 4 and rng.integers(0, 2):
            op = sorted(V._W_OPS)[int(rng.integers(0, 10))]
            f3, f7 = R_FUNCS[op]
            words.append(enc_r(0b0111011, f3, f7, r(), r(0), r(0)))
        elif cls == 4:
            f3, hi = [(0, None), (1, 0), (5, 0), (5, 0x400)][int(rng.integers(0, 4))]
            imm = int(rng.integers(-2048, 2048)) if hi is None else hi | int(rng.integers(0, 32))
            words.append(enc_i(0b0011011, f3, r(), r(0), imm))
        elif cls == 5:
            f3 = int(rng.integers(0, 7))
            words.append(enc_i(0b0000011, f3, r(), 20, 8 * int(rng.integers(0, 32))))
        elif cls == 6:
            words.append(enc_s(int(rng.integers(0, 4)), 20, r(0), 8 * int(rng.integers(0, 32))))
        elif cls == 7:
            # a branch over one filler instruction (never a jump's first word)
            f3 = [0, 1, 4, 5, 6, 7][int(rng.integers(0, 6))]
            words += [enc_b(f3, r(0), r(0), 8), enc_i(0b0010011, 0, r(), r(0), 1)]
        elif cls == 8:
            words += [enc_j(r(), 8), enc_i(0b0010011, 0, r(), r(0), 3)] if rng.integers(0, 2) \
                else [21 << 7 | 0b0010111]                              # JAL over one filler, or AUIPC
        else:
            # auipc x21, 0 ; jalr rd, x21, 12 ; filler: the jump skips the filler
            words += [21 << 7 | 0b0010111, enc_i(0b1100111, 0, r(), 21, 12), enc_i(0b0010011, 0, r(), r(0), 5)]
    return words


def run(seed, steps=256):
    m = Machine(program(seed, length=2 * steps))
    logs = [m.step() for _ in range(steps)]
    return m, logs


def cpu_ops(logs):
    return [V.cpu_operation(pc, ins, 4 * (i + 1), a, b, dst, nxt) for i, (pc, ins, a, b, dst, nxt) in enumerate(logs)]


def cpu32_rows(ops):
    """trace_builder.rs:1463-1540: the CPU32 row of a word instruction."""
    from test_vm_table_cpu32 import make_op  # noqa: F401  (same chip-result rule)
    out = []
    for op in ops:
        f = op.fields
        if not f.word_instr:
            continue
        base = dict(timestamp=op.timestamp, pc=op.pc, rs1=f.rs1, read_register1=f.read_register1, rv1=op.rv1,
                    rs2=f.rs2, read_register2=f.read_register2, rv2=op.rv2, imm=op.imm, rd=f.rd,
                    write_register=f.write_register, alu=f.alu, alu_flags=f.alu_flags, add=f.add, sub=f.sub,
                    half_instruction_length=f.half_instruction_length)
        c = V.Cpu32Operation(res=0, **base)
        _, _, arg1, _, arg2, _, _ = c.aux()
        code, signed = f.alu_flags & 0x1F, bool(f.alu_flags >> 5 & 1)
        s2, muldiv = bool(f.alu_flags >> 6 & 1), bool(f.alu_flags >> 7 & 1)
        if f.add:
            res = (arg1 + arg2) & M64
        elif f.sub:
            res = (arg1 - arg2) & M64
        elif code in (V.ALU_OP_SHIFT, V.ALU_OP_SHIFTW):
            o = V.ShiftOperation(arg1, arg2, s2, signed, True).aux()[6]
            res = o[0] | o[1] << 32
        elif code == V.ALU_OP_MUL:
            res = V.MulOperation(arg1, signed, arg2, s2).product()[0]
        else:
            d = V.DvrmOperation(arg1, arg2, signed)
            res = d.remainder() if muldiv else d.quotient()
        out.append(V.Cpu32Operation(res=res, **base))
    return out


def alu_chip_ops(ops):
    """The chip operations of the CPU's ALU rows (trace_builder.rs:750-790 and
    the EQ/BYTEWISE/MUL/DVRM collectors): keyed by the alu_flags op."""
    eq, lt, bw, sh, mu, dv = [], [], [], [], [], []
    for op in ops:
        f = op.fields
        if f.word_instr or not f.alu:
            continue
        code, signed = f.alu_flags & 0x1F, bool(f.alu_flags >> 5 & 1)
        s2, muldiv = bool(f.alu_flags >> 6 & 1), bool(f.alu_flags >> 7 & 1)
        a, b = op.rv1, op.arg2
        if code == V.ALU_OP_EQ:
            eq.append(V.EqOperation(a, b, s2))
        elif code == V.ALU_OP_LT:
            lt.append(V.LtOperation(a, b, signed, s2))
        elif code in (V.ALU_OP_AND, V.ALU_OP_OR, V.ALU_OP_XOR):
            bw.append(V.BytewiseOperation(a, b, code))
        elif code == V.ALU_OP_SHIFT:
            sh.append(V.ShiftOperation(a, b, s2, signed, False))
        elif code == V.ALU_OP_MUL:
            mu.append((V.MulOperation(a, signed, b, s2), muldiv))
        elif code == V.ALU_OP_DIVREM:
            dv.append((V.DvrmOperation(a, b, signed), muldiv))
    return eq, lt, bw, sh, mu, dv


def _only(num_cols, its, bus, senders=None):
    a = A.logup_air(num_cols, [it for it in its if it.bus_id == bus and (senders is None or it.is_sender == senders)])
    a.bus_descriptor = A.descriptor(a)
    return a


def _total(tables, label):
    tot = [0, 0, 0]
    for o in replay_air_tables(tables, 2, 3, 3, 2, 0, label):
        tot = [(x + int(y)) % P for x, y in zip(tot, o["table_contribution"])]
    return tuple(tot)


def test_cpu_constraints(ora):
    _, logs = run(1)
    ops = cpu_ops(logs)
    tr = V.generate_cpu_trace(ops)
    air = V.cpu_table_air(tr.shape[0])
    assert tr.shape == (256, 37) and (air.num_main, len(air.interactions)) == (37, 20)
    kinds = {ins.kind for _, ins, *_ in logs}
    assert kinds >= {"arith", "arith_imm", "arith_w", "arith_imm_w", "load", "store", "branch", "jal", "jalr", "lui",
                     "auipc"}, kinds
    assert any(o.branch_cond for o in ops if o.fields.alu) and any(not o.branch_cond for o in ops if o.fields.branch)
    assert not np.any(_constraint_values(ora, air, tr) % np.uint64(P))
    c = V.CpuCols
    alu = next(i for i, o in enumerate(ops) if o.fields.alu and not o.fields.branch and not o.fields.word_instr)
    br = next(i for i, o in enumerate(ops) if o.fields.branch and o.fields.alu)
    for row, col, idx in ((5, c.NEXT_PC_0, (32, 34)), (alu, c.RVD_0, (27, 28)), (br, c.BRANCH_COND, (31, 32))):
        bad = tr.copy()
        bad[row, col] ^= 1
        assert np.any(_constraint_values(ora, air, bad)[idx[0]:idx[1], row] % np.uint64(P)), (row, col)


def test_cpu_alu_sends_balance_against_chips():
    _, logs = run(2)
    ops = cpu_ops(logs)
    eq, lt, bw, sh, mu, dv = alu_chip_ops(ops)
    assert all((eq, lt, bw, sh, mu, dv))
    cpu = _only(V.CpuCols.NUM_COLUMNS, V.cpu_bus_interactions(), V.BUS_ALU)
    chips = [(_only(V.EqCols.NUM_COLUMNS, V.eq_bus_interactions(), V.BUS_ALU, False), V.generate_eq_trace(eq)),
             (_only(V.LtCols.NUM_COLUMNS, V.lt_bus_interactions(), V.BUS_ALU, False), V.generate_lt_trace(lt)),
             (_only(V.BytewiseCols.NUM_COLUMNS, V.bytewise_bus_interactions(), V.BUS_ALU, False),
              V.generate_bytewise_trace(bw)),
             (_only(V.ShiftCols.NUM_COLUMNS, V.shift_bus_interactions(), V.BUS_ALU, False), V.generate_shift_trace(sh)),
             (_only(V.MulCols.NUM_COLUMNS, V.mul_bus_interactions(), V.BUS_ALU, False), V.generate_mul_trace(mu)),
             (_only(V.DvrmCols.NUM_COLUMNS, V.dvrm_bus_interactions(), V.BUS_ALU, False), V.generate_dvrm_trace(dv))]
    assert _total([(cpu, V.generate_cpu_trace(ops))] + chips, b"cpu-alu") == (0, 0, 0)
    i = next(k for k, o in enumerate(ops) if o.fields.alu and (o.fields.alu_flags & 0x1F) == V.ALU_OP_MUL)
    bad = list(ops)
    bad[i] = V.CpuOperation(**{**ops[i].__dict__, "res": (ops[i].res + 1) & M64,
                               "rvd": (ops[i].res + 1) & M64})
    assert _total([(cpu, V.generate_cpu_trace(bad))] + chips, b"cpu-alu") != (0, 0, 0)


def test_cpu_memory_and_branch_sends_balance_against_chips():
    _, logs = run(3)
    ops = cpu_ops(logs)
    loads, stores, branches = [], [], []
    widths = {"b": 1, "bu": 1, "h": 2, "hu": 2, "w": 4, "wu": 4, "d": 8}
    for op, (_, ins, *_rest) in zip(ops, logs):
        if ins.kind == "load":
            loads.append(V.LoadOperation.of_memory(op.res, op.timestamp, widths[ins.width],
                                                   ins.width in ("b", "h", "w"), op.rvd))
        elif ins.kind == "store":
            stores.append(V.StoreOperation.of_width(op.res, op.timestamp, op.rv2, widths[ins.width]))
        if op.branch_cond:
            branches.append(V.BranchOperation(op.pc, op.imm, op.rv1, bool(op.fields.mem_flags & 1)))
    assert loads and stores and branches
    tr = V.generate_cpu_trace(ops)
    cpu_m = _only(V.CpuCols.NUM_COLUMNS, V.cpu_bus_interactions(), V.BUS_MEMORY_OP)
    ld = _only(V.LoadCols.NUM_COLUMNS, V.load_bus_interactions(), V.BUS_MEMORY_OP)
    st = _only(V.StoreCols.NUM_COLUMNS, V.store_bus_interactions(), V.BUS_MEMORY_OP)
    assert _total([(cpu_m, tr), (ld, V.generate_load_trace(loads)), (st, V.generate_store_trace(stores))],
                  b"cpu-mem") == (0, 0, 0)
    cpu_b = _only(V.CpuCols.NUM_COLUMNS, V.cpu_bus_interactions(), V.BUS_BRANCH)
    bra = _only(V.BranchCols.NUM_COLUMNS, V.branch_bus_interactions(), V.BUS_BRANCH)
    assert _total([(cpu_b, tr), (bra, V.generate_branch_trace(branches))], b"cpu-branch") == (0, 0, 0)
    bad = list(branches)
    bad[0] = V.BranchOperation(bad[0].pc, bad[0].offset + 4, bad[0].register, bad[0].jalr)
    assert _total([(cpu_b, tr), (bra, V.generate_branch_trace(bad))], b"cpu-branch") != (0, 0, 0)


def test_cpu_decode_and_cpu32_sends_balance():
    _, logs = run(4)
    ops = cpu_ops(logs)
    words = program(4, length=512)
    entries = V.decode_entries_from_words([(ENTRY, words)])
    c32 = cpu32_rows(ops)
    assert c32
    tr = V.generate_cpu_trace(ops)
    dec = V.decode_table_air(entries)
    dec.bus_descriptor = A.descriptor(dec)
    lookups = [o.pc for o in ops]  # every row: CPU rows fetch, word rows through CPU32
    tables = [(_only(V.CpuCols.NUM_COLUMNS, V.cpu_bus_interactions(), V.BUS_DECODE), tr),
              (_only(V.Cpu32Cols.NUM_COLUMNS, V.cpu32_bus_interactions(), V.BUS_DECODE), V.generate_cpu32_trace(c32)),
              (dec, V.generate_decode_trace(entries, lookups))]
    assert _total(tables, b"cpu-decode") == (0, 0, 0)
    tables = [(_only(V.CpuCols.NUM_COLUMNS, V.cpu_bus_interactions(), V.BUS_CPU32), tr),
              (_only(V.Cpu32Cols.NUM_COLUMNS, V.cpu32_bus_interactions(), V.BUS_CPU32), V.generate_cpu32_trace(c32))]
    assert _total(tables, b"cpu-cpu32") == (0, 0, 0)


def register_accesses(ops, c32):
    """The register reads and writes of CPU and CPU32 rows in timestamp order,
    with each register's previous value and timestamp (from the program-start
    image at timestamp 1)."""
    init = V.register_init_from_entry_point(ENTRY)
    state = {r: (1, init[2 * r] | init[2 * r + 1] << 32) for r in range(32)}
    state[255] = (1, ENTRY)
    events = []
    for op in ops:
        f = op.fields
        if f.word_instr:
            continue
        if f.read_register1 and f.rs1 != 0:
            events.append((op.timestamp, f.rs1, op.rv1, True))
        if f.read_register2 and f.rs2 != 0:
            events.append((op.timestamp + 1, f.rs2, op.rv2, True))
        if f.write_register and f.rd != 0:
            events.append((op.timestamp + 2, f.rd, op.rvd, False))
    for c in c32:
        rvd = c.aux()[6]
        if c.read_register1:
            events.append((c.timestamp, c.rs1, c.rv1, True))
        if c.read_register2:
            events.append((c.timestamp + 1, c.rs2, c.rv2, True))
        if c.write_register:
            events.append((c.timestamp + 2, c.rd, rvd, False))
    acc = []
    pc_reads = {}
    for ts, r, v, is_read in sorted(events):
        old_ts, old = state[r]
        if r == 255:
            # the x255 read of JAL/AUIPC: the pc token the previous row wrote at ts - 3
            old_ts, old = ts - 3, v
            pc_reads[ts] = v
        acc.append(V.RegisterAccess(2 * r, ts, v, old, old_ts, is_read))
        if r != 255:
            state[r] = (ts, v)
    return acc, state


def test_cpu_register_and_pc_tokens_close_against_register_table():
    m, logs = run(5)
    ops = cpu_ops(logs)
    c32 = cpu32_rows(ops)
    tr = V.generate_cpu_trace(ops)
    acc, state = register_accesses(ops, c32)
    assert any(a.base_address == 510 for a in acc)
    # MEMW bus: CPU and CPU32 register reads/writes against MEMW_REGISTER
    mem_tr = V.generate_memw_register_trace(acc)
    tables = [(_only(V.CpuCols.NUM_COLUMNS, V.cpu_bus_interactions(), V.BUS_MEMW), tr),
              (_only(V.Cpu32Cols.NUM_COLUMNS, V.cpu32_bus_interactions(), V.BUS_MEMW), V.generate_cpu32_trace(c32)),
              (_only(V.MemwRegisterCols.NUM_COLUMNS, V.memw_register_bus_interactions(), V.BUS_MEMW), mem_tr)]
    assert _total(tables, b"cpu-memw") == (0, 0, 0)
    # Memory bus: pc tokens and register tokens against REGISTER's image
    init = V.register_init_from_entry_point(ENTRY)
    final = {}
    for r in range(1, 32):
        t, v = state[r]
        if t > 1:
            final[2 * r], final[2 * r + 1] = (t, v & 0xFFFFFFFF), (t, v >> 32)
    final_pc_ts = V.cpu_final_pc_timestamp(tr.shape[0])
    final[510], final[511] = (final_pc_ts, m.pc & 0xFFFFFFFF), (final_pc_ts, m.pc >> 32)
    assert all(m.x[r] == state[r][1] for r in range(32))
    reg = V.register_table_air(init, V.fini_from_final_state(final, init), final_pc_ts)
    reg.bus_descriptor = A.descriptor(reg)
    tables = [(_only(V.CpuCols.NUM_COLUMNS, V.cpu_bus_interactions(), V.BUS_MEMORY), tr),
              (_only(V.MemwRegisterCols.NUM_COLUMNS, V.memw_register_bus_interactions(), V.BUS_MEMORY), mem_tr),
              (reg, V.generate_register_trace(final, init))]
    assert _total(tables, b"cpu-memory") == (0, 0, 0)
    bad = dict(final)
    bad[511] = (final_pc_ts, (m.pc >> 32) ^ 1)
    reg_bad = V.register_table_air(init, V.fini_from_final_state(bad, init), final_pc_ts)
    reg_bad.bus_descriptor = A.descriptor(reg_bad)
    assert _total(tables[:2] + [(reg_bad, V.generate_register_trace(bad, init))], b"cpu-memory") != (0, 0, 0)


def test_cpu_proof_verifies():
    _, logs = run(6, steps=128)
    tr = V.generate_cpu_trace(cpu_ops(logs))
    air = V.cpu_table_air(tr.shape[0])
    air.bus_descriptor = A.descriptor(air)
    proofs = replay_air_tables([(air, tr)], 4, 3, 2, 6, 0, b"cpu")
    assert verify_air_tables([air], proofs, 4, 3, 2, 6, 0, b"cpu")[0]
