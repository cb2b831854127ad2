"""The reference VM's DECODE table (prover/src/tables/decode.rs) with the RV64IM
decoder (executor/src/vm/instruction/decoding.rs) and the decode word
(prover/src/tables/types.rs:580-1110).

Pins independent of the restatement: an encoder written here from the RISC-V
instruction formats (R, I, S, B, U, J and the shift-immediate forms) produces
words the decoder maps back to the same fields, for every opcode and random
registers and immediates (negative ones included); encodings the reference
rejects raise; the decode-word layout facts of the reference's own tests
(decode_layout_tests.rs: field positions, build_alu_flags/build_mem_flags
values, ADD/SUB fast paths, branch and jump flags, load/store widths, ECALL).
Cross-table: the DECODE sends of CPU32 rows built from decoded word
instructions balance in one transcript against the DECODE table fed their
pcs, and stop balancing for a wrong register index; a proof with the fixed
columns in their own tree verifies with the verifier's own fixed-column
commitment and fails for another program."""
import numpy as np
import pytest

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from oracle.air_proof import precomputed_commitment, replay_air_tables, verify_air_tables
from test_vm_table_cpu32 import M64, make_op

P = 0xFFFFFFFF00000001
R_FUNCS = {"add": (0, 0), "sub": (0, 0x20), "sll": (1, 0), "slt": (2, 0), "sltu": (3, 0), "xor": (4, 0),
           "srl": (5, 0), "sra": (5, 0x20), "or": (6, 0), "and": (7, 0), "mul": (0, 1), "mulh": (1, 1),
           "mulhsu": (2, 1), "mulhu": (3, 1), "div": (4, 1), "divu": (5, 1), "rem": (6, 1), "remu": (7, 1)}


def enc_r(opc, f3, f7, rd, rs1, rs2):
    return f7 << 25 | rs2 << 20 | rs1 << 15 | f3 << 12 | rd << 7 | opc


def enc_i(opc, f3, rd, rs1, imm):
    return (imm & 0xFFF) << 20 | rs1 << 15 | f3 << 12 | rd << 7 | opc


def enc_s(f3, rs1, rs2, imm):
    imm &= 0xFFF
    return (imm >> 5) << 25 | rs2 << 20 | rs1 << 15 | f3 << 12 | (imm & 0x1F) << 7 | 0b0100011


def enc_b(f3, rs1, rs2, imm):
    imm &= 0x1FFF
    return ((imm >> 12) & 1) << 31 | ((imm >> 5) & 0x3F) << 25 | rs2 << 20 | rs1 << 15 | f3 << 12 \
        | ((imm >> 1) & 0xF) << 8 | ((imm >> 11) & 1) << 7 | 0b1100011


def enc_j(rd, imm):
    imm &= 0x1FFFFF
    return ((imm >> 20) & 1) << 31 | ((imm >> 1) & 0x3FF) << 21 | ((imm >> 11) & 1) << 20 \
        | ((imm >> 12) & 0xFF) << 12 | rd << 7 | 0b1101111


def random_program(k, seed):
    """(word, expected Instruction) pairs over every instruction class."""
    rng = np.random.default_rng(seed)
    r = lambda: int(rng.integers(0, 32))  # noqa: E731
    out = []
    for i in range(k):
        cls = i % 13
        rd, rs1, rs2 = r(), r(), r()
        if cls == 0:
            op = list(R_FUNCS)[i % len(R_FUNCS)]
            f3, f7 = R_FUNCS[op]
            out.append((enc_r(0b0110011, f3, f7, rd, rs1, rs2), V.Instruction("arith", rd, rs1, rs2, op=op)))
        elif cls == 1:
            op = sorted(V._W_OPS)[i % 10]
            f3, f7 = R_FUNCS[op]
            out.append((enc_r(0b0111011, f3, f7, rd, rs1, rs2), V.Instruction("arith_w", rd, rs1, rs2, op=op)))
        elif cls == 2:
            op, f3 = [("add", 0), ("slt", 2), ("sltu", 3), ("xor", 4), ("or", 6), ("and", 7)][i % 6]
            imm = int(rng.integers(-2048, 2048))
            out.append((enc_i(0b0010011, f3, rd, rs1, imm), V.Instruction("arith_imm", rd, rs1, imm=imm, op=op)))
        elif cls == 3:
            op, f3, hi = [("sll", 1, 0), ("srl", 5, 0), ("sra", 5, 0x400)][i % 3]
            sh = int(rng.integers(0, 64))
            out.append((enc_i(0b0010011, f3, rd, rs1, hi | sh), V.Instruction("arith_imm", rd, rs1, imm=sh, op=op)))
        elif cls == 4:
            op, f3, hi = [("add", 0, None), ("sll", 1, 0), ("srl", 5, 0), ("sra", 5, 0x400)][i % 4]
            if hi is None:
                imm = int(rng.integers(-2048, 2048))
                word = enc_i(0b0011011, 0, rd, rs1, imm)
            else:
                imm = int(rng.integers(0, 32))
                word = enc_i(0b0011011, f3, rd, rs1, hi | imm)
            out.append((word, V.Instruction("arith_imm_w", rd, rs1, imm=imm, op=op)))
        elif cls == 5:
            f3, width = [(0, "b"), (1, "h"), (2, "w"), (3, "d"), (4, "bu"), (5, "hu"), (6, "wu")][i % 7]
            imm = int(rng.integers(-2048, 2048))
            out.append((enc_i(0b0000011, f3, rd, rs1, imm), V.Instruction("load", rd, rs1, imm=imm, width=width)))
        elif cls == 6:
            f3, width = [(0, "b"), (1, "h"), (2, "w"), (3, "d")][i % 4]
            imm = int(rng.integers(-2048, 2048))
            out.append((enc_s(f3, rs1, rs2, imm), V.Instruction("store", 0, rs1, rs2, imm=imm, width=width)))
        elif cls == 7:
            f3, cond = [(0, "eq"), (1, "ne"), (4, "lt"), (5, "ge"), (6, "ltu"), (7, "geu")][i % 6]
            imm = 2 * int(rng.integers(-2048, 2048))
            out.append((enc_b(f3, rs1, rs2, imm), V.Instruction("branch", 0, rs1, rs2, imm=imm, cond=cond)))
        elif cls == 8:
            imm = 2 * int(rng.integers(-(1 << 19), 1 << 19))
            out.append((enc_j(rd, imm), V.Instruction("jal", rd, imm=imm)))
        elif cls == 9:
            imm = int(rng.integers(-2048, 2048))
            out.append((enc_i(0b1100111, 0, rd, rs1, imm), V.Instruction("jalr", rd, rs1, imm=imm)))
        elif cls == 10:
            hi = int(rng.integers(0, 1 << 20)) << 12
            opc, kind = [(0b0110111, "lui"), (0b0010111, "auipc")][i % 2]
            out.append((hi | rd << 7 | opc, V.Instruction(kind, rd, imm=hi)))
        elif cls == 11:
            out.append((0x00000073, V.Instruction("ecall")))
        else:
            out.append((0x0FF0000F, V.Instruction("fence")))
    return out


def test_decoder_round_trip():
    for word, want in random_program(900, 1):
        assert V.decode_instruction(word) == want, hex(word)
    assert V.decode_instruction(0x00100073) == V.Instruction("ecall")  # EBREAK decodes as the same instruction
    assert V.decode_instruction(enc_i(0b1110011, 2, 5, 0, 0xC00)) == V.Instruction("csr", 5, 0, op="csrrs", csr=0xC00)
    for bad in (0x0000007F,                                 # unknown opcode
                enc_r(0b0111011, 7, 0, 1, 2, 3),            # ANDW does not exist
                enc_i(0b0011011, 5, 1, 2, 0x200 | 3),       # SRxIW with a bad function field
                enc_i(0b0010011, 1, 1, 2, 0x400 | 3),       # SLLI with a bad function field
                enc_i(0b1110011, 3, 1, 2, 0),               # CSRRC rejected
                enc_i(0b1100111, 1, 1, 2, 0),               # JALR with func3 != 0
                enc_s(5, 1, 2, 0),                          # store width 5
                enc_b(2, 1, 2, 0)):                         # branch func3 2
        with pytest.raises(ValueError):
            V.decode_instruction(bad)


def test_decode_word_layout_facts():
    # decode_layout_tests.rs: flag builders, field placement, the instruction classes
    assert V.arith_decode("sltu", False)[0] == V.ALU_OP_LT and V.build_mem_flags(1, 0, 0, 0, 1) == 1 + 16
    assert V.ShrunkDecode(memory=True).pack() == 1 << 7
    assert V.ShrunkDecode(rs1=0xFF).pack() == 0xFF << 10 and V.ShrunkDecode(rd=0xAB).pack() == 0xAB << 26
    assert V.ShrunkDecode(half_instruction_length=2).pack() == 2 << 34
    assert V.ShrunkDecode(alu_flags=0xFF).pack() == 0xFF << 42 and V.ShrunkDecode(mem_flags=0xFF).pack() == 0xFF << 50
    full = V.ShrunkDecode(*([True] * 10), 255, 255, 255, 255, 255, 255).pack()
    assert full >> 58 == 0 and full == (1 << 58) - 1
    d = V.shrunk_decode(V.Instruction("arith", 3, 1, 2, op="add"))
    assert d.add and not d.alu and not d.sub and (d.rs1, d.rs2, d.rd, d.half_instruction_length) == (1, 2, 3, 2)
    assert d.read_register1 and d.read_register2 and d.write_register and d.alu_flags == 0
    d = V.shrunk_decode(V.Instruction("arith", 5, 6, 7, op="and"))
    assert d.alu and d.alu_flags == V.ALU_OP_AND
    d = V.shrunk_decode(V.Instruction("arith_w", 1, 2, 3, op="sra"))
    assert d.word_instr and d.alu_flags == V.ALU_OP_SHIFTW + 32 + 64
    d = V.shrunk_decode(V.Instruction("branch", 0, 1, 2, imm=8, cond="geu"))
    assert d.branch and d.alu and d.alu_flags == V.ALU_OP_LT + 64 and not d.write_register
    d = V.shrunk_decode(V.Instruction("jal", 1, imm=-4))
    assert d.branch and d.rs1 == 255 and d.read_register1 and d.mem_flags == 1 and not d.alu
    d = V.shrunk_decode(V.Instruction("load", 4, 2, imm=0, width="hu"))
    assert d.memory and d.add and d.mem_flags == 4 and d.write_register
    d = V.shrunk_decode(V.Instruction("store", 0, 2, 9, imm=0, width="w"))
    assert d.memory and d.mem_flags == 1 + 8 and d.read_register2 and not d.write_register
    d = V.shrunk_decode(V.Instruction("ecall"))
    assert d.ecall and d.rs1 == 17 and d.read_register1 and not d.write_register
    assert V.shrunk_decode(V.Instruction("fence")) == V.ShrunkDecode(add=True, half_instruction_length=2)
    assert V.instruction_imm(V.Instruction("lui", 1, imm=0xFFFFF000)) == M64 - 0xFFF
    assert V.instruction_imm(V.Instruction("store", 0, 1, 2, imm=-8)) == M64 - 7


def test_decode_rows_follow_the_reference_layout():
    prog = random_program(13, 2)
    entries = V.decode_entries_from_words([(0x1000, [w for w, _ in prog])])
    t = V.generate_decode_trace(entries, [0x1000, 0x1004, 0x1004, V.CPU_PADDING_PC, 0x999])
    c = V.DecodeCols
    assert t.shape == (16, 6)
    assert list(t[:3, c.MU]) == [1, 2, 0] and t[13, c.MU] == 1 and t[:, c.MU].sum() == 4
    assert all(int(t[r, c.PC_0]) == 0x1000 + 4 * r for r in range(13))
    assert (t[13:, c.PC_0] == 1).all() and not t[13:, c.PACKED_DECODE:c.MU].any()
    assert int(t[3, c.IMM_0]) | int(t[3, c.IMM_1]) << 32 == V.instruction_imm(prog[3][1])


def word_program(k, seed):
    """Word instructions with their CPU32 rows (fields from the decoded words)."""
    rng = np.random.default_rng(seed)
    words, ops = [], []
    for i in range(k):
        rd, rs1, rs2 = int(rng.integers(1, 32)), int(rng.integers(1, 32)), int(rng.integers(1, 32))
        if i % 3 == 2:
            op, f3, hi = [("add", 0, None), ("sll", 1, 0), ("srl", 5, 0), ("sra", 5, 0x400)][i % 4]
            imm = int(rng.integers(-2048, 2048)) if hi is None else int(rng.integers(0, 32))
            word = enc_i(0b0011011, f3, rd, rs1, imm if hi is None else hi | imm)
        else:
            op = sorted(V._W_OPS)[i % 10]
            f3, f7 = R_FUNCS[op]
            word = enc_r(0b0111011, f3, f7, rd, rs1, rs2)
        ins = V.decode_instruction(word)
        pc = 0x2000 + 4 * i
        operand = ins.imm if ins.kind == "arith_imm_w" else int(rng.integers(0, 2**64, dtype=np.uint64))
        op_row = make_op(ins.op, int(rng.integers(0, 2**64, dtype=np.uint64)), operand, ins.kind == "arith_imm_w",
                         ts=4 * (i + 1), pc=pc, rs1=ins.src1, rs2=ins.src2, rd=ins.dst)
        words.append(word)
        ops.append(op_row)
    return words, ops


def test_cpu32_decode_lookups_balance_against_decode():
    words, ops = word_program(60, 3)
    entries = V.decode_entries_from_words([(0x2000, words)])
    cpu = A.logup_air(V.Cpu32Cols.NUM_COLUMNS, [it for it in V.cpu32_bus_interactions() if it.bus_id == V.BUS_DECODE])
    dec = V.decode_table_air(entries)
    for a in (cpu, dec):
        a.bus_descriptor = A.descriptor(a)

    def total(ops):
        tables = [(cpu, V.generate_cpu32_trace(ops)), (dec, V.generate_decode_trace(entries, [o.pc for o in ops]))]
        tot = [0, 0, 0]
        for o in replay_air_tables(tables, 2, 3, 3, 2, 0, b"decode-bus"):
            tot = [(x + int(y)) % P for x, y in zip(tot, o["table_contribution"])]
        return tuple(tot)

    assert total(ops) == (0, 0, 0)
    bad = list(ops)
    bad[5] = V.Cpu32Operation(**{**bad[5].__dict__, "rd": (bad[5].rd % 31) + 1})
    assert total(bad) != (0, 0, 0)


def test_decode_proof_verifies_with_its_fixed_commitment():
    prog = random_program(50, 4)
    entries = V.decode_entries_from_words([(0x1000, [w for w, _ in prog])])
    dec = V.decode_table_air(entries)
    dec.bus_descriptor = A.descriptor(dec)
    tables = [(dec, V.generate_decode_trace(entries, [0x1000 + 4 * (i % 50) for i in range(300)]))]
    proofs = replay_air_tables(tables, 4, 3, 2, 6, 0, b"decode")
    assert proofs[0]["precomputed_root"] == precomputed_commitment(dec, 4, 3)
    assert verify_air_tables([dec], proofs, 4, 3, 2, 6, 0, b"decode")[0]
    other = V.decode_table_air(entries[:-1] + [(entries[-1][0], entries[-1][1] ^ 1, entries[-1][2])])
    other.bus_descriptor = A.descriptor(other)
    assert verify_air_tables([other], proofs, 4, 3, 2, 6, 0, b"decode") == (False, "precomputed commitment")
