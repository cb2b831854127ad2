"""The reference VM's HALT table (prover/src/tables/halt.rs): one row, 4
columns, 36 interactions, no constraints of its own.

Pins independent of the restatement: the Ecall receiver matches the ECALL
send of a CPU row executing sys_exit (a7 = 93) in one transcript; on the
Memory bus the pc token the halting row writes is consumed and replaced by the
padding pc 1, which the CPU padding rows carry on — the pc chain of a CPU
trace that halts and pads closes against HALT between the first row's token
and the last padding row's; the 31 register finalization sends equal, as a
multiset, what MEMW_REGISTER receives for writes of 0 to x1-x9 and x11-x31 and
a read of x10 holding 0 at timestamp 2^32 - 1, and differ when x10 holds a
nonzero exit code. The one-row proof replays and verifies on the CPU; the GPU
entry points take traces of at least two rows."""
from collections import Counter

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from oracle.air_proof import replay_air_tables, verify_air_tables
from test_vm_table_decode import enc_i
from test_vm_table_dvrm import bus_messages

P = 0xFFFFFFFF00000001


def halting_cpu():
    """A CPU trace: addi a7, x0, 93; addi x1, x0, 5; ecall; then padding rows."""
    words = [enc_i(0b0010011, 0, 17, 0, 93), enc_i(0b0010011, 0, 1, 0, 5), 0x00000073]
    pc, x, ops = 0x1000, [0] * 32, []
    for i, w in enumerate(words):
        ins = V.decode_instruction(w)
        a, b = x[ins.src1], x[ins.src2]
        dst = (a + V.instruction_imm(ins)) % 2**64 if ins.kind == "arith_imm" else 0
        if ins.kind == "ecall":
            a = x[17]
        ops.append(V.cpu_operation(pc, ins, 4 * (i + 1), a, b, dst, pc + 4))
        if ins.kind == "arith_imm" and ins.dst:
            x[ins.dst] = dst
        pc += 4
    return ops


def test_halt_closes_the_ecall_and_pc_chain_of_a_halting_cpu():
    ops = halting_cpu()
    cpu_tr = V.generate_cpu_trace(ops)          # 4 rows: three instructions and one padding row
    halt_ts, halt_next_pc = ops[-1].timestamp, ops[-1].next_pc
    halt_tr = V.generate_halt_trace(halt_ts, halt_next_pc)
    cpu_ecall = A.logup_air(V.CpuCols.NUM_COLUMNS, [it for it in V.cpu_bus_interactions() if it.bus_id == V.BUS_ECALL])
    halt_ecall = A.logup_air(V.HaltCols.NUM_COLUMNS, [it for it in V.halt_bus_interactions() if it.bus_id == V.BUS_ECALL])
    for a in (cpu_ecall, halt_ecall):
        a.bus_descriptor = A.descriptor(a)
    tot = [0, 0, 0]
    for o in replay_air_tables([(cpu_ecall, cpu_tr), (halt_ecall, halt_tr)], 2, 3, 0, 2, 0, b"halt-ecall"):
        tot = [(x + int(y)) % P for x, y in zip(tot, o["table_contribution"])]
    assert tuple(tot) == (0, 0, 0)
    # the pc chain: CPU rows and HALT telescope to the first row's token and the last padding row's
    cpu_mem = [it for it in V.cpu_bus_interactions() if it.bus_id == V.BUS_MEMORY]
    halt_mem = [it for it in V.halt_bus_interactions() if it.bus_id == V.BUS_MEMORY]
    net = bus_messages(cpu_mem, cpu_tr, range(cpu_tr.shape[0]))
    net.update(bus_messages(halt_mem, halt_tr, [0]))
    net = {k: v for k, v in net.items() if v}
    last_ts = 4 * cpu_tr.shape[0] + 1
    assert net == {(V.BUS_MEMORY, 1, 510, 0, 1, 0, 0x1000): 1, (V.BUS_MEMORY, 1, 511, 0, 1, 0, 0): 1,
                   (V.BUS_MEMORY, 1, 510, 0, last_ts, 0, V.CPU_PADDING_PC): -1,
                   (V.BUS_MEMORY, 1, 511, 0, last_ts, 0, 0): -1}


def test_halt_register_finalization_matches_memw_register():
    halt_tr = V.generate_halt_trace(40, 0x1010)
    sends = bus_messages([it for it in V.halt_bus_interactions() if it.bus_id == V.BUS_MEMW], halt_tr, [0])
    ts = V.HALT_FINAL_TIMESTAMP

    def receivers(exit_code):
        acc = [V.RegisterAccess(2 * r, ts, 0, exit_code if r == 10 else 7 * r, ts - 1, r == 10) for r in range(1, 32)]
        tr = V.generate_memw_register_trace(acc)
        its = [it for it in V.memw_register_bus_interactions() if it.bus_id == V.BUS_MEMW]
        return bus_messages(its, tr, range(len(acc)))

    got = Counter(sends)
    got.subtract(Counter({k: -v for k, v in receivers(0).items()}))
    assert not +got and not -got
    got = Counter(sends)
    got.subtract(Counter({k: -v for k, v in receivers(3).items()}))
    assert +got or -got
    assert len(V.halt_bus_interactions()) == 36


def test_halt_one_row_proof_verifies():
    air = V.halt_table_air()
    air.bus_descriptor = A.descriptor(air)
    tables = [(air, V.generate_halt_trace(4 * 1000, 0x2000))]
    proofs = replay_air_tables(tables, 4, 3, 0, 4, 0, b"halt")
    assert verify_air_tables([air], proofs, 4, 3, 0, 4, 0, b"halt")[0]
