"""The reference VM's COPY table (prover/src/tables/copy.rs): the streaming copy
of memcpy/memmove, memset and write(2), 41 columns, 40 constraints (degree 3),
33 interactions.

Pins independent of the restatement: the memory the rows' MEMW writes leave
behind equals Python's slice semantics (memmove copies through a snapshot even
when the ranges overlap, memset from an 8-byte seed fills the range, write
changes no memory) and the Commit bus carries exactly the written bytes at
consecutive indices from the running commit index; calls of 0, 1, 7, 8, 9, 23
and 64 bytes. Constraints vanish on the generated trace (padding included) and
break for an end flag on a row with bytes left, a wrong single flag and a wrong
count - step. The CopyNext messages cancel within the table (each call a
closed chain), and the pointer reads and the dst/dst_res read-write of x10/x254 balance on the
MEMW bus against MEMW_REGISTER. A proof replays and verifies."""
from collections import Counter

import numpy as np

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from oracle.air_proof import replay_air_tables, verify_air_tables
from test_vm_table_dvrm import bus_messages
from test_vm_table_eq import _constraint_values

P = 0xFFFFFFFF00000001
BASE = 0x5_0000


def calls(seed):
    """(functionality, timestamp, src, dst, count) calls over a byte memory, and
    the memory: copies (disjoint and overlapping both ways), memsets (seeded
    8 bytes at src, dst = src + 8) and writes at a running commit index."""
    rng = np.random.default_rng(seed)
    mem = {BASE + i: int(v) for i, v in enumerate(rng.integers(0, 256, 512))}
    out, ts, commit = [], 100, 0
    for count in (0, 1, 7, 8, 9, 23, 64):
        src = BASE + int(rng.integers(0, 200))
        out.append(("copy", ts, src, BASE + 300 + int(rng.integers(0, 100)), count))
        out.append(("copy", ts + 4, src, src + int(rng.integers(1, 12)), count))       # overlapping, forward
        out.append(("copy", ts + 8, src + 5, src, count))                            # overlapping, backward
        out.append(("set", ts + 12, src, src + 8, count))
        out.append(("write", ts + 16, src, commit, count))
        commit += count
        ts += 20
    return out, mem


def run(seed):
    """The rows of every call and the memory model the test keeps itself."""
    cl, mem = calls(seed)
    model = dict(mem)
    rows, committed = [], []
    for f, ts, src, dst, count in cl:
        rows += V.copy_call_rows(f, ts, src, dst, count, mem)
        data = [model.get(src + i, 0) for i in range(count)]
        if f == "copy":
            for i, v in enumerate(data):
                model[dst + i] = v
        elif f == "set":
            seed_bytes = [model.get(src + i, 0) for i in range(8)]
            for i in range(count):
                model[dst + i] = seed_bytes[i % 8]
        else:
            committed += [(dst + i, v) for i, v in enumerate(data)]
    return cl, rows, mem, model, committed


def test_copy_semantics():
    cl, rows, mem, model, committed = run(1)
    assert mem == model
    tr = V.generate_copy_trace(rows)
    its = V.copy_bus_interactions()
    got = bus_messages([it for it in its if it.bus_id == V.BUS_COMMIT], tr, range(tr.shape[0]))
    assert dict(got) == {(V.BUS_COMMIT, i, v): 1 for i, v in committed}
    # row shape: 8-byte rows, then single bytes; the zero-length call is one row
    for f, ts, src, dst, count in cl:
        widths = [r.width for r in rows if r.timestamp == ts and r.functionality == f]
        assert widths == ([8] * (count // 8) + [1] * (count % 8) if count else [1]), (f, count)


def test_copy_constraints(ora):
    _, rows, _, _, _ = run(2)
    air = V.copy_table_air()
    tr = V.generate_copy_trace(rows)
    assert (air.num_main, len(air.interactions)) == (41, 33)
    assert not np.any(_constraint_values(ora, air, tr) % np.uint64(P))
    c = V.CopyCols
    mid = next(i for i, r in enumerate(rows) if not r.end and r.width == 8 and r.count > 8)
    for col, row, val, idx in ((c.END, mid, 1, slice(32, 35)), (c.SINGLE, mid, 1, slice(10, 29)),
                               (c.COUNT_DECR, mid, 5, slice(14, 16))):
        bad = tr.copy()
        bad[row, col] = val
        assert np.any(_constraint_values(ora, air, bad)[idx, row] % np.uint64(P)), col


def test_copy_next_chains_close_and_registers_balance():
    cl, rows, _, _, _ = run(3)
    tr = V.generate_copy_trace(rows)
    its = V.copy_bus_interactions()
    chain = bus_messages([it for it in its if it.bus_id == V.BUS_COPY_NEXT], tr, range(tr.shape[0]))
    assert not {k: v for k, v in chain.items() if v}
    acc = []
    for f, ts, src, dst, count in cl:
        acc.append(V.RegisterAccess(22, ts, src, src, ts - 1, True))
        acc.append(V.RegisterAccess(24, ts, count, count, ts - 1, True))
        res = dst + count if f == "write" else dst
        acc.append(V.RegisterAccess(508 if f == "write" else 20, ts, res, dst, ts - 1, True))
        if f == "write":
            acc.append(V.RegisterAccess(20, ts, count, 1, ts - 1, True))
    regs = [it for it in its if it.bus_id == V.BUS_MEMW and it.multiplicity.kind == "Column"]
    assert len(regs) == 4
    cp = A.logup_air(V.CopyCols.NUM_COLUMNS, regs)
    mem = A.logup_air(V.MemwRegisterCols.NUM_COLUMNS, [it for it in V.memw_register_bus_interactions()
                                                       if it.bus_id == V.BUS_MEMW and not it.is_sender])
    for a in (cp, mem):
        a.bus_descriptor = A.descriptor(a)
    tot = [0, 0, 0]
    for o in replay_air_tables([(cp, tr), (mem, V.generate_memw_register_trace(acc))], 2, 3, 3, 2, 0, b"copy-regs"):
        tot = [(x + int(y)) % P for x, y in zip(tot, o["table_contribution"])]
    assert tuple(tot) == (0, 0, 0)
    # the RAM accesses: one read and one write per moving row, none for zero-length rows
    ram = bus_messages([it for it in its if it.bus_id == V.BUS_MEMW and it.multiplicity.kind == "Linear"], tr,
                       range(tr.shape[0]))
    moving = [r for r in rows if r.count]
    reads = sum(1 for r in moving)
    writes = sum(1 for r in moving if r.functionality != "write")
    assert sum(v for v in ram.values()) == reads + writes and all(v > 0 for v in ram.values())
    assert Counter(len(k) for k in ram.elements()) == Counter({25: reads, 17: writes})


def test_copy_proof_verifies():
    _, rows, _, _, _ = run(4)
    air = V.copy_table_air()
    air.bus_descriptor = A.descriptor(air)
    tables = [(air, V.generate_copy_trace(rows))]
    proofs = replay_air_tables(tables, 4, 3, 2, 6, 0, b"copy")
    assert verify_air_tables([air], proofs, 4, 3, 2, 6, 0, b"copy")[0]
