"""The reference VM's MEMW_A table (prover/src/tables/memw_aligned.rs): aligned
memory accesses, 29 columns, 8 constraints, 20 interactions.

The memory argument across three captured tables, in one transcript: a chain
of aligned reads and writes on a page (each slot always accessed at one width,
so its cells share a timestamp), replayed by a small memory model, gives
MEMW_A rows whose Memory tokens balance against PAGE's initial and final
tokens for the same page; the writes' MEMW receivers balance against the
STORE table's MEMW sends for the same stores, and the old_ts < ts checks
against the LT table. A wrong written byte, a missing write or a moved
timestamp unbalances the bus. Constraints vanish on the generated trace and
break for two width flags or a non-bit multiplicity; a proof of MEMW_A with
PAGE replays and verifies."""
import numpy as np
import pytest

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from oracle.air_proof import replay_air_tables, verify_air_tables
from test_vm_table_eq import _constraint_values

P = 0xFFFFFFFF00000001
BASE = 0x2_0000_0400 & ~1023
PAGE = 1024


def access_chain(k, seed, init=b""):
    """k aligned accesses on the page: the MemwOperations, the stores behind the
    writes, and the final (timestamp, value) of every touched byte."""
    rng = np.random.default_rng(seed)
    slots = {int(s) * 8: int(w) for s, w in zip(rng.choice(PAGE // 8, 24, replace=False),
                                                  rng.choice([1, 2, 4, 8], 24))}
    cells = {}
    ops, stores, ts = [], [], 0

    def cell(a):
        off = a - BASE
        return cells.get(a, (0, init[off] if off < len(init) else 0))

    for _ in range(k):
        off = int(rng.choice(list(slots)))
        w, addr = slots[off], BASE + off
        ts += int(rng.integers(1, 40))
        old = [cell(addr + i) for i in range(w)]
        assert len({t for t, _ in old}) == 1
        is_read = bool(rng.integers(0, 2))
        new = [v for _, v in old] if is_read else [int(x) for x in rng.integers(0, 256, w)]
        for i in range(w):
            cells[addr + i] = (ts, new[i])
        pad = [0] * (8 - w)
        ops.append(V.MemwOperation(False, addr, tuple(new + pad), ts, w, is_read, tuple([v for _, v in old] + pad),
                                   old[0][0]))
        if not is_read:
            stores.append(V.StoreOperation.of_width(addr, ts, int.from_bytes(bytes(new), "little"), w))
    return ops, stores, cells


def _desc(a):
    a.bus_descriptor = A.descriptor(a)
    return a


def _total(tables, label):
    out = replay_air_tables(tables, 2, 3, 2, 4, 0, label)
    tot = [0, 0, 0]
    for o in out:
        tot = [(x + int(y)) % P for x, y in zip(tot, o["table_contribution"])]
    return tuple(tot)


INIT = bytes(range(7, 250))


def test_memw_aligned_rows_and_constraints(ora):
    ops, _, _ = access_chain(30, 1, INIT)
    tr = V.generate_memw_aligned_trace(ops)
    c = V.MemwAlignedCols
    assert tr.shape == (32, 29)
    op = ops[4]
    assert int(tr[4, c.BASE_ADDRESS[0]]) | int(tr[4, c.BASE_ADDRESS[1]]) << 16 | int(tr[4, c.BASE_ADDRESS[2]]) << 32 \
        == op.base_address
    assert tuple(tr[4, c.WRITE2:c.WRITE8 + 1]) == op.write_flags()
    air = V.memw_aligned_table_air()
    assert len(air.interactions) == 20
    assert not np.any(_constraint_values(ora, air, tr) % np.uint64(P))
    for col, val, idx in ((c.WRITE4, 1, 7), (c.MU_READ, 2, 0)):
        bad = tr.copy()
        row = next(i for i, o in enumerate(ops) if o.width == 2) if col == c.WRITE4 else 3
        bad[row, col] = val
        assert np.any(_constraint_values(ora, air, bad)[idx, row] % np.uint64(P)), (col, idx)


def test_memory_argument_page_memw_aligned_store():
    ops, stores, cells = access_chain(60, 2, INIT)
    final = {a: tv for a, tv in cells.items()}
    page_mem = _desc(A.logup_air(V.PageCols.NUM_COLUMNS, V.page_bus_interactions(BASE)[1:]))
    mem = V.memw_aligned_bus_interactions()
    memw_mem = _desc(A.logup_air(V.MemwAlignedCols.NUM_COLUMNS, mem[2:18]))
    page_tr = V.generate_page_trace(BASE, INIT, final, PAGE)
    assert _total([(page_mem, page_tr), (memw_mem, V.generate_memw_aligned_trace(ops))], b"mem") == (0, 0, 0)
    # a written byte that differs from what the page finally holds
    w = next(i for i, o in enumerate(ops) if not o.is_read)
    bad = list(ops)
    bad[w] = V.MemwOperation(False, ops[w].base_address, (ops[w].value[0] ^ 1,) + ops[w].value[1:],
                             ops[w].timestamp, ops[w].width, False, ops[w].old, ops[w].old_timestamp)
    assert _total([(page_mem, page_tr), (memw_mem, V.generate_memw_aligned_trace(bad))], b"mem") != (0, 0, 0)
    # a dropped access breaks the chain
    assert _total([(page_mem, page_tr), (memw_mem, V.generate_memw_aligned_trace(ops[:-1]))], b"mem") != (0, 0, 0)
    # the writes' MEMW receivers against the STORE table's sends
    memw_w = _desc(A.logup_air(V.MemwAlignedCols.NUM_COLUMNS, mem[19:20]))
    store_s = _desc(A.logup_air(V.StoreCols.NUM_COLUMNS, V.store_bus_interactions()[:1]))
    assert _total([(memw_w, V.generate_memw_aligned_trace(ops)), (store_s, V.generate_store_trace(stores))],
                  b"memw") == (0, 0, 0)
    moved = [V.StoreOperation(s.base_address, s.timestamp + (i == 0), s.value, s.write2, s.write4, s.write8)
             for i, s in enumerate(stores)]
    assert _total([(memw_w, V.generate_memw_aligned_trace(ops)), (store_s, V.generate_store_trace(moved))],
                  b"memw") != (0, 0, 0)
    # old_ts < ts through the LT table
    memw_lt = _desc(A.logup_air(V.MemwAlignedCols.NUM_COLUMNS, mem[1:2]))
    lt_recv = _desc(A.logup_air(V.LtCols.NUM_COLUMNS, V.lt_bus_interactions()[-1:]))
    lts = [V.LtOperation(o.old_timestamp, o.timestamp, False, False) for o in ops]
    assert _total([(memw_lt, V.generate_memw_aligned_trace(ops)), (lt_recv, V.generate_lt_trace(lts))],
                  b"memw-lt") == (0, 0, 0)


def test_memw_aligned_proof_verifies():
    ops, _, cells = access_chain(40, 3, INIT)
    memw = _desc(V.memw_aligned_table_air())
    page = _desc(V.page_table_air(BASE, INIT, page_size=PAGE))
    tables = [(memw, V.generate_memw_aligned_trace(ops)), (page, V.generate_page_trace(BASE, INIT, cells, PAGE))]
    proofs = replay_air_tables(tables, 4, 3, 2, 6, 0, b"memw-a")
    assert verify_air_tables([memw, page], proofs, 4, 3, 2, 6, 0, b"memw-a")[0]
