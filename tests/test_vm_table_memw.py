"""The reference VM's MEMW table (prover/src/tables/memw.rs): the general
memory access — unaligned, cells with different old timestamps — 49 columns,
15 constraints, 26 interactions.

A memory model replays unaligned reads and writes of widths 1, 2, 4 and 8 on
two pages that meet at a 2^32 boundary of the address's low word (accesses
straddle it, so the per-cell carry bits are exercised), at random offsets, so
the cells of one access usually carry different old timestamps. In one
transcript the MEMW rows' Memory tokens balance against the two PAGE tables'
initial and final tokens, and the per-cell old_ts < ts checks against the LT
table; a wrong carry bit, a wrong old timestamp or a dropped access
unbalances the bus. Constraints vanish on generated traces and break for a
non-bit carry; a proof with both PAGE tables replays and verifies."""
import numpy as np

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from oracle.air_proof import replay_air_tables, verify_air_tables
from test_vm_table_eq import _constraint_values

P = 0xFFFFFFFF00000001
PAGE = 1024
BASES = (0x5_FFFF_FC00, 0x6_0000_0000)   # two pages meeting at a 2^32 low-word boundary
INITS = (bytes(range(3, 200)), b"")


def unaligned_chain(k, seed):
    rng = np.random.default_rng(seed)
    cells, ops, ts = {}, [], 0

    def cell(a):
        for base, init in zip(BASES, INITS):
            if base <= a < base + PAGE:
                off = a - base
                return cells.get(a, (0, init[off] if off < len(init) else 0))
        raise AssertionError(hex(a))

    for i in range(k):
        w = int(rng.choice([1, 2, 4, 8]))
        if i % 4 == 0:
            addr = BASES[1] - int(rng.integers(1, 8))  # straddles the boundary
        else:
            addr = BASES[0] + PAGE - 64 + int(rng.integers(0, 120))
        addr = min(addr, BASES[1] + PAGE - w)
        ts += int(rng.integers(1, 30))
        old = [cell(addr + j) for j in range(w)]
        is_read = bool(rng.integers(0, 2))
        new = [v for _, v in old] if is_read else [int(x) for x in rng.integers(0, 256, w)]
        for j in range(w):
            cells[addr + j] = (ts, new[j])
        pad = [0] * (8 - w)
        ops.append(V.MemwGeneralOperation(False, addr, tuple(new + pad), ts, w, is_read,
                                          tuple([v for _, v in old] + pad), tuple([t for t, _ in old] + pad)))
    finals = [{a: tv for a, tv in cells.items() if b <= a < b + PAGE} for b in BASES]
    return ops, finals


def _desc(a):
    a.bus_descriptor = A.descriptor(a)
    return a


def _total(tables, label):
    out = replay_air_tables(tables, 2, 3, 2, 4, 0, label)
    tot = [0, 0, 0]
    for o in out:
        tot = [(x + int(y)) % P for x, y in zip(tot, o["table_contribution"])]
    return tuple(tot)


def _pages(finals, memory_only=True):
    out = []
    for base, init, fin in zip(BASES, INITS, finals):
        a = _desc(A.logup_air(V.PageCols.NUM_COLUMNS, V.page_bus_interactions(base)[1:])) if memory_only else \
            _desc(V.page_table_air(base, init, page_size=PAGE))
        out.append((a, V.generate_page_trace(base, init, fin, PAGE)))
    return out


def test_memw_rows_and_constraints(ora):
    ops, _ = unaligned_chain(40, 1)
    tr = V.generate_memw_trace(ops)
    c = V.MemwCols
    straddle = [r for r, o in enumerate(ops) if (o.base_address & 0xFFFFFFFF) + o.width > 1 << 32]
    assert straddle
    r = straddle[0]
    lo = ops[r].base_address & 0xFFFFFFFF
    assert [int(tr[r, col]) for col in c.CARRY] == [int(lo + i + 1 >= 1 << 32) for i in range(7)]
    air = V.memw_table_air()
    assert len(air.interactions) == 26
    assert not np.any(_constraint_values(ora, air, tr) % np.uint64(P))
    bad = tr.copy()
    bad[r, c.CARRY[2]] = 2
    assert np.any(_constraint_values(ora, air, bad)[6, r] % np.uint64(P))


def test_memory_argument_across_a_word_boundary():
    ops, finals = unaligned_chain(80, 2)
    its = V.memw_bus_interactions()
    memw_mem = _desc(A.logup_air(V.MemwCols.NUM_COLUMNS, its[:16]))
    pages = _pages(finals)
    tr = V.generate_memw_trace(ops)
    assert _total([(memw_mem, tr)] + pages, b"memw") == (0, 0, 0)
    c = V.MemwCols
    r = next(r for r, o in enumerate(ops) if (o.base_address & 0xFFFFFFFF) + o.width > 1 << 32)
    k = next(i for i in range(7) if tr[r, c.CARRY[i]] == 1)
    bad = tr.copy()
    bad[r, c.CARRY[k]] = 0  # the cell lands on the wrong side of the boundary
    assert _total([(memw_mem, bad)] + pages, b"memw") != (0, 0, 0)
    bad = tr.copy()
    bad[r, c.old_timestamp(0)] += 1
    assert _total([(memw_mem, bad)] + pages, b"memw") != (0, 0, 0)
    assert _total([(memw_mem, V.generate_memw_trace(ops[1:]))] + pages, b"memw") != (0, 0, 0)
    # per-cell old_ts < ts through the LT table
    memw_lt = _desc(A.logup_air(V.MemwCols.NUM_COLUMNS, its[18:26]))
    lt_recv = _desc(A.logup_air(V.LtCols.NUM_COLUMNS, V.lt_bus_interactions()[-1:]))
    lts = [V.LtOperation(o.old_timestamp[j], o.timestamp, False, False) for o in ops for j in range(o.width)]
    assert _total([(memw_lt, tr), (lt_recv, V.generate_lt_trace(lts))], b"memw-lt") == (0, 0, 0)


def test_memw_proof_verifies():
    ops, finals = unaligned_chain(30, 3)
    memw = _desc(V.memw_table_air())
    pages = _pages(finals, memory_only=False)
    tables = [(memw, V.generate_memw_trace(ops))] + pages
    proofs = replay_air_tables(tables, 4, 3, 2, 6, 0, b"memw")
    assert verify_air_tables([a for a, _ in tables], proofs, 4, 3, 2, 6, 0, b"memw")[0]
