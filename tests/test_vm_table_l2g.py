"""The reference VM's LOCAL_TO_GLOBAL table (prover/src/tables/local_to_global.rs,
continuation.rs:150-200): the per-epoch memory boundary claims of a
continuation, 9 columns, an epoch-local AIR (Memory bus, range and ordering
checks) and a global AIR (GlobalMemory bus) over the same trace.

Pins independent of the restatement: over a multi-epoch run whose cells are
written in several epochs (some never after genesis), each init claim names
the previous writer's epoch and value, read back from a plain replay of the
writes kept here; on the GlobalMemory bus the L2G tables of every epoch
close, in one transcript, against GLOBAL_MEMORY's genesis and final tokens
of the page (a small page), and stop closing for a wrong fini value; inside
an epoch, the L2G Memory-bus tokens equal PAGE's for the same touched cells
(the same bookend, sparse instead of dense); the ordering value of an honest
row is a 20-bit value and that of a backward edge is not; mu a bit. Proofs of
both AIRs over one trace verify."""
import numpy as np

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from oracle.air_proof import replay_air_tables, verify_air_tables
from test_vm_table_dvrm import bus_messages
from test_vm_table_eq import _constraint_values

P = 0xFFFFFFFF00000001
PAGE = 0x7000_0000
SIZE = 256


def scenario(seed, epochs=4):
    """Initial bytes of a small page and, per epoch, the cells it touches with
    their end value and end timestamp."""
    rng = np.random.default_rng(seed)
    init = [int(v) for v in rng.integers(0, 256, SIZE)]
    touched = []
    for e in range(epochs):
        cells = sorted(set(int(x) for x in rng.integers(0, SIZE // 2, 24)))
        touched.append([(PAGE + c, int(rng.integers(0, 256)), int(rng.integers(1, 1 << 20))) for c in cells])
    return init, touched


def test_boundaries_follow_the_last_writer():
    init, touched = scenario(1)
    bds = V.epoch_boundaries({PAGE + i: v for i, v in enumerate(init)}, touched)
    last = {PAGE + i: (0, v) for i, v in enumerate(init)}
    for e, (rows, cells) in enumerate(zip(bds, touched)):
        assert [b.address for b in rows] == [a for a, _, _ in cells]
        for b, (a, v, ts) in zip(rows, cells):
            assert (b.init_epoch, b.init_value) == last[a] and (b.fini_epoch, b.fini_value, b.fini_timestamp) == \
                (e + 1, v, ts)
            last[a] = (e + 1, v)
    tr = V.generate_l2g_trace(bds[1])
    assert tr.shape[0] == 1 << (len(bds[1]) - 1).bit_length() and tr[:len(bds[1]), V.L2gCols.MU].all()
    assert V.generate_l2g_trace([]).shape == (1, 9)


def test_global_memory_bus_closes_over_the_epochs():
    init, touched = scenario(2)
    bds = V.epoch_boundaries({PAGE + i: v for i, v in enumerate(init)}, touched)
    final = {}
    for rows in bds:
        for b in rows:
            final[b.address] = (b.fini_value, b.fini_epoch)
    gm = V.global_memory_table_air(PAGE, init, page_size=SIZE)
    gm.bus_descriptor = A.descriptor(gm)
    gm_tr = V.generate_global_memory_trace(PAGE, init, final, page_size=SIZE)

    def total(epoch_rows):
        tables = [(gm, gm_tr)]
        for e, rows in enumerate(epoch_rows):
            air = V.l2g_global_air(e + 1)
            air.bus_descriptor = A.descriptor(air)
            tables.append((air, V.generate_l2g_trace(rows)))
        tot = [0, 0, 0]
        for o in replay_air_tables(tables, 2, 3, 2, 2, 0, b"l2g-global"):
            tot = [(x + int(y)) % P for x, y in zip(tot, o["table_contribution"])]
        return tuple(tot)

    assert total(bds) == (0, 0, 0)
    bad = [list(r) for r in bds]
    b = bad[1][3]
    bad[1][3] = V.CellBoundary(b.address, b.init_value, b.init_epoch, b.init_timestamp, b.fini_value ^ 1, b.fini_epoch,
                               b.fini_timestamp)
    assert total(bad) != (0, 0, 0)


def test_epoch_memory_tokens_equal_page_bookends(ora):
    init, touched = scenario(3, epochs=1)
    rows = V.epoch_boundaries({PAGE + i: v for i, v in enumerate(init)}, touched)[0]
    air = V.l2g_memory_air(1)
    tr = V.generate_l2g_trace(rows)
    got = bus_messages([it for it in air.interactions if it.bus_id == V.BUS_MEMORY], tr, range(tr.shape[0]))
    page_tr = V.generate_page_trace(PAGE, init, {a: (ts, v) for a, v, ts in touched[0]}, page_size=SIZE)
    page_its = [it for it in V.page_bus_interactions(PAGE) if it.bus_id == V.BUS_MEMORY]
    want = bus_messages(page_its, page_tr, [a - PAGE for a, _, _ in touched[0]])
    assert got == want and len(got) == 2 * len(rows)
    # ordering: the IS_B20 value of every honest row is in range, a backward edge's is not
    its = V.l2g_memory_bus_interactions(1)
    b20 = bus_messages([it for it in its if it.bus_id == V.BUS_IS_B20], tr, range(len(rows)))
    assert all(msg[1] < 1 << 20 for msg in b20)
    back = V.generate_l2g_trace([V.CellBoundary(PAGE, 1, 3, 0, 2, 1, 5)])
    msg = next(iter(bus_messages([it for it in its if it.bus_id == V.BUS_IS_B20], back, [0])))
    assert msg[1] >= 1 << 20
    assert not np.any(_constraint_values(ora, air, tr) % np.uint64(P))
    bad = tr.copy()
    bad[0, V.L2gCols.MU] = 2
    assert np.any(_constraint_values(ora, air, bad)[0, 0] % np.uint64(P))


def test_l2g_proofs_verify():
    init, touched = scenario(4, epochs=3)
    rows = V.epoch_boundaries({PAGE + i: v for i, v in enumerate(init)}, touched)[2]
    tr = V.generate_l2g_trace(rows)
    mem, glob = V.l2g_memory_air(3), V.l2g_global_air(3)
    for a in (mem, glob):
        a.bus_descriptor = A.descriptor(a)
    proofs = replay_air_tables([(mem, tr), (glob, tr)], 4, 3, 2, 6, 0, b"l2g")
    assert verify_air_tables([mem, glob], proofs, 4, 3, 2, 6, 0, b"l2g")[0]
