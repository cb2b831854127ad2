"""The reference VM's PAGE and GLOBAL_MEMORY tables (prover/src/tables/page.rs,
global_memory.rs) as preprocessed AIRs.

Pins: the fixed OFFSET + INIT columns of a zero-initialised 2^18-byte page, and
OFFSET alone (a private-input page, and GLOBAL_MEMORY's), committed by the
oracle, give the reference's static commitments (page.rs:416-455, extracted
to tests/golden/kats.json) for blowups 2, 4 and 8 at coset offset 3.
Buses: an untouched page's initial and final Memory tokens cancel (PAGE's
Memory interactions alone contribute zero), and so do GLOBAL_MEMORY's genesis and final tokens;
a touched page balances against a table that carries the same tokens with
the address as plain columns (the first access consuming the initial token,
the last producing the final one), and stops balancing when one final byte
or one timestamp differs. A proof of PAGE (small page) with that table
replays and verifies, with the verifier's own commitment of the fixed
columns; a verifier expecting other initial bytes refuses it."""
import json
import os

import numpy as np
import pytest

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from lambda_vm_amd.logup import BusInteraction, BusValue, LinearTerm, Multiplicity, Packing
from oracle.air_proof import precomputed_commitment, replay_air_tables, verify_air_tables

P = 0xFFFFFFFF00000001
KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kats.json")))["preprocessed_commitments"]


@pytest.mark.parametrize("blowup", [2, 4, 8])
def test_page_static_commitments(ora, blowup):
    g = KATS["coset_offset"]
    zero = V.page_precomputed(None, 2)
    assert bytes(ora.lde_commit(zero, blowup, g, want_lde=False)[1][0]) == bytes(KATS["page_zero"][str(blowup)])
    private = V.page_precomputed(b"\x07" * 100, 1)
    assert bytes(ora.lde_commit(private, blowup, g, want_lde=False)[1][0]) == bytes(KATS["page_private"][str(blowup)])
    if blowup == 2:
        data = V.page_precomputed(b"\x00\x01", 2)  # one nonzero initial byte
        assert bytes(ora.lde_commit(data, blowup, g, want_lde=False)[1][0]) != bytes(KATS["page_zero"]["2"])


def test_page_rows_follow_the_reference_layout():
    tr = V.generate_page_trace(3 << 18, b"\x05\x06", {(3 << 18) + 1: (0x1_0000_0007, 9)})
    c = V.PageCols
    assert tr.shape == (1 << 18, 5) and tr[-1, c.OFFSET] == (1 << 18) - 1
    assert list(tr[0]) == [0, 5, 5, 0, 0] and list(tr[1]) == [1, 6, 9, 7, 1] and list(tr[2]) == [2, 0, 0, 0, 0]
    g = V.generate_global_memory_trace(0, b"\x05", {3: (8, 2)}, page_size=16)
    assert list(g[0]) == [0, 5, 5, 0] and list(g[3]) == [3, 0, 8, 2]
    with pytest.raises(ValueError):
        V.generate_page_trace(5, None, {})
    with pytest.raises(ValueError):
        V.generate_page_trace(0, None, {1 << 18: (1, 1)})
    with pytest.raises(ValueError):
        V.generate_page_trace(0, None, {1: (1, 256)})


BASE = 0x7_0000_4000  # a page base with a nonzero high word
PAGE = 1024


def access_chain_table(page_base, init, final):
    """Each touched byte as one row [addr_lo, addr_hi, init, fini, ts_lo, ts_hi]
    that sends the initial token and receives the final one."""
    rows = []
    for addr, (ts, val) in sorted(final.items()):
        off = addr - page_base
        iv = init[off] if off < len(init) else 0
        rows.append([addr & 0xFFFFFFFF, addr >> 32, iv, val, ts & 0xFFFFFFFF, ts >> 32])
    while len(rows) < 4 or len(rows) & (len(rows) - 1):
        rows.append([0, 0, 0, 0, 0, 0])
    d = lambda col: BusValue.packed(Packing.DIRECT, col)  # noqa: E731
    k = lambda v: BusValue.of_linear([LinearTerm(v)])  # noqa: E731
    live = Multiplicity("Column", (6,))
    its = [BusInteraction.sender(V.BUS_MEMORY, live, [k(0), d(0), d(1), k(0), k(0), d(2)]),
           BusInteraction.receiver(V.BUS_MEMORY, live, [k(0), d(0), d(1), d(4), d(5), d(3)])]
    tr = np.zeros((len(rows), 7), dtype=np.uint64)
    for i, r in enumerate(rows):
        tr[i, :6] = r
        tr[i, 6] = int(any(r))
    a = A.logup_air(7, its)
    a.bus_descriptor = A.descriptor(a)
    return a, tr


def _contribution(tables, label):
    out = replay_air_tables(tables, 2, 3, 2, 4, 0, label)
    tot = [0, 0, 0]
    for o in out:
        tot = [(x + int(y)) % P for x, y in zip(tot, o["table_contribution"])]
    return tuple(tot)


def _with_desc(a):
    a.bus_descriptor = A.descriptor(a)
    return a


def _page_memory_only():
    """PAGE's two Memory interactions (its ARE_BYTES range check is balanced
    by the BITWISE table, not here)."""
    return _with_desc(A.logup_air(V.PageCols.NUM_COLUMNS, V.page_bus_interactions(BASE)[1:]))


def test_untouched_page_and_genesis_cancel():
    init = bytes(range(1, 200))
    page = _page_memory_only()
    assert _contribution([(page, V.generate_page_trace(BASE, init, {}, PAGE))], b"page") == (0, 0, 0)
    gm = _with_desc(V.global_memory_table_air(BASE, init, page_size=PAGE))
    assert _contribution([(gm, V.generate_global_memory_trace(BASE, init, {}, PAGE))], b"gm") == (0, 0, 0)
    touched = V.generate_global_memory_trace(BASE, init, {BASE + 3: (9, 4)}, PAGE)
    assert _contribution([(gm, touched)], b"gm") != (0, 0, 0)


def test_touched_page_balances_against_its_accesses():
    rng = np.random.default_rng(4)
    init = bytes(rng.integers(0, 256, 300, dtype=np.uint8))
    final = {BASE + int(o): (int(t), int(v)) for o, t, v in
             zip(rng.choice(PAGE, 40, replace=False), rng.integers(4, 2**40, 40), rng.integers(0, 256, 40))}
    page = _page_memory_only()
    chain = access_chain_table(BASE, init, final)
    assert _contribution([(page, V.generate_page_trace(BASE, init, final, PAGE)), chain], b"pg") == (0, 0, 0)
    for bad in ({**final, BASE + 7: (5, 1)},  # an extra touched byte the accesses do not know
                {a: (t + (a == min(final)), v) for a, (t, v) in final.items()}):  # one timestamp moved
        assert _contribution([(page, V.generate_page_trace(BASE, init, bad, PAGE)), chain], b"pg") != (0, 0, 0)


def test_page_proof_verifies_with_its_fixed_commitment():
    rng = np.random.default_rng(5)
    init = bytes(rng.integers(0, 256, 64, dtype=np.uint8))
    final = {BASE + int(o): (int(t), int(v)) for o, t, v in
             zip(rng.choice(PAGE, 12, replace=False), rng.integers(4, 2**33, 12), rng.integers(0, 256, 12))}
    page = _with_desc(V.page_table_air(BASE, init, page_size=PAGE))
    chain_air, chain_tr = access_chain_table(BASE, init, final)
    tables = [(page, V.generate_page_trace(BASE, init, final, PAGE)), (chain_air, chain_tr)]
    proofs = replay_air_tables(tables, 4, 3, 2, 6, 0, b"page-proof")
    assert proofs[0]["precomputed_root"] == precomputed_commitment(page, 4, 3)
    assert verify_air_tables([page, chain_air], proofs, 4, 3, 2, 6, 0, b"page-proof")[0]
    other = _with_desc(V.page_table_air(BASE, bytes([init[0] ^ 1]) + init[1:], page_size=PAGE))
    assert verify_air_tables([other, chain_air], proofs, 4, 3, 2, 6, 0, b"page-proof") == (False, "precomputed commitment")
