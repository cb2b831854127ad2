"""The reference VM's ECSM chip (prover/src/tables/ecsm.rs): one row per
secp256k1 scalar-multiplication ecall, 667 columns, 413 constraints, 579
interactions.

Pins independent of the restatement: the inputs the reference rejects (k = 0,
k >= N, xG >= p, an x with no curve point) raise; the written xR is x(k·G) of
a right-to-left ladder written in the ECDAS test; the two carry arrays sit in
their IS_HALFWORD windows. Constraints vanish on generated traces and break
for a non-bit scalar bit, a wrong overflow addend (the xG < p, k < N, xR < p
checks) and a wrong carry. Cross-table: ECSM and the ECDAS chain of the same
ecalls balance in one transcript on the Ecdas bus (ECSM seeds G at the round
below the top bit and drains the result at round -1) and on the Bit bus (ECSM
receives every set bit of k and sends the top one, ECDAS sends the others),
and stop balancing when one result byte differs; ECSM's pointer reads of
x10-x12 balance on the MEMW bus against MEMW_REGISTER; its byte, halfword and
ZERO checks equal, as a multiset, the BITWISE receivers of the same lookups."""
from collections import Counter

import numpy as np
import pytest

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from oracle.air_proof import replay_air_tables, verify_air_tables
from test_vm_table_dvrm import bus_messages
from test_vm_table_ecdas import GX, GY, ladder
from test_vm_table_eq import _constraint_values

P = 0xFFFFFFFF00000001
KS = [0x1D, V.SECP_N - 2, 0xC0FFEE_0000_1111_2222_3333_4444_5555_6666_7777_8888_9999_AAAA_BBBB]
X7 = ladder(7, (GX, GY))[0]   # another valid x-coordinate


def ecsm_ops(ks, xg=GX):
    return [V.EcsmOperation(4000 + 16 * i, 0x40000 + 0x100 * i, 0x40040 + 0x100 * i, 0x40080 + 0x100 * i, k, xg)
            for i, k in enumerate(ks)]


def ecdas_rows(ops):
    return [(op.timestamp, s) for op in ops for s in op.witness()["steps"]]


def test_ecsm_witness():
    for k in KS:
        for xg in (GX, X7):
            w = V.EcsmOperation(4, 0, 0, 0, k, xg).witness()
            yg = V.secp_recover_y(xg)
            assert (w["xr"], w["yr"]) == ladder(k, (xg, yg))
            assert w["len_k"] == k.bit_length() - 1
            for cc, off in zip((w["c0"], w["c1"]), V.ECSM_CARRY_OFFSETS):
                assert all(0 <= v + off < 1 << 16 for v in cc[:63]) and cc[63] == 0
    for k, xg in ((0, GX), (V.SECP_N, GX), (3, V.SECP_P), (3, 5)):
        with pytest.raises(ValueError):
            V.EcsmOperation(4, 0, 0, 0, k, xg).witness()


def test_ecsm_constraints(ora):
    air = V.ecsm_table_air()
    tr = V.generate_ecsm_trace(ecsm_ops(KS))
    assert (air.num_main, len(air.interactions)) == (667, 579)
    assert not np.any(_constraint_values(ora, air, tr) % np.uint64(P))
    c = V.EcsmCols
    for col, val, rows in ((c.K + 3, 2, slice(4, 5)), (c.K_SUB_N + 5, None, slice(389, 413)),
                           (c.C1 + 7, None, slice(323, 388))):
        bad = tr.copy()
        bad[1, col] = val if val is not None else (int(bad[1, col]) + 1) % P
        assert np.any(_constraint_values(ora, air, bad)[rows, 1] % np.uint64(P)), col


def _total(tables, label):
    tot = [0, 0, 0]
    for o in replay_air_tables(tables, 2, 3, 3, 2, 0, label):
        tot = [(x + int(y)) % P for x, y in zip(tot, o["table_contribution"])]
    return tuple(tot)


def test_ecsm_and_ecdas_close_the_ecdas_and_bit_buses():
    ops = ecsm_ops(KS)
    buses = (V.BUS_ECDAS, V.BUS_BIT)
    ecsm = A.logup_air(V.EcsmCols.NUM_COLUMNS, [it for it in V.ecsm_bus_interactions() if it.bus_id in buses])
    ecdas = A.logup_air(V.EcdasCols.NUM_COLUMNS, [it for it in V.ecdas_bus_interactions() if it.bus_id in buses])
    for a in (ecsm, ecdas):
        a.bus_descriptor = A.descriptor(a)
    das = V.generate_ecdas_trace(ecdas_rows(ops))
    assert _total([(ecsm, V.generate_ecsm_trace(ops)), (ecdas, das)], b"ecsm-ecdas") == (0, 0, 0)
    bad = V.generate_ecsm_trace(ops)
    bad[2, V.EcsmCols.XR + 5] ^= 1
    assert _total([(ecsm, bad), (ecdas, das)], b"ecsm-ecdas") != (0, 0, 0)


def test_ecsm_pointer_reads_balance_against_memw_register():
    ops = ecsm_ops(KS[:2])
    acc = []
    for op in ops:
        for off, reg, val in ((0, 11, op.addr_xg), (1, 12, op.addr_k), (2, 10, op.addr_xr)):
            ts = op.timestamp + off
            acc.append(V.RegisterAccess(2 * reg, ts, val, val, ts - 1, True))
    regs = [it for it in V.ecsm_bus_interactions() if it.bus_id == V.BUS_MEMW and it.values[8].linear
            and it.values[8].linear[0].coefficient == 1 and it.values[8].linear[0].column is None]
    assert len(regs) == 3
    ecsm = A.logup_air(V.EcsmCols.NUM_COLUMNS, regs)
    mem = A.logup_air(V.MemwRegisterCols.NUM_COLUMNS, [it for it in V.memw_register_bus_interactions()
                                                       if it.bus_id == V.BUS_MEMW])
    for a in (ecsm, mem):
        a.bus_descriptor = A.descriptor(a)
    assert _total([(ecsm, V.generate_ecsm_trace(ops)), (mem, V.generate_memw_register_trace(acc))],
                  b"ecsm-regs") == (0, 0, 0)


def test_ecsm_checks_match_bitwise():
    ops = ecsm_ops(KS)
    tr = V.generate_ecsm_trace(ops)
    sends = [it for it in V.ecsm_bus_interactions() if it.bus_id in (V.BUS_ARE_BYTES, V.BUS_IS_HALFWORD, V.BUS_ZERO)]
    got = bus_messages(sends, tr, range(tr.shape[0]))
    lookups = []
    for op in ops:
        w = op.witness()
        for v in V._le(w["x2"], 32) + V._le(w["q0"], 32) + V._le(w["yg"], 32) + V._le(w["q1"], 33):
            lookups.append(("are_bytes", v, 0, 0))
        for cc, off in zip((w["c0"], w["c1"]), V.ECSM_CARRY_OFFSETS):
            lookups += [("is_half", (v + off) & 0xFF, (v + off) >> 8, 0) for v in cc[:63]]
        for v in w["sub"]:
            lookups += [("is_half", (v >> (16 * j)) & 0xFF, (v >> (16 * j + 8)) & 0xFF, 0) for j in range(16)]
        ksum = sum(V._le(op.k, 32))
        lookups.append(("zero", ksum & 0xFF, (ksum >> 8) & 0xFF, ksum >> 16))
    bw = V.generate_bitwise_trace(lookups)
    used = sorted({V.bitwise_row_index(x, y, z) for _, x, y, z in lookups})
    want = bus_messages(V.bitwise_bus_interactions(), bw, used)
    got.subtract(Counter({key: -v for key, v in want.items()}))
    assert not +got and not -got


def test_ecsm_proof_verifies():
    air = V.ecsm_table_air()
    air.bus_descriptor = A.descriptor(air)
    tables = [(air, V.generate_ecsm_trace(ecsm_ops(KS)))]
    proofs = replay_air_tables(tables, 4, 3, 2, 4, 0, b"ecsm")
    assert verify_air_tables([air], proofs, 4, 3, 2, 4, 0, b"ecsm")[0]
