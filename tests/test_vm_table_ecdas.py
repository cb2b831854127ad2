"""The reference VM's ECDAS chip (prover/src/tables/ecdas.rs) with the
secp256k1 double-and-add witness of crypto/ecsm (curve.rs, witness.rs): 521
columns, 200 constraints (degree 3), 388 interactions.

Pins independent of the restatement: the reference's own known answers
(crypto/ecsm/src/tests/lib_tests.rs: the generator's even y, x(2G), x(3G),
(N - 1)·G sharing x with G); for random scalars the double-and-add result
equals k·G computed here by a right-to-left ladder of its own; the recovered
y satisfies the curve equation and is even, and a non-residue (Euler's
criterion, computed here) has none. The carries of the three convolution
relations sit inside their IS_HALFWORD windows. Constraints vanish and break
for a wrong λ byte, a wrong carry and op·next_op = 1. Bus checks: over one
scalar multiplication the Ecdas-bus messages telescope to the seed (G, the
top round, a double) and the result, and the Bit bus carries exactly the set
bits of k below its top bit; the ARE_BYTES and IS_HALFWORD sends equal, as a
multiset, the BITWISE receivers of the same lookups."""
from collections import Counter

import numpy as np
import pytest

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from oracle.air_proof import replay_air_tables, verify_air_tables
from test_vm_table_dvrm import bus_messages
from test_vm_table_eq import _constraint_values

P = 0xFFFFFFFF00000001
GX = 0x79BE667EF9DCBBAC55A06295CE870B07029BFCDB2DCE28D959F2815B16F81798
GY = 0x483ADA7726A3C4655DA4FBFC0E1108A8FD17B448A68554199C47D08FFB10D4B8
X2G = 0xC6047F9441ED7D6D3045406E95C07CD85C778E4B8CEF3CA7ABAC09B95C709EE5
X3G = 0xF9308A019258C31049344F85F89D5229B531C845836F99B08601F113BCE036F9
SP = V.SECP_P


def _add(p1, p2):
    """Affine addition on y^2 = x^3 + 7 (None is the point at infinity)."""
    if p1 is None:
        return p2
    if p2 is None:
        return p1
    (x1, y1), (x2, y2) = p1, p2
    if x1 == x2 and (y1 + y2) % SP == 0:
        return None
    if p1 == p2:
        lam = 3 * x1 * x1 * pow(2 * y1, SP - 2, SP) % SP
    else:
        lam = (y2 - y1) * pow(x2 - x1, SP - 2, SP) % SP
    x3 = (lam * lam - x1 - x2) % SP
    return x3, (lam * (x1 - x3) - y1) % SP


def ladder(k, pt):
    """k·pt, right to left."""
    acc = None
    while k:
        if k & 1:
            acc = _add(acc, pt)
        pt = _add(pt, pt)
        k >>= 1
    return acc


def test_secp256k1_known_answers():
    assert V.secp_recover_y(GX) == GY and GY % 2 == 0
    assert V.ecdas_steps(2, GX)[1][0] == X2G and V.ecdas_steps(3, GX)[1][0] == X3G
    assert V.ecdas_steps(V.SECP_N - 1, GX)[1][0] == GX
    assert V.ecdas_steps(1, GX) == ([], (GX, GY))
    for x in range(1, 40):
        rhs = (x ** 3 + 7) % SP
        residue = pow(rhs, (SP - 1) // 2, SP) == 1
        y = V.secp_recover_y(x)
        assert (y is not None) == residue, x
        if y is not None:
            assert y * y % SP == rhs and y % 2 == 0
    for bad in ((0, GX), (V.SECP_N, GX), (5, SP)):
        with pytest.raises(ValueError):
            V.ecdas_steps(*bad)


def test_double_and_add_matches_ladder():
    rng = np.random.default_rng(1)
    for _ in range(6):
        k = int.from_bytes(rng.bytes(32), "little") % V.SECP_N or 1
        steps, res = V.ecdas_steps(k, GX)
        assert res == ladder(k, (GX, GY))
        assert len(steps) == (k.bit_length() - 1) + bin(k).count("1") - 1
        for s in steps:
            _, cs = s.witness()
            for cc, off in zip(cs, V.ECDAS_CARRY_OFFSETS):
                assert all(0 <= v + off < 1 << 16 for v in cc[:63]) and cc[63] == 0


def ecdas_ops(ks, seed=0):
    ops = []
    for n, k in enumerate(ks):
        steps, _ = V.ecdas_steps(k, GX)
        ops += [(1000 + 4 * n, s) for s in steps]
    return ops


def test_ecdas_constraints(ora):
    air = V.ecdas_table_air()
    ops = ecdas_ops([0xDEADBEEF12345, 7])
    tr = V.generate_ecdas_trace(ops)
    assert (air.num_main, len(air.interactions)) == (521, 388)
    assert not np.any(_constraint_values(ora, air, tr) % np.uint64(P))
    c = V.EcdasCols
    for col, rows in ((c.LAMBDA + 3, slice(5, 70)), (c.C1 + 10, slice(70, 135))):
        bad = tr.copy()
        bad[2, col] = (int(bad[2, col]) + 1) % P
        assert np.any(_constraint_values(ora, air, bad)[rows, 2] % np.uint64(P))
    bad = tr.copy()
    bad[0, c.OP], bad[0, c.NEXT_OP] = 1, 1
    assert np.any(_constraint_values(ora, air, bad)[3, 0] % np.uint64(P))


def test_ecdas_bus_telescopes_and_bits_match_scalar():
    k = 0xB5E1_7A3F_0000_0001_C0FF_EE00_1234_5678_9ABC_DEF0
    ops = ecdas_ops([k])
    tr = V.generate_ecdas_trace(ops)
    its = V.ecdas_bus_interactions()
    got = bus_messages([it for it in its if it.bus_id == V.BUS_ECDAS], tr, range(tr.shape[0]))
    got = {key: v for key, v in got.items() if v}
    le = lambda v: tuple(V._le(v, 32))  # noqa: E731
    m = k.bit_length() - 1
    res = ladder(k, (GX, GY))
    seed = (V.BUS_ECDAS, 0, 1000, 0) + le(GX) + le(GY) + le(GX) + le(GY) + (m - 1, 0)
    final = (V.BUS_ECDAS, 0, 1000, 0) + le(res[0]) + le(res[1]) + le(GX) + le(GY) + (P - 1, 0)
    assert got == {seed: -1, final: 1}
    bits = bus_messages([it for it in its if it.bus_id == V.BUS_BIT], tr, range(tr.shape[0]))
    assert dict(bits) == {(V.BUS_BIT, 1000, 0, j): 1 for j in range(m) if k >> j & 1}


def test_ecdas_byte_and_carry_checks_match_bitwise():
    ops = ecdas_ops([12345678901234567])
    tr = V.generate_ecdas_trace(ops)
    sends = [it for it in V.ecdas_bus_interactions() if it.bus_id in (V.BUS_ARE_BYTES, V.BUS_IS_HALFWORD)]
    got = bus_messages(sends, tr, range(tr.shape[0]))
    lookups = []
    for _, s in ops:
        qs, cs = s.witness()
        for v in [s.round] + V._le(s.lam, 32) + qs[0] + V._le(s.xr, 32) + qs[1] + V._le(s.yr, 32) + qs[2]:
            lookups.append(("are_bytes", v, 0, 0))
        for cc, off in zip(cs, V.ECDAS_CARRY_OFFSETS):
            lookups += [("is_half", (v + off) & 0xFF, (v + off) >> 8, 0) for v in cc[:63]]
    bw = V.generate_bitwise_trace(lookups)
    used = sorted({V.bitwise_row_index(x, y, z) for _, x, y, z in lookups})
    want = bus_messages(V.bitwise_bus_interactions(), bw, used)
    got.subtract(Counter({key: -v for key, v in want.items()}))
    assert not +got and not -got


def test_ecdas_proof_verifies():
    air = V.ecdas_table_air()
    air.bus_descriptor = A.descriptor(air)
    tables = [(air, V.generate_ecdas_trace(ecdas_ops([0x1F2E3D4C5B6A7988])))]
    proofs = replay_air_tables(tables, 4, 3, 2, 4, 0, b"ecdas")
    assert verify_air_tables([air], proofs, 4, 3, 2, 4, 0, b"ecdas")[0]
