"""The reference VM's KECCAK_RND round chip (prover/src/tables/keccak_rnd.rs):
24 rows per Keccak-f[1600], 1480 columns, 141 constraints, 1031 interactions.

Pins independent of the restatement: the permutation the 24 rows compute
equals the one inside Python's hashlib — SHAKE128 of a one-block message is
the first 21 lanes of Keccak-f of the padded block, SHA3-256 the first 4 —
and each row's input state is the previous row's output. Constraints vanish
on generated traces and break for a wrong rho or theta shift cell, a carry
that is not a bit and a multiplicity that is not a bit.
Cross-table: on the Keccak and KeccakRc buses the KECCAK core, KECCAK_RND and
KECCAK_RC tables balance in one transcript (the core's round-0 send and
round-24 receive against the round chip's chain, the round constants against
the fixed table) and stop balancing for a wrong output state; the round chip's
BYTE_ALU and ARE_BYTES sends equal, as a multiset of bus messages, what the
BITWISE table receives for the lookups of an independent round-by-round
Keccak (trace_builder.rs collect_bitwise_from_keccak)."""
import hashlib
from collections import Counter

import numpy as np

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from oracle.air_proof import replay_air_tables, verify_air_tables
from test_vm_table_dvrm import bus_messages
from test_vm_table_eq import _constraint_values

P = 0xFFFFFFFF00000001
M64 = 2**64 - 1


def _lanes(block: bytes) -> list:
    return [int.from_bytes(block[8 * i:8 * i + 8], "little") for i in range(25)]


def _padded(msg: bytes, rate: int, domain: int) -> bytes:
    """One absorbed block (pad10*1 with the domain bits) in a 200-byte state."""
    blk = bytearray(msg) + bytes([domain]) + bytes(rate - len(msg) - 1)
    blk[rate - 1] |= 0x80
    return bytes(blk) + bytes(200 - rate)


def _final_state(tr, k):
    """Output lanes of permutation k: the round-23 row's iota and chi bytes."""
    c = V.KeccakRndCols
    row = tr[24 * k + 23]
    out = []
    for y in range(5):
        for x in range(5):
            cols = [c.IOTA + b for b in range(8)] if (x, y) == (0, 0) else [c.lane(c.CHI, x, y, b) for b in range(8)]
            out.append(sum(int(row[col]) << (8 * b) for b, col in enumerate(cols)))
    return out


def keccak_f(lanes):
    """Keccak-f[1600] through the round chip's own trace generator."""
    tr = V.generate_keccak_rnd_trace([V.KeccakRoundOperation(0, tuple(lanes), (0,) * 25)])
    return _final_state(tr, 0)


def rnd_ops(k, seed):
    rng = np.random.default_rng(seed)
    ops = []
    for i in range(k):
        inp = [int(v) for v in rng.integers(0, 2**64, 25, dtype=np.uint64)]
        ops.append(V.KeccakRoundOperation(1000 + 4 * i + (i << 33), tuple(inp), tuple(keccak_f(inp))))
    return ops


def test_rounds_compute_keccak_f():
    for msg in (b"", b"abc", bytes(range(100))):
        out = keccak_f(_lanes(_padded(msg, 168, 0x1F)))
        want = hashlib.shake_128(msg).digest(168)
        assert b"".join(v.to_bytes(8, "little") for v in out[:21]) == want, msg
        out = keccak_f(_lanes(_padded(msg, 136, 0x06)))
        assert b"".join(v.to_bytes(8, "little") for v in out[:4]) == hashlib.sha3_256(msg).digest(), msg
    # the chain: each row's start is the previous row's output
    ops = rnd_ops(2, 1)
    tr = V.generate_keccak_rnd_trace(ops)
    c = V.KeccakRndCols
    assert tr.shape == (64, 1480) and list(tr[48:, c.MU]) == [0] * 16
    for r in range(1, 48):
        if r % 24 == 0:
            continue
        prev, row = tr[r - 1], tr[r]
        for x in range(5):
            for y in range(5):
                for b in range(8):
                    src = c.IOTA + b if (x, y) == (0, 0) else c.lane(c.CHI, x, y, b)
                    assert row[c.lane(c.START, x, y, b)] == prev[src]
    assert _final_state(tr, 1) == list(ops[1].output)


def test_keccak_rnd_constraints(ora):
    air = V.keccak_rnd_table_air()
    tr = V.generate_keccak_rnd_trace(rnd_ops(1, 2))
    assert (air.num_main, len(air.interactions)) == (1480, 1031)
    ev = _constraint_values(ora, air, tr) % np.uint64(P)
    assert not np.any(ev)
    c = V.KeccakRndCols
    for col, idx in ((c.lane(c.ROT_LEFT, 1, 2, 3), 40 + 4 * (1 * 5 + 2) + 1), (c.cxz_left(3, 4), 20 + 3 * 4 + 2)):
        bad = tr.copy()
        bad[5, col] += 1
        assert np.any(_constraint_values(ora, air, bad)[idx, 5] % np.uint64(P))
    bad = tr.copy()
    bad[7, c.cxz_right(2, 1)] = 2
    assert np.any(_constraint_values(ora, air, bad)[2 * 4 + 1, 7] % np.uint64(P))
    bad = tr.copy()
    bad[3, c.MU] = 2
    assert np.any(_constraint_values(ora, air, bad)[140, 3] % np.uint64(P))


def test_keccak_bus_balances_core_round_chip_and_constants():
    ops = rnd_ops(2, 3)
    core_ops = [V.KeccakOperation(op.timestamp, 0x10000 + 256 * i, op.input, op.output) for i, op in enumerate(ops)]
    core = A.logup_air(V.KeccakCols.NUM_COLUMNS, [it for it in V.keccak_bus_interactions()
                                                  if it.bus_id == V.BUS_KECCAK])
    rnd = A.logup_air(V.KeccakRndCols.NUM_COLUMNS, [it for it in V.keccak_rnd_bus_interactions()
                                                    if it.bus_id in (V.BUS_KECCAK, V.BUS_KECCAK_RC)])
    rc = A.logup_air(V.KeccakRcCols.NUM_COLUMNS, V.keccak_rc_bus_interactions())
    for a in (core, rnd, rc):
        a.bus_descriptor = A.descriptor(a)
    rc_tr = np.zeros((V.KeccakRcCols.NUM_ROWS, V.KeccakRcCols.NUM_COLUMNS), dtype=np.uint64)
    rc_tr[:, :V.KeccakRcCols.NUM_PRECOMPUTED] = V.keccak_rc_precomputed()
    rc_tr[:24, V.KeccakRcCols.MU] = len(ops)

    def total(core_ops):
        tables = [(core, V.generate_keccak_trace(core_ops)), (rnd, V.generate_keccak_rnd_trace(ops)), (rc, rc_tr)]
        tot = [0, 0, 0]
        for o in replay_air_tables(tables, 2, 3, 3, 2, 0, b"keccak-bus"):
            tot = [(a + int(b)) % P for a, b in zip(tot, o["table_contribution"])]
        return tuple(tot)

    assert total(core_ops) == (0, 0, 0)
    wrong = list(ops[1].output)
    wrong[7] ^= 1 << 40
    bad = [core_ops[0], V.KeccakOperation(core_ops[1].timestamp, core_ops[1].state_addr, ops[1].input, tuple(wrong))]
    assert total(bad) != (0, 0, 0)


def round_lookups(state, rnd):
    """BITWISE lookups of one round, computed lane-wise (trace_builder.rs
    collect_bitwise_from_keccak): the parity chain's XORs, the Cxz_left
    pairs, D's XORs, theta's XORs, the (rot_left, rot_right) pairs, chi's ANDs
    and XORs and iota's XORs; returns them and the next state."""
    byte = lambda v, b: (v >> (8 * b)) & 0xFF  # noqa: E731
    out = []
    cpar = [0] * 5
    for x in range(5):
        acc = state[x]
        for y in range(1, 5):
            out += [("xor", byte(acc, b), byte(state[x + 5 * y], b), 0) for b in range(8)]
            acc ^= state[x + 5 * y]
        cpar[x] = acc
    order = []  # reorder the chain lookups as the interactions list them: x, then stage, then byte
    for x in range(5):
        order += out[32 * x:32 * x + 32]
    out = order
    for x in range(5):
        shifted = sum(((cpar[x] >> (16 * h) << 1) & 0xFFFF) << (16 * h) for h in range(4))
        out += [("are_bytes", byte(shifted, 2 * i), byte(shifted, 2 * i + 1), 0) for i in range(4)]
    rot = [((v << 1) | (v >> 63)) & M64 for v in cpar]
    d = [cpar[(x + 4) % 5] ^ rot[(x + 1) % 5] for x in range(5)]
    for x in range(5):
        out += [("xor", byte(cpar[(x + 4) % 5], b), byte(rot[(x + 1) % 5], b), 0) for b in range(8)]
    theta = [state[i] ^ d[i % 5] for i in range(25)]
    for x in range(5):
        for y in range(5):
            out += [("xor", byte(state[x + 5 * y], b), byte(d[x], b), 0) for b in range(8)]
    for x in range(5):
        for y in range(5):
            rnc = V.KECCAK_RHO[x][y] % 16
            v = theta[x + 5 * y]
            left = sum(((v >> (16 * h) << rnc) & 0xFFFF) << (16 * h) for h in range(4))
            right = sum((((v >> (16 * h)) & 0xFFFF) >> (16 - rnc) if rnc else 0) << (16 * h) for h in range(4))
            out += [("are_bytes", byte(left, b), byte(right, b), 0) for b in range(8)]
    pi = [0] * 25
    for x in range(5):
        for y in range(5):
            r = V.KECCAK_RHO[x][y]
            pi[y + 5 * ((2 * x + 3 * y) % 5)] = ((theta[x + 5 * y] << r) | (theta[x + 5 * y] >> (64 - r))) & M64 if r \
                else theta[x + 5 * y]
    ands = {}
    for x in range(5):
        for y in range(5):
            a, b_ = pi[(x + 1) % 5 + 5 * y], pi[(x + 2) % 5 + 5 * y]
            ands[(x, y)] = ~a & b_ & M64
            out += [("and", 255 - byte(a, b), byte(b_, b), 0) for b in range(8)]
    nxt = [0] * 25
    for x in range(5):
        for y in range(5):
            out += [("xor", byte(pi[x + 5 * y], b), byte(ands[(x, y)], b), 0) for b in range(8)]
            nxt[x + 5 * y] = pi[x + 5 * y] ^ ands[(x, y)]
    rc = V.KECCAK_ROUND_CONSTANTS[rnd]
    out += [("xor", byte(nxt[0], b), byte(rc, b), 0) for b in range(8)]
    nxt[0] ^= rc
    return out, nxt


def test_keccak_rnd_bitwise_lookups_match_bitwise_receivers():
    ops = rnd_ops(1, 4)
    tr = V.generate_keccak_rnd_trace(ops)
    sends = [it for it in V.keccak_rnd_bus_interactions() if it.bus_id in (V.BUS_BYTE_ALU, V.BUS_ARE_BYTES)]
    assert len(sends) == 1028
    lookups = []
    for op in ops:
        state = list(op.input)
        for rnd in range(24):
            lk, state = round_lookups(state, rnd)
            lookups += lk
        assert state == list(op.output)
    lookups += [("are_bytes", 0, 0, 0)] * (220 * (tr.shape[0] - 24 * len(ops)))
    got = bus_messages(sends, tr, range(tr.shape[0]))
    bw = V.generate_bitwise_trace(lookups)
    used = sorted({V.bitwise_row_index(x, y, z) for _, x, y, z in lookups})
    want = bus_messages(V.bitwise_bus_interactions(), bw, used)
    got.subtract(Counter({k: -v for k, v in want.items()}))
    assert not +got and not -got, (len(+got), len(-got))


def test_keccak_rnd_proof_verifies():
    air = V.keccak_rnd_table_air()
    air.bus_descriptor = A.descriptor(air)
    tables = [(air, V.generate_keccak_rnd_trace(rnd_ops(1, 5)))]
    proofs = replay_air_tables(tables, 2, 3, 2, 4, 0, b"keccak-rnd")
    assert verify_air_tables([air], proofs, 2, 3, 2, 4, 0, b"keccak-rnd")[0]
