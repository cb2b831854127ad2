"""The reference VM's DVRM table (prover/src/tables/dvrm.rs): 64-bit
DIV/DIVU/REM/REMU, 34 columns, 19 constraints, 35 interactions.

Pins independent of the restatement: the known answers of the reference's own
tests (prover/src/tests/dvrm_tests.rs: quotients and remainders, sign flags,
absolute values, n - r, division by zero, MIN / -1, the padding row, 35
interactions); for random and extreme operands, q·d + r = n over the integers
with |r| < |d| and r of n's sign (the definition of truncating division) or,
for d = 0, q = 2^64 - 1 and r = n. Constraints vanish on generated traces and
break for a forged division-by-zero remainder (the carry chain of
n - r + r = n, dvrm_tests.rs test_dvrm_rejects_false_div_by_zero_remainder),
a wrong sign_q and sign bits on unsigned rows.
Cross-table: DVRM's LT and MUL sends on the ALU bus balance in one transcript
against the LT and MUL tables' receivers fed the operations the trace builder
derives from DVRM (trace_builder.rs:3453-3466), and its BITWISE-bound sends
(ARE_BYTES, IS_HALFWORD, MSB16, ZERO — the NEG carries included) equal, as a
multiset of bus messages, what the BITWISE table receives for the lookups
trace_builder.rs collect_bitwise_from_dvrm lists."""
from collections import Counter

import numpy as np
import pytest

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from oracle.air_proof import replay_air_tables, verify_air_tables
from test_vm_table_eq import _constraint_values

P = 0xFFFFFFFF00000001
M64 = 2**64 - 1
MIN = 1 << 63
EXTREMES = [0, 1, 2, 3, M64, M64 - 1, MIN, MIN - 1, MIN + 1, 0xFFFF, 0x1_0000, 0xFFFF_FFFF, 1 << 32]


def dvrm_ops(k, seed):
    rng = np.random.default_rng(seed)
    ops = []
    for i in range(k):
        n = EXTREMES[i % len(EXTREMES)] if i % 3 == 0 else int(rng.integers(0, 2**64, dtype=np.uint64))
        if i % 4 == 0:
            d = EXTREMES[(i // 4) % len(EXTREMES)]
        elif i % 4 == 1:
            d = int(rng.integers(1, 1 << 20))  # small divisors: large quotients
        else:
            d = int(rng.integers(0, 2**64, dtype=np.uint64))
        ops.append((V.DvrmOperation(n, d, bool(i & 2)), bool(i % 3 == 1)))
    ops.append((V.DvrmOperation(MIN, M64, True), False))  # MIN / -1
    ops.append((V.DvrmOperation(MIN, M64, True), True))
    ops.append((V.DvrmOperation(5, 0, True), True))
    ops.append((V.DvrmOperation(M64 - 4, 0, True), False))
    return ops


def _s(v):
    return v - 2**64 if v >> 63 else v


# (n, d, signed) -> (q, r), from dvrm_tests.rs:21-100 and :374-392
KATS = [((20, 3, False), (6, 2)), ((100, 10, False), (10, 0)), ((3, 10, False), (0, 3)),
        ((M64, 2, False), (M64 // 2, 1)), ((20, 3, True), (6, 2)),
        (((-20) & M64, 3, True), ((-6) & M64, (-2) & M64)), ((20, (-3) & M64, True), ((-6) & M64, 2)),
        (((-20) & M64, (-3) & M64, True), (6, (-2) & M64)), ((42, 0, False), (M64, 42)),
        (((-5) & M64, 0, True), (M64, (-5) & M64)), ((0, 0, False), (M64, 0)), ((MIN, M64, True), (MIN, 0)),
        ((MIN, M64, False), (0, MIN)), ((0, 12345, False), (0, 0)), ((0, 12345, True), (0, 0)),
        ((12345, 1, False), (12345, 0))]


def test_dvrm_known_answers():
    for (n, d, s), (q, r) in KATS:
        op = V.DvrmOperation(n, d, s)
        assert (op.quotient(), op.remainder()) == (q, r), (n, d, s)
    # sign flags, absolute values and n - r (dvrm_tests.rs:105-163)
    flags = lambda op: (op.sign_n(), op.sign_d(), op.sign_r(), op.sign_q())  # noqa: E731
    assert flags(V.DvrmOperation(M64, M64, False)) == (False, False, False, False)
    assert flags(V.DvrmOperation(20, 3, True)) == (False, False, False, True)
    assert flags(V.DvrmOperation((-20) & M64, 3, True)) == (True, False, True, True)
    assert flags(V.DvrmOperation(MIN, M64, True))[::3] == (True, False)
    assert V.DvrmOperation(MIN, M64, True).sign_d()
    for n, d, s, ar, ad in ((20, 3, False, 2, 3), ((-20) & M64, 3, True, 2, 3), (20, (-3) & M64, True, 2, 3)):
        assert (V.DvrmOperation(n, d, s).abs_r(), V.DvrmOperation(n, d, s).abs_d()) == (ar, ad)
    assert V.DvrmOperation(20, 3, False).n_sub_r() == 18 and not V.DvrmOperation(20, 3, False).sign_n_sub_r()
    op = V.DvrmOperation((-20) & M64, 3, True)
    assert op.n_sub_r() == (-18) & M64 and op.sign_n_sub_r()
    assert len(V.dvrm_bus_interactions()) == 35
    # trace columns of division by zero and of overflow (dvrm_tests.rs:319-371)
    c = V.DvrmCols
    t = V.generate_dvrm_trace([(V.DvrmOperation(42, 0, False), False)])
    assert t.shape == (4, 34)
    assert (t[0, c.DIV_BY_ZERO], t[0, c.OVERFLOW]) == (1, 0)
    assert list(t[0, c.Q[0]:c.Q[3] + 1]) == [0xFFFF] * 4 and list(t[0, c.R[0]:c.R[3] + 1]) == [42, 0, 0, 0]
    t = V.generate_dvrm_trace([(V.DvrmOperation(MIN, M64, True), False)])
    assert (t[0, c.DIV_BY_ZERO], t[0, c.OVERFLOW], t[0, c.SIGN_Q]) == (0, 1, 0)
    assert list(t[0, c.Q[0]:c.Q[3] + 1]) == [0, 0, 0, 0x8000] and not t[0, c.R[0]:c.R[3] + 1].any()
    # the padding-only trace (dvrm_tests.rs:393-420)
    t = V.generate_dvrm_trace([])
    assert t.shape == (4, 34) and not t.any()
    # a count above MU_MAX splits over rows (dvrm_tests.rs:592-614)
    rows = V.dedup_dvrm_rows([(V.DvrmOperation(7, 2, False), False)] * 300 + [(V.DvrmOperation(7, 2, False), True)])
    assert [(mq, mr) for _, mq, mr in rows] == [(255, 1), (45, 0)]
    assert len(V.dedup_dvrm_rows([(V.DvrmOperation(7, 2, False), False), (V.DvrmOperation(7, 2, True), False)])) == 2


def test_dvrm_results():
    ops = dvrm_ops(400, 1)
    tr = V.generate_dvrm_trace(ops)
    c = V.DvrmCols
    rows = V.dedup_dvrm_rows(ops)
    word = lambda r, cols: sum(int(tr[r, col]) << (16 * i) for i, col in enumerate(cols))  # noqa: E731
    for r, (op, _, _) in enumerate(rows):
        q, rem = word(r, c.Q), word(r, c.R)
        if op.d == 0:
            assert (q, rem) == (M64, op.n)
            continue
        n, d = (_s(op.n), _s(op.d)) if op.signed else (op.n, op.d)
        qi, ri = (_s(q), _s(rem)) if op.signed else (q, rem)
        if op.signed and op.n == MIN and op.d == M64:
            assert (q, rem) == (MIN, 0)  # the quotient 2^63 wraps to MIN
            continue
        assert qi * d + ri == n, op
        assert abs(ri) < abs(d) and (ri == 0 or (ri < 0) == (n < 0)), op
        assert word(r, c.N_SUB_R) == (op.n - rem) & M64
    assert sum(mq for _, mq, _ in rows) == sum(1 for _, w in ops if not w)
    assert sum(mr for _, _, mr in rows) == sum(1 for _, w in ops if w)


def test_dvrm_constraints(ora):
    air = V.dvrm_table_air()
    ops = dvrm_ops(120, 2)
    tr = V.generate_dvrm_trace(ops)
    assert (air.num_main, len(air.interactions)) == (34, 35)
    assert not np.any(_constraint_values(ora, air, tr) % np.uint64(P))
    c = V.DvrmCols
    rows = V.dedup_dvrm_rows(ops)
    # a forged remainder on a division by zero breaks the carry chain n - r + r = n
    z = next(i for i, (op, _, _) in enumerate(rows) if op.d == 0 and not op.signed)
    bad = tr.copy()
    bad[z, c.R[0]] = 999 if bad[z, c.R[0]] != 999 else 998
    assert np.any(_constraint_values(ora, air, bad)[7:11, z] % np.uint64(P))
    # a quotient other than all ones on a division by zero
    bad = tr.copy()
    bad[z, c.Q[1]] = 0
    assert np.any(_constraint_values(ora, air, bad)[16, z] % np.uint64(P))
    s = next(i for i, (op, _, _) in enumerate(rows) if op.signed and not op.is_overflow())
    bad = tr.copy()
    bad[s, c.SIGN_Q] = 0
    assert np.any(_constraint_values(ora, air, bad)[6, s] % np.uint64(P))
    u = next(i for i, (op, _, _) in enumerate(rows) if not op.signed and op.n >> 63)
    bad = tr.copy()
    bad[u, c.SIGN_N] = 1
    assert np.any(_constraint_values(ora, air, bad)[12, u] % np.uint64(P))


def derived_lt_mul_ops(ops):
    """trace_builder.rs:3453-3466: per DVRM operation one LT[|r|, |d|] and the
    lo and hi MUL of d (signed as the division) by q (signed unless MIN / -1)."""
    lt = [V.LtOperation(op.abs_r(), op.abs_d(), False) for op, _ in ops]
    mul = []
    for op, _ in ops:
        m = V.MulOperation(op.d, op.signed, op.quotient(), op.sign_q())
        mul += [(m, False), (m, True)]
    return lt, mul


def test_dvrm_alu_lookups_balance_against_lt_and_mul():
    ops = dvrm_ops(60, 3)
    dv = A.logup_air(V.DvrmCols.NUM_COLUMNS, [it for it in V.dvrm_bus_interactions()
                                              if it.bus_id == V.BUS_ALU and it.is_sender])
    lt = A.logup_air(V.LtCols.NUM_COLUMNS, [it for it in V.lt_bus_interactions() if it.bus_id == V.BUS_ALU])
    mul = A.logup_air(V.MulCols.NUM_COLUMNS, [it for it in V.mul_bus_interactions() if it.bus_id == V.BUS_ALU])
    for a in (dv, lt, mul):
        a.bus_descriptor = A.descriptor(a)

    def total(lt_ops, mul_ops):
        tables = [(dv, V.generate_dvrm_trace(ops)), (lt, V.generate_lt_trace(lt_ops)),
                  (mul, V.generate_mul_trace(mul_ops))]
        tot = [0, 0, 0]
        for o in replay_air_tables(tables, 2, 3, 3, 2, 0, b"dvrm-alu"):
            tot = [(a + int(b)) % P for a, b in zip(tot, o["table_contribution"])]
        return tuple(tot)

    lt_ops, mul_ops = derived_lt_mul_ops(ops)
    assert total(lt_ops, mul_ops) == (0, 0, 0)
    # a MUL by a quotient one too large no longer balances
    i = next(k for k, (m, hi) in enumerate(mul_ops) if hi and m.rhs < M64)
    bad = list(mul_ops)
    m = bad[i][0]
    bad[i] = (V.MulOperation(m.lhs, m.lhs_signed, m.rhs + 1, m.rhs_signed), True)
    assert total(lt_ops, bad) != (0, 0, 0)


def _field(v):
    return v % P


def _eval_value(bv, row):
    if bv.packing is not None:
        return [_field(sum(cf * int(row[col]) for cf, col in el)) for el in bv.packing.elements(bv.start_column)]
    return [_field(sum(t.coefficient * (1 if t.column is None else int(row[t.column])) for t in bv.linear))]


def _eval_mult(m, row):
    if m.kind == "Linear":  # a field value read back as a small signed count
        v = sum(t.coefficient * (1 if t.column is None else int(row[t.column])) for t in m.linear) % P
        return v - P if v > P // 2 else v
    col = lambda i: int(row[m.cols[i]])  # noqa: E731
    return {"One": lambda: 1, "Column": lambda: col(0), "Sum": lambda: col(0) + col(1),
            "Sum3": lambda: col(0) + col(1) + col(2), "Negated": lambda: 1 - col(0),
            "Diff": lambda: col(0) - col(1)}[m.kind]()


def bus_messages(its, trace, rows):
    """Signed count of every bus message (bus id, elements) the interactions
    put on their buses over the given rows: sends count up, receives down."""
    out = Counter()
    for r in rows:
        row = trace[r]
        for it in its:
            mu = _eval_mult(it.multiplicity, row)
            if mu:
                msg = (it.bus_id,) + tuple(x for v in it.values for x in _eval_value(v, row))
                out[msg] += mu if it.is_sender else -mu
    return out


def bitwise_lookups(ops, rows, height):
    """trace_builder.rs:2116-2283 collect_bitwise_from_dvrm: per operation the
    IS_HALF of the n, d, r, n - r and q halves and the two ZERO checks; per
    trace row ARE_BYTES[mu_q, mu_r] (padding rows too), and on signed rows the
    three MSB16 and on sign_r / sign_d rows the NEG ZERO pairs."""
    out = []
    halves = lambda v: [(v >> s) & 0xFFFF for s in (0, 16, 32, 48)]  # noqa: E731
    for op, _ in ops:
        for v in (op.n, op.d, op.remainder(), op.n_sub_r(), op.quotient()):
            out += [("is_half", h & 0xFF, h >> 8, 0) for h in halves(v)]
        nh, dh = halves(op.n), halves(op.d)
        zero = sum(nh) + 262141 - 32769 * int(op.sign_n()) - sum(dh)
        for v in (zero, sum(dh)):
            out.append(("zero", v & 0xFF, (v >> 8) & 0xFF, v >> 16))
    out += [("are_bytes", 0, 0, 0)] * (height - len(rows))
    for op, mu_q, mu_r in rows:
        out.append(("are_bytes", mu_q, mu_r, 0))
        if op.signed:
            for v in (op.n, op.remainder(), op.d):
                top = v >> 48
                out.append(("msb16", top & 0xFF, top >> 8, 0))
        for flag, v in ((op.sign_r(), op.remainder()), (op.sign_d(), op.d)):
            if flag:
                h = halves(v)
                for s in (h[0] + h[1], sum(h)):
                    out.append(("zero", s & 0xFF, (s >> 8) & 0xFF, s >> 16))
    return out


def test_dvrm_bitwise_lookups_match_bitwise_receivers():
    ops = dvrm_ops(80, 4)
    tr = V.generate_dvrm_trace(ops)
    rows = V.dedup_dvrm_rows(ops)
    sends = [it for it in V.dvrm_bus_interactions() if it.bus_id != V.BUS_ALU]
    got = bus_messages(sends, tr, range(tr.shape[0]))
    lookups = bitwise_lookups(ops, rows, tr.shape[0])
    bw = V.generate_bitwise_trace(lookups)
    used = sorted({V.bitwise_row_index(x, y, z) for _, x, y, z in lookups})
    want = bus_messages(V.bitwise_bus_interactions(), bw, used)
    got.subtract(Counter({k: -v for k, v in want.items()}))
    assert not +got and not -got, (+got, -got)
    # every ZERO send is a 20-bit value with the right flag
    assert all(m[1] < 1 << 20 and m[2] == int(m[1] == 0) for m in want if m[0] == V.BUS_ZERO)
    # a remainder off by one moves its IS_HALF and NEG lookups
    bad = tr.copy()
    bad[0, V.DvrmCols.R[0]] += 1
    moved = bus_messages(sends, bad, range(tr.shape[0]))
    moved.subtract(Counter({k: -v for k, v in want.items()}))
    assert +moved or -moved


def test_dvrm_proof_verifies():
    air = V.dvrm_table_air()
    air.bus_descriptor = A.descriptor(air)
    tables = [(air, V.generate_dvrm_trace(dvrm_ops(50, 5)))]
    proofs = replay_air_tables(tables, 4, 3, 2, 6, 0, b"dvrm")
    assert verify_air_tables([air], proofs, 4, 3, 2, 6, 0, b"dvrm")[0]
