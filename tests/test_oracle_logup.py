"""Pins of the oracle's LogUp auxiliary-trace restatement (CPU only).

- The descriptor fingerprints equal a direct big-integer evaluation of the
  reference's packing formulas (crypto/stark/src/lookup.rs:250-345
  `accumulate_fingerprint_with`), over the interaction set of the reference's
  own parity test (crypto/stark/src/logup_gpu.rs:600-650): every packing plus
  a signed linear value. A wrong alpha index, shift or sign fails here.
- Term columns equal sum of +-m / fp with each fingerprint inverted on its own
  (no batch inversion), and the accumulated column follows
  acc[0] = 0, acc[i+1] = acc[i] + rowsum[i] - L/N (lookup.rs:1790-1834).
- The running sum closes (acc[n-1] + rowsum[n-1] - L/N = 0) and a balanced
  bus (one table sends a multiset, another receives it with multiplicities)
  has table contributions summing to zero.
"""
import random

import numpy as np
import pytest

from lambda_vm_amd import workload
from lambda_vm_amd.logup import (BusInteraction, BusValue, LinearTerm, Multiplicity, Packing, ONE,
                                 build_fingerprint_descriptor, split_interactions)
from pyref import P, f3add as _add, f3inv as _inv, f3mul as _mul, f3scale as _scale, f3sub as _sub


def T(x):
    return tuple(int(v) % P for v in x)


def f3add(a, b):
    return T(_add(a, b))


def f3sub(a, b):
    return T(_sub(a, b))


def f3mul(a, b):
    return T(_mul(a, b))


def f3scale(a, s):
    return T(_scale(a, s))


def f3inv(a):
    return T(_inv(a))


SHIFTS = {0: 1, 1: 1 << 8, 2: 1 << 16, 3: 1 << 24}


def packing_value(packing, c, col):
    """Bus elements of a packed value, written from the reference's combining
    formulas (lookup.rs:161-205 doc formulas)."""
    w2 = lambda x: (col(x) + col(x + 1) * (1 << 16)) % P  # noqa: E731
    w4 = lambda x: sum(col(x + i) * SHIFTS[i] for i in range(4)) % P  # noqa: E731
    v = packing.value
    if v == "Direct":
        return [col(c)]
    if v == "Word2L":
        return [w2(c)]
    if v == "Word4L":
        return [w4(c)]
    if v == "DWordWL":
        return [col(c), col(c + 1)]
    if v == "DWordHHW":
        return [col(c), w2(c + 1)]
    if v == "DWordWHH":
        return [w2(c), col(c + 2)]
    if v == "DWordHL":
        return [w2(c), w2(c + 2)]
    if v == "DWordBL":
        return [w4(c), w4(c + 4)]
    if v == "QuadHL":
        return [w2(c + 2 * i) for i in range(4)]
    if v == "QuadWL":
        return [col(c + i) for i in range(4)]
    raise ValueError(v)


def direct_fingerprint(it, row_vals, alpha, z):
    col = lambda c: int(row_vals[c]) % P  # noqa: E731
    elems = []
    for bv in it.values:
        if bv.packing is not None:
            elems += packing_value(bv.packing, bv.start_column, col)
        else:
            s = 0
            for t in bv.linear:
                s += t.coefficient if t.column is None else t.coefficient * col(t.column)
            elems.append(s % P)
    lc = (it.bus_id % P, 0, 0)
    apow = (1, 0, 0)
    for e in elems:
        apow = f3mul(apow, alpha)
        lc = f3add(lc, f3scale(apow, e))
    return f3sub(z, lc)


def direct_multiplicity(it, row_vals):
    m = it.multiplicity
    col = lambda c: int(row_vals[c]) % P  # noqa: E731
    if m.kind == "One":
        v = 1
    elif m.kind == "Column":
        v = col(m.cols[0])
    elif m.kind == "Sum":
        v = col(m.cols[0]) + col(m.cols[1])
    elif m.kind == "Negated":
        v = 1 - col(m.cols[0])
    elif m.kind == "Diff":
        v = col(m.cols[0]) - col(m.cols[1])
    elif m.kind == "Sum3":
        v = sum(col(c) for c in m.cols)
    else:
        v = sum(t.coefficient if t.column is None else t.coefficient * col(t.column) for t in m.linear)
    return v % P


def alpha_powers(alpha, count):
    out, cur = [], (1, 0, 0)
    for _ in range(count):
        out.append(cur)
        cur = f3mul(cur, alpha)
    return np.array(out, dtype=np.uint64)


def reference_interactions():
    """The interaction set of logup_gpu.rs's descriptor parity test."""
    one = ONE
    return [
        BusInteraction.sender(0, one, [BusValue.packed(Packing.DIRECT, 0)]),
        BusInteraction.sender(1, one, [BusValue.packed(Packing.WORD2L, 1)]),
        BusInteraction.sender(2, one, [BusValue.packed(Packing.WORD4L, 1)]),
        BusInteraction.sender(3, one, [BusValue.packed(Packing.DWORD_WL, 2)]),
        BusInteraction.sender(4, one, [BusValue.packed(Packing.DWORD_HHW, 0)]),
        BusInteraction.sender(5, one, [BusValue.packed(Packing.DWORD_WHH, 0)]),
        BusInteraction.sender(6, one, [BusValue.packed(Packing.DWORD_HL, 0)]),
        BusInteraction.sender(7, one, [BusValue.packed(Packing.QUAD_WL, 0)]),
        BusInteraction.sender(8, one, [BusValue.packed(Packing.QUAD_HL, 0)]),
        BusInteraction.sender(9, one, [BusValue.of_linear([LinearTerm(3, 1), LinearTerm(-2, 4), LinearTerm(42)]),
                                       BusValue.column(5)]),
    ]


def mixed_interactions():
    """Senders and receivers with every multiplicity kind."""
    return [
        BusInteraction.sender(11, Multiplicity("Column", (6,)), [BusValue.packed(Packing.WORD2L, 0)]),
        BusInteraction.receiver(12, ONE, [BusValue.packed(Packing.DWORD_BL, 0), BusValue.column(7)]),
        BusInteraction.sender(13, Multiplicity("Sum", (2, 3)), [BusValue.column(1)]),
        BusInteraction.receiver(14, Multiplicity("Negated", (4,)), [BusValue.packed(Packing.QUAD_WL, 3)]),
        BusInteraction.sender(15, Multiplicity("Diff", (5, 6)), [BusValue.packed(Packing.DWORD_HHW, 2)]),
        BusInteraction.receiver(16, Multiplicity("Sum3", (0, 1, 2)), [BusValue.column(0), BusValue.column(7)]),
        BusInteraction.sender(17, Multiplicity("Linear", linear=(LinearTerm(-5, 1), LinearTerm(9))),
                              [BusValue.of_linear([LinearTerm(7, 2), LinearTerm(-1)])]),
    ]


def rand_fp3(rng):
    return (rng.randrange(P), rng.randrange(P), rng.randrange(P))


def test_descriptor_fingerprints_match_packing_formulas(ora):
    rng = random.Random(5)
    its = reference_interactions()
    d = build_fingerprint_descriptor(its)
    n, m = 16, 8
    main = np.array([[rng.getrandbits(64) for _ in range(m)] for _ in range(n)], dtype=np.uint64)
    alpha, z = rand_fp3(rng), rand_fp3(rng)
    ap = alpha_powers(alpha, max(it.num_bus_elements() for it in its))
    assert d.alpha_powers_len == ap.shape[0]
    fps = ora.logup_fingerprints(d, main, ap, z)
    for k, it in enumerate(its):
        for r in range(n):
            got = tuple(int(x) % P for x in fps[k, r])
            assert got == direct_fingerprint(it, main[r], alpha, z), (k, r)


@pytest.mark.parametrize("count", [1, 2, 3, 4, 5, 7])
def test_aux_matches_direct_inversion(ora, count):
    rng = random.Random(10 + count)
    its = mixed_interactions()[:count]
    d = build_fingerprint_descriptor(its)
    pairs, absorbed = split_interactions(count)
    assert d.num_out_cols == pairs + 1
    n, m = 32, 8
    main = np.array([[rng.getrandbits(64) for _ in range(m)] for _ in range(n)], dtype=np.uint64)
    alpha, z = rand_fp3(rng), rand_fp3(rng)
    ap = alpha_powers(alpha, d.alpha_powers_len)
    aux, L, terms = ora.logup_aux(d, main, ap, z)
    # term column c at row r: sum over its interactions of m / fp, each inverted alone
    cols = [[2 * p, 2 * p + 1] for p in range(pairs)] + [list(range(count - absorbed, count))]
    want_terms = []
    for ks in cols:
        col = []
        for r in range(n):
            t = (0, 0, 0)
            for k in ks:
                inv = f3inv(direct_fingerprint(its[k], main[r], alpha, z))
                mult = direct_multiplicity(its[k], main[r])
                t = f3add(t, f3scale(inv, mult if its[k].is_sender else (P - mult) % P))
            col.append(t)
        want_terms.append(col)
    for c in range(len(cols)):
        for r in range(n):
            assert tuple(int(x) % P for x in terms[c, r]) == want_terms[c][r], (c, r)
    Lw = (0, 0, 0)
    for r in range(n):
        for c in range(len(cols)):
            Lw = f3add(Lw, want_terms[c][r])
    assert tuple(int(x) % P for x in L) == Lw
    off = f3scale(Lw, pow(n, P - 2, P))
    acc = (0, 0, 0)
    for r in range(n):
        for c in range(pairs):
            assert tuple(int(x) % P for x in aux[r, c]) == want_terms[c][r]
        assert tuple(int(x) % P for x in aux[r, pairs]) == acc, r
        rs = (0, 0, 0)
        for c in range(len(cols)):
            rs = f3add(rs, want_terms[c][r])
        acc = f3sub(f3add(acc, rs), off)
    assert acc == (0, 0, 0)  # the running sum closes


def test_balanced_bus_sums_to_zero(ora):
    # table A sends (x, y) pairs with multiplicity 1; table B receives each
    # distinct pair once with multiplicity = how often A sent it
    rng = random.Random(21)
    n = 64
    values = [(rng.randrange(50), rng.randrange(7)) for _ in range(n)]
    a = np.array([[x, y, 0] for x, y in values], dtype=np.uint64)
    counts = {}
    for v in values:
        counts[v] = counts.get(v, 0) + 1
    distinct = sorted(counts)
    b = np.zeros((n, 3), dtype=np.uint64)
    for i, v in enumerate(distinct):
        b[i] = (v[0], v[1], counts[v])
    for i in range(len(distinct), n):  # padding rows: multiplicity 0
        b[i] = (distinct[0][0], distinct[0][1], 0)
    send = [BusInteraction.sender(3, ONE, [BusValue.column(0), BusValue.column(1)])]
    recv = [BusInteraction.receiver(3, Multiplicity("Column", (2,)), [BusValue.column(0), BusValue.column(1)])]
    alpha, z = rand_fp3(rng), rand_fp3(rng)
    da, db = build_fingerprint_descriptor(send), build_fingerprint_descriptor(recv)
    _, La, _ = ora.logup_aux(da, a, alpha_powers(alpha, da.alpha_powers_len), z)
    _, Lb, _ = ora.logup_aux(db, b, alpha_powers(alpha, db.alpha_powers_len), z)
    total = f3add(tuple(int(x) % P for x in La), tuple(int(x) % P for x in Lb))
    assert total == (0, 0, 0)
    assert tuple(int(x) % P for x in La) != (0, 0, 0)
    # a single wrong multiplicity breaks the balance
    b[0, 2] += 1
    _, Lb2, _ = ora.logup_aux(db, b, alpha_powers(alpha, db.alpha_powers_len), z)
    assert f3add(tuple(int(x) % P for x in La), tuple(int(x) % P for x in Lb2)) != (0, 0, 0)


def test_descriptor_layout():
    its = mixed_interactions()
    d = build_fingerprint_descriptor(its)
    # 7 interactions: 3 committed pairs + 1 absorbed
    assert d.num_out_cols == 4
    assert d.out_col_offsets.tolist() == [0, 2, 4, 6, 7]
    assert d.out_col_interactions.tolist() == [0, 1, 2, 3, 4, 5, 6]
    # element alpha indices start at 1 for every interaction
    for k in range(d.num_interactions):
        e0 = d.elem_offsets[k]
        assert d.elem_alpha_idx[e0] == 1
    # receivers carry the negated multiplicity
    assert int(d.mult_const[1]) == P - 1
    assert workload.P == P
