"""AIR-table proofs on the CPU (oracle replay of every round) checked by the
verifier restatement: honest proofs of FibonacciMultiColumn, Quadratic and a
4-bus LogUp table (alone, together, with proof of work) are accepted; a proof
with any one value changed, or the proof of a trace that violates its AIR, is
rejected. This pins the verifier against the prover replay from both sides:
a verifier that skipped a check, or a replay that dropped a term, fails here."""
import copy

import numpy as np
import pytest

from lambda_vm_amd import air as A
from lambda_vm_amd import workload
from oracle.air_proof import replay_air_tables, verify_air_tables

P = workload.P
OPTS = dict(blowup=4, offset=3, fri_k=2, num_queries=6, grinding=0)


def _tables(n=64, which=("fib", "quad", "logup"), valid=True, seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for kind in which:
        if kind == "fib":
            a = A.fibonacci_multi_column(4)
            seeds = rng.integers(0, 2**62, (4, 2)).astype(np.uint64)
            tr = A.fibonacci_trace(n, 4, seeds) if valid else workload.trace_table(7, 1, n, 4)
            A.with_boundaries_from_trace(a, tr)
        elif kind == "quad":
            a = A.quadratic_multi(3)
            tr = A.quadratic_trace(n, 3, rng.integers(0, 2**62, 3).astype(np.uint64))
            A.with_boundaries_from_trace(a, tr)
        else:
            a = A.logup_air(24, A.four_bus_interactions())
            a.bus_descriptor = A.descriptor(a)
            tr = workload.trace_table(8, 2, n, 24)
        out.append((a, np.ascontiguousarray(tr, dtype=np.uint64)))
    return out


def _prove(tables, label=b"air-proof-test", **kw):
    o = dict(OPTS, **kw)
    return replay_air_tables(tables, o["blowup"], o["offset"], o["fri_k"], o["num_queries"], o["grinding"], label)


def _verify(airs, proofs, label=b"air-proof-test", **kw):
    o = dict(OPTS, **kw)
    return verify_air_tables(airs, proofs, o["blowup"], o["offset"], o["fri_k"], o["num_queries"], o["grinding"],
                             label)


@pytest.fixture(scope="module")
def three():
    tables = _tables()
    return tables, _prove(tables)


def test_three_tables_accepted(three):
    tables, proofs = three
    ok, why = _verify([a for a, _ in tables], proofs)
    assert ok, why
    # the LogUp table carries its aux commitment and contribution; the others do not
    assert proofs[2]["aux_root"] is not None and proofs[0]["aux_root"] is None
    assert len(proofs[0]["parts_ood"]) == 1 and len(proofs[1]["parts_ood"]) == 2


@pytest.mark.parametrize("kind", ["fib", "quad", "logup"])
def test_single_table_accepted(kind):
    tables = _tables(which=(kind,), seed=11)
    ok, why = _verify([a for a, _ in tables], _prove(tables))
    assert ok, why


def test_grinding_accepted_and_checked():
    tables = _tables(which=("quad",), seed=5)
    proofs = _prove(tables, grinding=6)
    assert proofs[0]["nonce"] is not None
    assert _verify([a for a, _ in tables], proofs, grinding=6)[0]
    bad = copy.deepcopy(proofs)
    bad[0]["nonce"] += 1
    assert not _verify([a for a, _ in tables], bad, grinding=6)[0]


def _tamper(proofs, t, field, index, delta=1):
    bad = copy.deepcopy(proofs)
    v = bad[t][field]
    if isinstance(v, np.ndarray):
        flat = v.reshape(-1)
        flat[index] = (int(flat[index]) + delta) % P
    elif isinstance(v, list) and isinstance(v[0], (tuple, list)):
        row = list(v[index])
        row[0] = (int(row[0]) + delta) % P
        v[index] = tuple(row)
    else:
        raise AssertionError(field)
    return bad


@pytest.mark.parametrize("t,field,index", [(0, "ood_block0", 1), (1, "parts_ood", 1), (2, "ood_block0", 26),
                                           (0, "fri_coeffs", 0), (1, "main_rows", 5), (2, "aux_rows", 3),
                                           (0, "parts_rows", 2)])
def test_tampered_value_rejected(three, t, field, index):
    tables, proofs = three
    ok, _ = _verify([a for a, _ in tables], _tamper(proofs, t, field, index))
    assert not ok


def test_tampered_fri_sym_and_contribution_rejected(three):
    tables, proofs = three
    bad = copy.deepcopy(proofs)
    bad[1]["fri_sym"][0][2] = ((bad[1]["fri_sym"][0][2][0] + 1) % P,) + tuple(bad[1]["fri_sym"][0][2][1:])
    assert not _verify([a for a, _ in tables], bad)[0]
    bad = copy.deepcopy(proofs)
    L = bad[2]["table_contribution"]
    bad[2]["table_contribution"] = ((L[0] + 1) % P, L[1], L[2])
    assert not _verify([a for a, _ in tables], bad)[0]


def test_trace_violating_its_air_rejected():
    tables = _tables(which=("fib",), valid=False, seed=2)
    ok, why = _verify([a for a, _ in tables], _prove(tables))
    assert not ok and why


def test_deep_weights_mismatch_rejected(monkeypatch):
    """A prover that weights the trace terms differently from the verifier
    (columns rotated) yields a DEEP codeword whose queried values disagree with
    the verifier's reconstruction: the FRI check rejects it."""
    import oracle.air_proof as AP
    tables = _tables(which=("quad",), seed=7)
    orig = AP.gamma_exponents

    def rotated(W, K, cols):
        e = orig(W, K, cols)
        return {((j + 1) % W, k): v for (j, k), v in e.items()}
    monkeypatch.setattr(AP, "gamma_exponents", rotated)
    proofs = _prove(tables)
    monkeypatch.setattr(AP, "gamma_exponents", orig)
    ok, why = _verify([a for a, _ in tables], proofs)
    assert not ok and "FRI" in why


class _Shape:
    def __init__(self, **kw):
        self.__dict__.update(kw)


def test_composition_part_rule_matches_reference():
    """Parts = composition_poly_degree_bound(n) / n. AirWithBuses
    (crypto/stark/src/lookup.rs:1063-1078): n * (max_degree - 1).max(1), so
    max degree 1..5 commits 1, 1, 2, 3, 4 parts; the example AIRs their own
    bounds (fibonacci_multi_column.rs:93: n -> 1; quadratic_air.rs:149: 2n -> 2)."""
    import oracle.air_proof as AP
    expected = {1: 1, 2: 1, 3: 2, 4: 3, 5: 4}
    for d, want in expected.items():
        shape = _Shape(bus=True, max_degree=d, example_parts=None)
        assert AP.num_parts(shape) == want
        air = A.logup_air(24, A.four_bus_interactions())
        air.max_degree = d
        assert air.num_parts == want
    assert A.fibonacci_multi_column(3).num_parts == 1 and A.quadratic_multi(3).num_parts == 2
    for a in (A.fibonacci_multi_column(3), A.quadratic_multi(3)):
        assert AP.num_parts(a) == a.num_parts
    # the LogUp degree alone: one absorbed interaction is degree 2 (one part),
    # committed pairs degree 3 (two parts); a degree-4 base constraint three
    assert A.logup_air(24, A.four_bus_interactions()[:1]).num_parts == 1
    assert A.logup_air(24, A.four_bus_interactions()).num_parts == 2
    assert A.product_bus_air(A.four_bus_interactions()).num_parts == 3


def test_pruned_openings_layout():
    """Surviving OOD openings (crypto/stark/src/ood.rs:44-48
    num_surviving_trace_openings): step*W + (K - step)*C with step 1. The bus
    table (24 main + 4 aux columns, K = 2) opens only the accumulated column,
    aux index 3, at z*g: C = 1, and its gamma power comes right after the
    current row's (W); the Fibonacci table (4 columns, K = 3) opens all four at
    both next rows: powers W + 2j + k - 1."""
    import oracle.air_proof as AP
    tables = _tables(which=("fib", "logup"), seed=13)
    proofs = _prove(tables)
    fib, bus = tables[0][0], tables[1][0]
    assert AP.next_row_cols(bus) == [27] and AP.frame_rows(bus) == 2
    assert np.asarray(proofs[1]["ood_block1"]).shape == (1, 1, 3)
    assert np.asarray(proofs[0]["ood_block1"]).shape == (4, 2, 3)
    e = AP.gamma_exponents(28, 2, [27])
    assert len(e) == 28 + 1 and e[(27, 1)] == 28 and all(e[(j, 0)] == j for j in range(28))
    e = AP.gamma_exponents(4, 3, range(4))
    assert len(e) == 4 + 2 * 4 and all(e[(j, k)] == 4 + 2 * j + k - 1 for j in range(4) for k in (1, 2))
    assert fib.next_row_cols() == AP.next_row_cols(fib) and bus.next_row_cols() == AP.next_row_cols(bus)


def test_unpruned_bus_proof_rejected(monkeypatch):
    """A bus-table proof that opens every column at z*g (the trait default)
    has the wrong next-row block for the AIR: the verifier refuses it."""
    import oracle.air_proof as AP
    tables = _tables(which=("logup",), seed=17)
    monkeypatch.setattr(AP, "next_row_cols", lambda air: list(range(air.num_main + air.num_aux)))
    proofs = _prove(tables)
    monkeypatch.undo()
    ok, why = _verify([a for a, _ in tables], proofs)
    assert not ok and "OOD shape" in why


def _product_tables(n, interactions, seed):
    a = A.product_bus_air(interactions)
    a.bus_descriptor = A.descriptor(a) if interactions else None
    return [(a, np.ascontiguousarray(A.product_trace(n, seed), dtype=np.uint64))]


@pytest.mark.parametrize("which", ["degree2_bus", "degree4_bus", "degree4_no_bus"])
def test_part_counts_prove_and_verify(which):
    if which == "degree2_bus":
        a = A.logup_air(24, A.four_bus_interactions()[:1])
        a.bus_descriptor = A.descriptor(a)
        tables = [(a, np.ascontiguousarray(workload.trace_table(8, 5, 64, 24), dtype=np.uint64))]
        want = 1
    else:
        tables = _product_tables(64, A.four_bus_interactions() if which == "degree4_bus" else [], seed=4)
        want = 3
    proofs = _prove(tables)
    assert len(proofs[0]["parts_ood"]) == want
    ok, why = _verify([a for a, _ in tables], proofs)
    assert ok, why


def test_degree4_violating_trace_rejected():
    tables = _product_tables(64, A.four_bus_interactions(), seed=6)
    tables[0][1][5, 24] ^= 1
    ok, why = _verify([a for a, _ in tables], _prove(tables))
    assert not ok and why


def _horner(coeffs, x):
    acc = (0, 0, 0)
    for c in reversed(coeffs):
        acc = tuple((a * x + int(v)) % P for a, v in zip(acc, c))
    return acc


def _poly_horner_fp3(coeffs, x3):
    """Fp3 Horner with an Fp3 point."""
    from oracle.air_proof import f3add, f3mul
    acc = (0, 0, 0)
    for c in reversed(coeffs):
        acc = f3add(f3mul(acc, x3), tuple(int(v) % P for v in c))
    return acc


@pytest.mark.parametrize("num_parts", [2, 3, 4, 5])
def test_break_in_parts_recombines(ora, num_parts):
    """The parts of break_in_parts recombine: H(x) = sum_i x^i H_i(x^P) at every
    coset point x, with each H_i read back from its LDE by an inverse FFT and
    evaluated by Horner; each part has at most ceil(N/P) coefficients. For
    P = 2 the parts equal the degree-2 split's (H0, H1), built independently."""
    from oracle.air_proof import break_in_parts_lde
    N, g = 32, 3
    H = workload.trace_table(12, num_parts, N, 3)
    parts = break_in_parts_lde(H, num_parts, g)
    wN = ora.root_of_unity(5)
    ginv = pow(g, P - 2, P)
    ninv = pow(N, P - 2, P)
    coeffs = []
    for i in range(num_parts):
        a = ora.fft_two_half(parts[:, 3 * i:3 * i + 3], inverse=True)
        c = [tuple(int(v) * ninv * pow(ginv, k, P) % P for v in a[k]) for k in range(N)]
        L = -(-N // num_parts)
        assert all(x == (0, 0, 0) for x in c[L:])
        coeffs.append(c[:L])
    for j in range(N):
        x = g * pow(wN, j, P) % P
        xp = pow(x, num_parts, P)
        total = (0, 0, 0)
        for i in range(num_parts):
            v = _horner(coeffs[i], xp)
            total = tuple((t + pow(x, i, P) * w) % P for t, w in zip(total, v))
        assert total == tuple(int(v) % P for v in H[j])
    if num_parts == 2:
        halves = ora.decompose_d2(H, g)
        d2_lde, _ = ora.lde_commit(halves, 2, ginv)
        assert np.array_equal(d2_lde % np.uint64(P), parts % np.uint64(P))
