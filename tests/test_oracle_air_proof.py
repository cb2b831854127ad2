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
    orig = AP.gamma_index
    monkeypatch.setattr(AP, "gamma_index", lambda W, K, j, k: orig(W, K, (j + 1) % W, k))
    proofs = _prove(tables)
    monkeypatch.setattr(AP, "gamma_index", orig)
    ok, why = _verify([a for a, _ in tables], proofs)
    assert not ok and "FRI" in why
