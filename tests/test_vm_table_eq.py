"""The reference VM's EQ table (prover/src/tables/eq.rs) captured as an AIR:
on a trace from its generator every transition constraint (the ADD carry
pair, the bit checks, res = eq XOR invert, and the LogUp constraints of its
six bus interactions) vanishes at every trace row, a trace with a wrong
result or a wrong difference half does not, and the oracle's replay of a full
proof of the table is accepted by the verifier restatement (and rejected for
the corrupted trace)."""
import numpy as np
import pytest

from lambda_vm_amd import vm_tables as V
from lambda_vm_amd import air as A
from oracle.air_proof import f3powers, replay_air_tables, verify_air_tables

P = 0xFFFFFFFF00000001


def _ops(k, seed):
    rng = np.random.default_rng(seed)
    ops = []
    for i in range(k):
        a = int(rng.integers(0, 2**63)) * 2 + int(rng.integers(0, 2))
        b = a if i % 3 == 0 else int(rng.integers(0, 2**63))
        ops.append(V.EqOperation(a, b, bool(i % 2)))
    ops.append(V.EqOperation(5, 7, False))  # a < b: the difference wraps
    ops.append(V.EqOperation(2**64 - 1, 0, True))
    return ops


def _constraint_values(ora, air, trace, seed=1):
    """All constraint values at every trace row (aux built with seeded challenges)."""
    rng = np.random.default_rng(seed)
    z = [int(x) for x in rng.integers(0, 2**63, 3)]
    al = [int(x) for x in rng.integers(0, 2**63, 3)]
    d = A.descriptor(air)
    ap = f3powers(tuple(al), d.alpha_powers_len)
    aux, L, _ = ora.logup_aux(d, trace, np.array(ap, dtype=np.uint64), np.array(z, dtype=np.uint64))
    n = trace.shape[0]
    toff = np.array([int(x) * pow(n, P - 2, P) % P for x in L], dtype=np.uint64)
    # the trace itself as a blowup-1 "LDE" in natural order: frame step 1
    return ora.eval_constraints(air.device, trace, aux.reshape(n, -1), 1, [z, al], ap, toff)


def test_constraints_vanish_on_generated_trace(ora):
    air = V.eq_table_air()
    tr = V.generate_eq_trace(_ops(11, 3))
    assert tr.shape == (16, 12)
    ev = _constraint_values(ora, air, tr)
    assert not np.any(ev % np.uint64(P)), np.argwhere(ev % np.uint64(P))[:5]


@pytest.mark.parametrize("col,row,delta", [(V.EqCols.RES, 1, 1), (V.EqCols.DIFF_2, 4, 1), (V.EqCols.EQ, 0, 1)])
def test_corrupted_trace_breaks_constraints(ora, col, row, delta):
    air = V.eq_table_air()
    tr = V.generate_eq_trace(_ops(11, 3))
    tr[row, col] = (int(tr[row, col]) + delta) % P
    ev = _constraint_values(ora, air, tr)
    assert np.any(ev[:5, row] % np.uint64(P))  # one of the table's own constraints fails at that row


def test_eq_table_proof_verifies_and_corruption_rejected():
    air = V.eq_table_air()
    air.bus_descriptor = A.descriptor(air)
    tr = V.generate_eq_trace(_ops(50, 9))
    proofs = replay_air_tables([(air, tr)], 4, 3, 2, 6, 0, b"eq-table")
    ok, why = verify_air_tables([air], proofs, 4, 3, 2, 6, 0, b"eq-table")
    assert ok, why
    bad = tr.copy()
    bad[7, V.EqCols.RES] ^= np.uint64(1)
    proofs = replay_air_tables([(air, bad)], 4, 3, 2, 6, 0, b"eq-table")
    assert not verify_air_tables([air], proofs, 4, 3, 2, 6, 0, b"eq-table")[0]
