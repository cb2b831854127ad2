"""Preprocessed tables: the fixed columns of KECCAK_RC and BITWISE, as
restated in lambda_vm_amd.vm_tables, committed by the oracle (coset LDE of the
columns, row-pair bit-reversed Keccak leaves over the column subset, Merkle
tree) give the reference's static commitments (prover/src/tables/
keccak_rc.rs:97-116, bitwise.rs:199-218, extracted to tests/golden/kats.json)
for blowups 2, 4 and 8 at coset offset 3. The subset tree of the whole table's
LDE over the fixed columns gives the same root; any changed fixed value or a
shifted column range does not."""
import json
import os

import numpy as np
import pytest

from lambda_vm_amd import vm_tables as V

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kats.json")))["preprocessed_commitments"]
OFFSET = KATS["coset_offset"]


@pytest.mark.parametrize("blowup", [2, 4, 8])
def test_keccak_rc_static_commitment(ora, blowup):
    want = bytes(KATS["keccak_rc"][str(blowup)])
    _, nodes = ora.lde_commit(V.keccak_rc_precomputed(), blowup, OFFSET, want_lde=False)
    assert bytes(nodes[0]) == want
    # the fixed columns' subset tree over the LDE of the whole table (MU after them)
    full = V.generate_keccak_rc_trace(num_keccak_ops=7)
    lde, _ = ora.lde_commit(full, blowup, OFFSET)
    c = V.KeccakRcCols
    assert bytes(ora.commit_rows_bit_reversed(lde, 1, 0, c.NUM_PRECOMPUTED)[0]) == want
    assert bytes(ora.commit_rows_bit_reversed(lde, 1, 1, c.NUM_PRECOMPUTED + 1)[0]) != want
    bad = V.keccak_rc_precomputed()
    bad[5, 3] ^= 1
    assert bytes(ora.lde_commit(bad, blowup, OFFSET, want_lde=False)[1][0]) != want


@pytest.mark.parametrize("blowup", [2, 4, 8])
def test_bitwise_static_commitment(ora, blowup):
    want = bytes(KATS["bitwise"][str(blowup)])
    pre = V.bitwise_precomputed()
    _, nodes = ora.lde_commit(pre, blowup, OFFSET, want_lde=False)
    assert bytes(nodes[0]) == want
    if blowup == 2:
        pre[V.bitwise_row_index(3, 200, 9), V.BitwiseCols.SLLC] += 1
        assert bytes(ora.lde_commit(pre, blowup, OFFSET, want_lde=False)[1][0]) != want


def test_bitwise_rows_follow_the_formulas():
    pre = V.bitwise_precomputed()
    c = V.BitwiseCols
    rng = np.random.default_rng(3)
    for x, y, z in list(rng.integers(0, [256, 256, 16], size=(200, 3))) + [(0, 0, 0), (255, 255, 15), (0, 0, 1)]:
        x, y, z = int(x), int(y), int(z)
        r = pre[V.bitwise_row_index(x, y, z)]
        hw = x + 256 * y
        assert list(r[:3]) == [x, y, z]
        assert (r[c.AND], r[c.OR], r[c.XOR]) == (x & y, x | y, x ^ y)
        assert r[c.MSB8] == x >> 7 and r[c.MSB16] == hw >> 15
        assert r[c.ZERO] == int(x == y == z == 0)
        assert r[c.SLL] == (hw << z) % 65536 and r[c.SLLC] == (hw >> (16 - z) if z else 0)


def _keccak_rc_air():
    from lambda_vm_amd import air as A
    a = V.keccak_rc_table_air()
    a.bus_descriptor = A.descriptor(a)
    return a


def test_preprocessed_proof_replay_verifies_and_rejects_tampering(ora):
    """The oracle's replay of a KECCAK_RC proof (beside a plain table in one
    transcript) carries the static commitment as its precomputed root, and
    the verifier restatement accepts it; a changed fixed value in an opening,
    a precomputed root that is not the verifier's own commitment, or a fixed
    opening moved into the main rows is refused; the canonical bytes keep the
    precomputed root and openings."""
    import copy

    from lambda_vm_amd import air as A
    from lambda_vm_amd import proof_io
    from lambda_vm_amd.prover import AirTableProof
    from oracle.air_proof import replay_air_tables, verify_air_tables

    rc = _keccak_rc_air()
    fib = A.fibonacci_multi_column(4)
    tr_fib = A.fibonacci_trace(32, 4, np.arange(8, dtype=np.uint64).reshape(4, 2) + 1)
    A.with_boundaries_from_trace(fib, tr_fib)
    tables = [(rc, V.generate_keccak_rc_trace(num_keccak_ops=5)), (fib, tr_fib)]
    args = (4, OFFSET, 2, 8, 0, b"preprocessed")
    proofs = replay_air_tables(tables, *args)
    airs = [rc, fib]
    ok, why = verify_air_tables(airs, proofs, *args)
    assert ok, why
    assert proofs[0]["precomputed_root"] == bytes(KATS["keccak_rc"]["4"])
    assert proofs[1]["precomputed_root"] is None
    assert proofs[0]["main_rows"].shape[2] == 1 and proofs[0]["precomputed_rows"].shape[2] == 9
    bad = copy.deepcopy(proofs)
    bad[0]["precomputed_rows"][2, 1, 4] += 1
    assert verify_air_tables(airs, bad, *args) == (False, "table 0: precomputed opening")
    bad = copy.deepcopy(proofs)
    bad[0]["precomputed_root"] = bytes(32)
    assert verify_air_tables(airs, bad, *args) == (False, "precomputed commitment")
    bad = copy.deepcopy(proofs)
    bad[0]["main_rows"] = np.concatenate([bad[0]["precomputed_rows"], bad[0]["main_rows"]], axis=2)
    assert not verify_air_tables(airs, bad, *args)[0]
    data = proof_io.serialize_multi_proof([AirTableProof(**p) for p in proofs])
    back = proof_io.deserialize_multi_proof(data)
    assert back[0].precomputed_root == proofs[0]["precomputed_root"] and back[1].precomputed_root is None
    assert np.array_equal(back[0].precomputed_rows, proofs[0]["precomputed_rows"])
    assert verify_air_tables(airs, back, *args)[0]
    assert proof_io.serialize_multi_proof(back) == data
