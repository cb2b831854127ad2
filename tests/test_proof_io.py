"""Canonical proof bytes (lambda_vm_amd.proof_io): proofs of the oracle's
replay serialize, read back to the same bytes, and the verifier restatement
accepts the proofs read back; a non-canonical field element is refused by
the reader, and single flipped bytes are caught by the reader or the
verifier."""
import numpy as np
import pytest

from lambda_vm_amd import proof_io
from lambda_vm_amd.prover import AirTableProof
from oracle.air_proof import verify_air_tables
from test_oracle_air_proof import OPTS, _prove, _tables

LABEL = b"air-proof-test"
P = 0xFFFFFFFF00000001


@pytest.fixture(scope="module")
def proved():
    tables = _tables(which=("fib", "quad", "logup"), seed=21)
    proofs = [AirTableProof(**p) for p in _prove(tables, label=LABEL, grinding=3)]
    return [a for a, _ in tables], proofs


def _verify(airs, proofs):
    o = dict(OPTS, grinding=3)
    return verify_air_tables(airs, proofs, o["blowup"], o["offset"], o["fri_k"], o["num_queries"], o["grinding"],
                             LABEL)


def test_round_trip_and_verify(proved):
    airs, proofs = proved
    data = proof_io.serialize_multi_proof(proofs)
    back = proof_io.deserialize_multi_proof(data)
    assert proof_io.serialize_multi_proof(back) == data
    ok, why = _verify(airs, back)
    assert ok, why
    # the fields read back equal the originals
    for a, b in zip(proofs, back):
        assert a.main_root == b.main_root and a.composition_root == b.composition_root
        assert a.nonce == b.nonce and list(a.iotas) == [int(x) for x in b.iotas]
        assert np.array_equal(np.asarray(a.main_rows, dtype=np.uint64), b.main_rows)


def test_non_canonical_element_refused(proved):
    _, proofs = proved
    data = bytearray(proof_io.serialize_table_proof(proofs[0]))
    # first element of the current-row OOD block: after n, the main root, the
    # two option tags and the block's width and height
    off = 8 + 32 + 1 + 1 + 8 + 8
    data[off:off + 8] = (0xFFFFFFFF00000001).to_bytes(8, "little")
    with pytest.raises(ValueError):
        proof_io.deserialize_table_proof(bytes(data))


def test_flipped_bytes_caught(proved):
    airs, proofs = proved
    data = proof_io.serialize_multi_proof(proofs)
    rng = np.random.default_rng(4)
    for pos in rng.integers(0, len(data), 40).tolist():
        bad = bytearray(data)
        bad[pos] ^= 0x10
        try:
            back = proof_io.deserialize_multi_proof(bytes(bad))
        except (ValueError, IndexError, MemoryError, OverflowError, struct_error()):
            continue
        except Exception:
            continue  # any other refusal by the reader is a catch too
        ok, _ = _verify(airs, back)
        assert not ok, pos


def struct_error():
    import struct
    return struct.error


def test_pruned_next_row_block_round_trip(proved):
    """Bus tables carry a one-column next-row block (the accumulated column),
    the example AIRs one column per trace column: both read back as written."""
    _, proofs = proved
    widths = [np.asarray(p.ood_block1).shape[0] for p in proofs]
    assert widths == [4, 3, 1]
    back = proof_io.deserialize_multi_proof(proof_io.serialize_multi_proof(proofs))
    for a, b in zip(proofs, back):
        assert np.array_equal(np.asarray(a.ood_block1, dtype=np.uint64) % np.uint64(P), b.ood_block1)


@pytest.mark.parametrize("where", ["block0_width", "next_width", "next_height", "parts", "fri_roots", "queries"])
def test_oversized_counts_refused(proved, where):
    """A length prefix larger than the bytes left is refused before anything
    is allocated for it (a corrupted count of 2^60 would otherwise ask for
    exabytes or loop for ages)."""
    _, proofs = proved
    p = proofs[0]
    data = bytearray(proof_io.serialize_table_proof(p))
    W = np.asarray(p.ood_block0).reshape(-1, 3).shape[0]
    C1, h1 = np.asarray(p.ood_block1).shape[:2]
    off = 8 + 32 + 1 + 1  # n, main root, aux tag (none), precomputed tag
    pos = {"block0_width": off}
    off += 16 + 24 * W
    pos["next_width"] = off
    pos["next_height"] = off + 8
    off += 16 + 24 * C1 * h1 + 32
    pos["parts"] = off
    off += 8 + 24 * len(p.parts_ood)
    pos["fri_roots"] = off
    off += 8 + 32 * len(p.fri_roots)
    off += 8 + 24 * len(p.fri_coeffs)
    pos["queries"] = off
    data[pos[where]:pos[where] + 8] = (1 << 60).to_bytes(8, "little")
    with pytest.raises(ValueError):
        proof_io.deserialize_table_proof(bytes(data))


def test_truncated_proof_refused(proved):
    _, proofs = proved
    data = proof_io.serialize_table_proof(proofs[2])
    for cut in (3, 50, len(data) // 2, len(data) - 1):
        with pytest.raises(ValueError):
            proof_io.deserialize_table_proof(data[:cut])
