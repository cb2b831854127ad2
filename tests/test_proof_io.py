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
