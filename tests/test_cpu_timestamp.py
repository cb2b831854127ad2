"""The CPU table's fixed TIMESTAMP column (prover/src/tables/cpu.rs:161-165:
row i holds 4·(i + 1)), committed by the oracle like every fixed column,
gives the reference's static commitments for every height 2^2..2^19 and
blowups 2, 4 and 8 at coset offset 3 (cpu_timestamp_commitments.rs, extracted
to tests/golden/kats.json): 54 known answers for the LDE + bit-reversed
row-pair Keccak tree of one column, from four rows up. The height helpers
follow cpu.rs:72-121."""
import json
import os

import pytest

from lambda_vm_amd import vm_tables as V

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kats.json")))["preprocessed_commitments"]


@pytest.mark.parametrize("blowup", [2, 4, 8])
def test_cpu_timestamp_static_commitments(ora, blowup):
    want = KATS["cpu_timestamp"][str(blowup)]
    assert sorted(map(int, want)) == list(range(2, 20))
    for log_h, root in want.items():
        col = V.cpu_timestamp_precomputed(1 << int(log_h))
        got = bytes(ora.lde_commit(col, blowup, KATS["coset_offset"], want_lde=False)[1][0])
        assert got == bytes(root), (blowup, log_h)
    col = V.cpu_timestamp_precomputed(1 << 5)
    col[3, 0] += 1
    assert bytes(ora.lde_commit(col, blowup, 3, want_lde=False)[1][0]) != bytes(want["5"])


def test_cpu_height_rules():
    assert V.cpu_final_pc_timestamp(8) == 33
    assert V.cpu_timestamp_bases([8, 4, 16]) == [0, 32, 48]
    assert V.cpu_table_heights(10, 4) == [4, 4, 4] and V.cpu_table_heights(9, 8) == [8, 4]
    assert V.cpu_table_heights(3, 1 << 20) == [4] and V.cpu_table_heights(0, 8) == []
    V.cpu_validate_heights([8, 4], 8)
    for bad in ([6], [2], [16]):
        with pytest.raises(ValueError):
            V.cpu_validate_heights(bad, 8)
