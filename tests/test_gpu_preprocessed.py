"""Preprocessed tables on the GPU (lvm_coset_lde_split_trees_dev): one coset
LDE of the whole table and two subset trees. The fixed-columns root equals
the reference's static commitment (prover/src/tables/keccak_rc.rs:97-116,
bitwise.rs:199-218, the CPU table's TIMESTAMP column at every height 2^2..2^19
from cpu_timestamp_commitments.rs; tests/golden/kats.json) and the oracle's subset tree, node
for node; the multiplicity-columns root equals the oracle's subset tree over
the remaining columns; opened rows are the whole LDE rows."""
import json
import os

import numpy as np
import pytest

from lambda_vm_amd import vm_tables as V
from lambda_vm_amd import workload

pytestmark = pytest.mark.gpu
KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kats.json")))["preprocessed_commitments"]
OFFSET = KATS["coset_offset"]
P = workload.P


def canon(a):
    a = np.asarray(a, dtype=np.uint64)
    return np.where(a >= np.uint64(P), a - np.uint64(P), a)


def _check(gpu, ora, trace, split, blowup, want):
    n, m = trace.shape
    print(f"[preprocessed] {n} x {m} split {split} blowup {blowup}", flush=True)
    d = gpu.DeviceBuffer.from_array(trace)
    w = gpu.coset_weights_dev(n, OFFSET)
    h, pre_root, pre_nodes = gpu.coset_lde_split_trees_dev(d, n, m, blowup, w, split)
    d.free()
    w.free()
    assert pre_root == want
    lde, _ = ora.lde_commit(trace, blowup, OFFSET)
    ref_pre = ora.commit_rows_bit_reversed(lde, 1, 0, split)
    assert np.array_equal(pre_nodes, ref_pre)
    ref_mult = ora.commit_rows_bit_reversed(lde, 1, split, m)
    assert h.root == bytes(ref_mult[0])
    # openings: whole rows, multiplicity-tree paths; fixed-tree paths from the host nodes
    N = n * blowup
    pos = [0, 1, N // 2 - 1, (N // 2) // 3]
    paths, rows = h.open(pos)
    bits = N.bit_length() - 1
    for q, p in enumerate(pos):
        r0 = int(format(2 * p, f"0{bits}b")[::-1], 2)
        assert np.array_equal(canon(rows[q, 0]), canon(lde[r0]))
        assert ora.merkle_verify(h.root, bytes(ref_mult[N // 2 - 1 + p]), p, paths[q])
    pre_paths = gpu.merkle_paths_from_nodes(pre_nodes, pos)
    for q, p in enumerate(pos):
        assert ora.merkle_verify(pre_root, bytes(ref_pre[N // 2 - 1 + p]), p, pre_paths[q])
    h.close()


@pytest.mark.parametrize("blowup", [2, 4, 8])
def test_keccak_rc_split_commitment(gpu, ora, blowup):
    trace = V.generate_keccak_rc_trace(num_keccak_ops=11)
    _check(gpu, ora, trace, V.KeccakRcCols.NUM_PRECOMPUTED, blowup, bytes(KATS["keccak_rc"][str(blowup)]))


@pytest.mark.parametrize("blowup", [2, 4])
def test_bitwise_split_commitment(gpu, ora, blowup):
    rng = np.random.default_rng(blowup)
    kinds = list(V.BITWISE_MU_COLUMN)
    lookups = [(kinds[int(k)], int(x), int(y), int(z)) for k, x, y, z in
               zip(rng.integers(0, len(kinds), 3000), rng.integers(0, 256, 3000), rng.integers(0, 256, 3000),
                   rng.integers(0, 16, 3000))]
    trace = V.generate_bitwise_trace(lookups)
    _check(gpu, ora, trace, V.BitwiseCols.NUM_PRECOMPUTED, blowup, bytes(KATS["bitwise"][str(blowup)]))


@pytest.mark.parametrize("blowup", [2, 4, 8])
def test_cpu_timestamp_split_commitments(gpu, blowup):
    """The CPU table's fixed TIMESTAMP column beside a multiplicity-like column,
    every height 2^2..2^19: the fixed-columns root is the reference's static
    commitment (cpu_timestamp_commitments.rs)."""
    want = KATS["cpu_timestamp"][str(blowup)]
    rng = np.random.default_rng(blowup)
    print(f"[preprocessed] CPU TIMESTAMP 2^2..2^19, blowup {blowup}", flush=True)
    for log_h in range(2, 20):
        n = 1 << log_h
        trace = np.concatenate([V.cpu_timestamp_precomputed(n), rng.integers(0, 2**63, (n, 1), dtype=np.uint64)], 1)
        d = gpu.DeviceBuffer.from_array(np.ascontiguousarray(trace))
        w = gpu.coset_weights_dev(n, OFFSET)
        h, pre_root, _ = gpu.coset_lde_split_trees_dev(d, n, 2, blowup, w, 1, want_precomputed_nodes=False)
        d.free()
        w.free()
        h.close()
        assert pre_root == bytes(want[str(log_h)]), (blowup, log_h)
