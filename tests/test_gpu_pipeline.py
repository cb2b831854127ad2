"""End-to-end single-table rounds (commit, OOD, DEEP, FRI, queries, openings)
on the GPU path against a replay of the same rounds by the oracle."""
import numpy as np
import pytest

from lambda_vm_amd import workload

pytestmark = pytest.mark.gpu
P = workload.P


def canon(a):
    a = np.asarray(a, dtype=np.uint64)
    return np.where(a >= np.uint64(P), a - np.uint64(P), a)


@pytest.mark.parametrize("log_n,m,blowup,g,k", [(10, 8, 4, 7, 3), (12, 16, 2, 3, 6), (14, 5, 8, 7, 4),
                                                (16, 64, 4, 7, 6)])
def test_single_table_rounds_match_oracle(gpu, ora, log_n, m, blowup, g, k):
    from lambda_vm_amd import pipeline
    from oracle.pipeline import replay_single_table

    n = 1 << log_n
    tr = workload.trace_table(1, log_n, n, m)
    d_tr = gpu.DeviceBuffer.from_array(tr)
    d_w = gpu.coset_weights_dev(n, g)
    proof = pipeline.prove_single_table(d_tr, n, m, blowup, g, d_w, fri_final_log_degree=k, num_queries=12,
                                        label=b"e2e")
    ref = replay_single_table(tr, blowup, g, k, 12, b"e2e")
    assert proof.root == ref["root"]
    assert proof.z == ref["z"]
    assert np.array_equal(canon(proof.ood), canon(ref["ood"]))
    assert np.array_equal(proof.fri_roots, ref["roots"])
    assert np.array_equal(canon(proof.fri_coeffs), canon(ref["coeffs"]))
    assert proof.iotas.tolist() == ref["iotas"]
    assert proof.transcript_state == ref["state"]
    # trace openings verify against the committed root and carry the LDE rows
    N = n * blowup
    bits = N.bit_length() - 1
    for q, p in enumerate(ref["iotas"]):
        leaf = bytes(ref["nodes"][N // 2 - 1 + p])
        assert ora.merkle_verify(ref["root"], leaf, p, proof.trace_paths[q])
        r0 = int(format(2 * p, f"0{bits}b")[::-1], 2)
        assert np.array_equal(canon(proof.trace_rows[q, 0]), canon(ref["lde"][r0]))
    # FRI symmetric evaluations of every layer
    off = 0
    for li in range(len(ref["roots"])):
        size = N >> (li + 1)
        ev = ref["layers"][off:off + size]
        off += size
        for q, iota in enumerate(ref["iotas"]):
            assert np.array_equal(canon(proof.fri_sym[li, q]), canon(ev[(iota >> li) ^ 1]))


@pytest.mark.parametrize("log_n,m,blowup,k,q", [(8, 8, 4, 2, 10), (11, 13, 2, 3, 7), (6, 4, 4, 6, 5),
                                              (13, 33, 8, 4, 30)])
def test_native_rounds_equal_stepwise(gpu, log_n, m, blowup, k, q):
    """lvm_prove_committed_table gives every field the host-driven sequence
    of library calls gives, and leaves the transcript in the same state."""
    from lambda_vm_amd import pipeline

    n = 1 << log_n
    tr = workload.trace_table(4, log_n, n, m)
    d_tr = gpu.DeviceBuffer.from_array(tr)
    d_w = gpu.coset_weights_dev(n, 3)
    out = []
    for native in (True, False):
        h = gpu.coset_lde_dev(d_tr, n, m, blowup, d_w)
        t = gpu.Transcript(b"rounds")
        t.append_bytes(h.root)
        if native:
            p = pipeline.prove_committed_table(h, t, n, m, blowup, 3, k, q)
        else:
            p, fri = pipeline.prove_committed_table_steps(h, t, n, m, blowup, 3, k, q, keep=True)
            fri.close()
        h.close()
        out.append(p)
    a, b = out
    assert a.root == b.root and a.z == b.z and a.transcript_state == b.transcript_state
    for f in ("ood", "fri_roots", "fri_coeffs", "iotas", "fri_sym", "trace_paths", "trace_rows"):
        assert np.array_equal(canon(getattr(a, f)) if getattr(a, f).dtype == np.uint64 else getattr(a, f),
                              canon(getattr(b, f)) if getattr(b, f).dtype == np.uint64 else getattr(b, f)), f
    assert len(a.fri_paths) == len(b.fri_paths)
    for x, y in zip(a.fri_paths, b.fri_paths):
        assert np.array_equal(x, y)
