"""GPU parity for the field kernels, NTT, coset LDE + Keccak Merkle commit
(through the C ABI) against the oracle on the same seeded inputs."""
import random

import numpy as np
import pytest

from lambda_vm_amd import workload

pytestmark = pytest.mark.gpu
P = workload.P


def canon(a):
    a = np.asarray(a, dtype=np.uint64)
    return np.where(a >= np.uint64(P), a - np.uint64(P), a)


def _edge(rng, k):
    edges = [0, 1, 2, P - 1, P, P + 1, 2**64 - 1, 2**64 - 2, 0xFFFFFFFF, 0x100000000]
    return np.array([rng.choice(edges) if rng.random() < 0.3 else rng.getrandbits(64) for _ in range(k)],
                    dtype=np.uint64)


def test_goldilocks_ops_raw_equal(gpu, ora):
    # raw limbs, non-canonical inputs included (math-cuda/tests/goldilocks.rs:1-3)
    rng = random.Random(3)
    a, b = _edge(rng, 5000), _edge(rng, 5000)
    for op in ("add", "sub", "mul", "sqr", "neg"):
        assert np.array_equal(gpu.gl_op(op, a, b), ora.gl_op(op, a, b)), op
    nz = a[canon(a) != 0]
    assert np.array_equal(gpu.gl_op("inv", nz), ora.gl_op("inv", nz))


def test_fp3_ops_raw_equal(gpu, ora):
    rng = random.Random(4)
    a, b = _edge(rng, 3000).reshape(-1, 3), _edge(rng, 3000).reshape(-1, 3)
    for op in ("add", "sub", "mul", "sqr", "neg"):
        assert np.array_equal(gpu.fp3_op(op, a, b), ora.fp3_op(op, a, b)), op
    nz = a[(canon(a) != 0).any(axis=1)]
    assert np.array_equal(gpu.fp3_op("inv", nz), ora.fp3_op("inv", nz))


@pytest.mark.parametrize("log_n", [1, 2, 3, 5, 8, 9, 12, 16, 17])
@pytest.mark.parametrize("m", [1, 5, 16, 17])
def test_ntt_matches_oracle(gpu, ora, log_n, m):
    n = 1 << log_n
    x = workload.trace_table(9, log_n, n, m)
    assert np.array_equal(canon(gpu.ntt_forward(x)), canon(ora.fft_two_half(x)))
    assert np.array_equal(canon(gpu.ntt_inverse(x)), canon(ora.fft_two_half(x, inverse=True)))


def test_ntt_closed_form(gpu, ora):
    n = 1 << 10
    c = np.zeros(n, dtype=np.uint64)
    c[0] = c[1] = 1
    w = ora.root_of_unity(10)
    want = [(1 + pow(w, i, P)) % P for i in range(n)]
    assert canon(gpu.ntt_forward(c)).tolist() == want


def _lde_case(gpu, ora, n, m, blowup, offset, cfg=2, table=0):
    tr = workload.trace_table(cfg, table, n, m)
    w = ora.lde_weights(n, offset)
    h, host = gpu.coset_lde_row_major_with_merkle_tree_keep(tr, n, m, blowup, w, retain_host_lde=True)
    ref_lde, ref_nodes = ora.lde_commit(tr, blowup, offset)
    return tr, w, h, host, ref_lde, ref_nodes


@pytest.mark.parametrize("n,m,blowup", [(2, 1, 1), (2, 3, 2), (4, 1, 4), (8, 16, 2), (16, 17, 8), (64, 5, 4),
                                        (256, 33, 2), (512, 8, 4), (1024, 2, 16), (4096, 24, 4), (1 << 13, 7, 2)])
def test_lde_commit_small_shapes(gpu, ora, n, m, blowup):
    tr, w, h, host, ref_lde, ref_nodes = _lde_case(gpu, ora, n, m, blowup, 7)
    assert np.array_equal(canon(host), canon(ref_lde))
    assert h.root == bytes(ref_nodes[0])
    assert np.array_equal(h.nodes(), ref_nodes)
    h.close()


@pytest.mark.parametrize("offset", [7, 3])
def test_plain_commit_2_16x32(gpu, ora, offset):
    # 2^16 rows x 32 columns, blowup 4, coset offset 7 (multiplicative
    # generator) and 3 (the production coset offset): root and every LDE
    # column bit-exact (canonical)
    tr, w, h, host, ref_lde, ref_nodes = _lde_case(gpu, ora, 1 << 16, 32, 4, offset, cfg=0)
    assert h.root == bytes(ref_nodes[0])
    assert np.array_equal(canon(host), canon(ref_lde))
    # openings: auth paths verify against the root, opened rows are the
    # bit-reversed LDE rows of the reference layout
    rng = np.random.default_rng(1)
    pos = rng.integers(0, h.lde_rows // 2, 40).astype(np.uint64)
    paths, rows = h.open(pos)
    bits = (h.lde_rows).bit_length() - 1
    for k, p in enumerate(pos.tolist()):
        r0 = int(format(2 * p, f"0{bits}b")[::-1], 2)
        r1 = int(format(2 * p + 1, f"0{bits}b")[::-1], 2)
        assert np.array_equal(canon(rows[k, 0]), canon(ref_lde[r0]))
        assert np.array_equal(canon(rows[k, 1]), canon(ref_lde[r1]))
        assert ora.merkle_verify(h.root, bytes(ref_nodes[h.lde_rows // 2 - 1 + p]), p, paths[k])
    h.close()


def test_lde_ext3_commit(gpu, ora):
    n, m, blowup = 256, 3, 4
    tr = workload.trace_table(3, 0, n, 3 * m)
    w = ora.lde_weights(n, 3)
    h, host = gpu.coset_lde_ext3_row_major_with_merkle_tree_keep(tr, n, m, blowup, w)
    ref = ora.coset_lde_row_major(tr, blowup, 3)
    nodes = ora.commit_rows_bit_reversed(ref, elem_words=3)
    assert np.array_equal(canon(host), canon(ref))
    assert h.root == bytes(nodes[0])


def test_device_resident_input_same_root(gpu, ora):
    n, m, blowup = 1 << 12, 9, 4
    tr = workload.trace_table(4, 0, n, m)
    w = ora.lde_weights(n, 7)
    dt = gpu.DeviceBuffer.from_array(tr)
    dw = gpu.DeviceBuffer.from_array(w)
    h = gpu.coset_lde_dev(dt, n, m, blowup, dw)
    _, ref_nodes = ora.lde_commit(tr, blowup, 7, want_lde=False)
    assert h.root == bytes(ref_nodes[0])
    assert np.array_equal(canon(h.download()), canon(ora.coset_lde_row_major(tr, blowup, 7)))


@pytest.mark.parametrize("num", [1, 2, 3, 5, 1000, 4097])
def test_build_merkle_tree(gpu, ora, num):
    rng = np.random.default_rng(num)
    leaves = rng.integers(0, 256, size=(num, 32), dtype=np.uint8)
    assert np.array_equal(gpu.build_merkle_tree(leaves), ora.merkle_build(leaves))


def test_injected_fault_is_a_status_not_an_abort(gpu, ora):
    from lambda_vm_amd._native import LvmError

    x = workload.trace_table(5, 0, 64, 2)
    gpu.inject_fault_after(0)
    with pytest.raises(LvmError) as ei:
        gpu.ntt_forward(x)
    assert ei.value.code == 4
    # the next call runs normally
    assert np.array_equal(canon(gpu.ntt_forward(x)), canon(ora.fft_two_half(x)))


def test_bad_shapes_are_rejected(gpu):
    from lambda_vm_amd._native import LvmError

    with pytest.raises(LvmError):
        gpu.coset_lde_row_major_with_merkle_tree_keep(np.zeros(3 * 2, np.uint64), 3, 2, 4, np.zeros(3, np.uint64))
    with pytest.raises(LvmError):
        gpu.coset_lde_row_major_with_merkle_tree_keep(np.zeros(8, np.uint64), 4, 2, 3, np.zeros(4, np.uint64))


def test_device_coset_weights(gpu, ora):
    for n, g in ((1, 7), (8, 3), (1 << 12, 7)):
        b = gpu.coset_weights_dev(n, g)
        assert np.array_equal(canon(b.to_array(np.uint64, (n,))), canon(ora.lde_weights(n, g)))
