"""GPU FRI (fold + pair-leaf Keccak + layer trees, commit phase with the host
transcript, query phase) against the oracle."""
import numpy as np
import pytest

from lambda_vm_amd import workload

pytestmark = pytest.mark.gpu
P = np.uint64(workload.P)


def canon(a):
    a = np.asarray(a, dtype=np.uint64)
    return np.where(a >= P, a - P, a)


def _codeword(ora, log_n, blowup_log, offset, seed):
    """bit-reversed Fp3 codeword of a random polynomial of degree < N/blowup"""
    N = 1 << log_n
    n = N >> blowup_log
    coeffs = workload.fp3_values(seed, n)  # (n, 3)
    # evaluate componentwise via the base LDE of the three coefficient columns
    t = ora.fft_two_half(coeffs, inverse=False)  # evaluations of sum c_i x^i at w_n^k
    lde = ora.coset_lde_row_major(t, 1 << blowup_log, offset)  # (N, 3) natural order
    br = np.array([int(format(i, f"0{log_n}b")[::-1], 2) for i in range(N)]) if log_n else np.zeros(1, int)
    return np.ascontiguousarray(lde[br])


def test_single_layer_fold_commit(gpu, ora):
    log_n, g = 10, 7
    cw = workload.fp3_values(11, 1 << log_n)
    zeta = workload.fp3_values(12, 1)[0]
    f = gpu.GpuFri(cw, g)
    root = f.fold_and_commit_layer(zeta)
    tw = ora.fri_inv_twiddles(g, 1 << log_n)
    folded = ora.fri_fold(cw, zeta, tw)
    cur, off = f.current()
    assert np.array_equal(canon(cur), canon(folded))
    assert off == pow(g, 2, workload.P)
    nodes = ora.merkle_build(ora.fri_layer_leaves(folded))
    assert root == bytes(nodes[0])
    ev, gn = f.layer(0)
    assert np.array_equal(gn, nodes)
    f.close()


@pytest.mark.parametrize("log_n,blowup_log,k,offset", [(3, 1, 1, 7), (4, 2, 1, 3), (8, 2, 3, 7), (12, 1, 7, 3),
                                                       (14, 2, 6, 7), (16, 3, 2, 3), (18, 2, 7, 7)])
def test_commit_phase_matches_oracle(gpu, ora, log_n, blowup_log, k, offset):
    cw = _codeword(ora, log_n, blowup_log, offset, 100 + log_n)
    ta, tb = gpu.Transcript(b"fri-test"), ora.Transcript(b"fri-test")
    f = gpu.GpuFri(cw, offset)
    roots, coeffs = f.commit_phase(ta, blowup_log, k)
    rroots, rcoeffs, rlayers = ora.fri_commit(cw, tb, offset, blowup_log, k, keep_layers=True)
    assert roots.shape == rroots.shape and np.array_equal(roots, rroots)
    assert np.array_equal(canon(coeffs), canon(rcoeffs))
    assert ta.state() == tb.state()
    # queries: symmetric evaluations and authentication paths per layer
    if f.num_layers():
        rng = np.random.default_rng(log_n)
        iotas = rng.integers(0, (1 << log_n) // 2, 16).astype(np.uint64)
        sym, paths = f.query(iotas)
        off = 0
        for li in range(f.num_layers()):
            size = (1 << log_n) >> (li + 1)
            ev = rlayers[off:off + size]
            off += size
            nodes = ora.merkle_build(ora.fri_layer_leaves(ev))
            leaves = ora.fri_layer_leaves(ev)
            for q, iota in enumerate(iotas.tolist()):
                idx = iota >> li
                assert np.array_equal(canon(sym[li, q]), canon(ev[idx ^ 1]))
                assert ora.merkle_verify(bytes(nodes[0]), bytes(leaves[idx >> 1]), idx >> 1, paths[li][q])
    f.close()


def test_low_degree_codeword_terminal_is_exact(gpu, ora):
    # a codeword of degree < 2^k folds to a terminal whose coefficients are the
    # folded polynomial: the oracle pins that identity; here the GPU result
    # must equal the oracle and the high coefficients beyond degree 2^k vanish
    log_n, blowup_log, k = 12, 2, 4
    cw = _codeword(ora, log_n, blowup_log, 7, 7)
    ta, tb = gpu.Transcript(b"x"), ora.Transcript(b"x")
    f = gpu.GpuFri(cw, 7)
    _, coeffs = f.commit_phase(ta, blowup_log, k)
    _, rcoeffs, _ = ora.fri_commit(cw, tb, 7, blowup_log, k)
    assert np.array_equal(canon(coeffs), canon(rcoeffs))


def test_fault_mid_commit_restores_transcript(gpu, ora):
    log_n = 12
    cw = _codeword(ora, log_n, 2, 7, 5)
    t = gpu.Transcript(b"snap")
    before = t.state()
    f = gpu.GpuFri(cw, 7)
    from lambda_vm_amd._native import LvmError

    gpu.inject_fault_after(3)
    with pytest.raises(LvmError):
        f.commit_phase(t, 2, 3)
    gpu.inject_fault_after(-1)
    assert t.state() == before
    f.close()


def test_fri_on_device_codeword(gpu, ora):
    log_n = 13
    cw = _codeword(ora, log_n, 2, 3, 9)
    dev = gpu.DeviceBuffer.from_array(cw)
    f = gpu.GpuFri(None, 3, on_device=dev, length=1 << log_n)
    ta, tb = gpu.Transcript(b"d"), ora.Transcript(b"d")
    roots, coeffs = f.commit_phase(ta, 2, 5)
    rroots, rcoeffs, _ = ora.fri_commit(cw, tb, 3, 2, 5)
    assert np.array_equal(roots, rroots) and np.array_equal(canon(coeffs), canon(rcoeffs))
