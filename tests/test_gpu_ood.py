"""GPU batch inversion, barycentric OOD evaluation and DEEP composition
against the oracle on the same seeded inputs."""
import numpy as np
import pytest

import pyref
from lambda_vm_amd import workload

pytestmark = pytest.mark.gpu
P = np.uint64(workload.P)


def canon(a):
    a = np.asarray(a, dtype=np.uint64)
    return np.where(a >= P, a - P, a)


def bitrev_perm(log_n):
    return np.array([int(format(i, f"0{log_n}b")[::-1], 2) for i in range(1 << log_n)]) if log_n else np.zeros(1, int)


@pytest.mark.parametrize("n", [1, 2, 3, 255, 256, 257, 10000, 1 << 18])
def test_batch_inverse(gpu, ora, n):
    x = workload.fp3_values(31 + n, n)
    got = gpu.batch_inverse_ext3(x)
    assert np.array_equal(canon(got), canon(ora.batch_inverse_fp3(x)))
    if n <= 300:
        for a, b in zip(x.tolist(), got.tolist()):
            assert pyref.f3mul(a, b) == [1, 0, 0]


@pytest.mark.parametrize("log_n,m,blowup,offset,K", [(4, 3, 4, 7, 1), (8, 5, 2, 3, 2), (10, 64, 4, 7, 2),
                                                     (12, 17, 8, 3, 8), (14, 70, 4, 7, 3)])
def test_barycentric_multi(gpu, ora, log_n, m, blowup, offset, K):
    n = 1 << log_n
    tr = workload.trace_table(6, log_n, n, m)
    w = ora.lde_weights(n, offset)
    h, host = gpu.coset_lde_row_major_with_merkle_tree_keep(tr, n, m, blowup, w)
    zs = workload.fp3_values(77 + log_n, K)
    got = gpu.barycentric_base_multi(h, offset, zs)
    for k in range(K):
        ref = ora.barycentric_base(host, n, blowup, offset, zs[k])
        assert np.array_equal(canon(got[k]), canon(ref)), k
    h.close()


def test_barycentric_equals_trace_interpolant(gpu, ora):
    # the OOD value is the trace interpolant at z (Horner on the coefficients)
    n, m, blowup, g = 32, 2, 4, 7
    tr = workload.trace_table(8, 0, n, m)
    h, _ = gpu.coset_lde_row_major_with_merkle_tree_keep(tr, n, m, blowup, ora.lde_weights(n, g))
    z = [5, 6, 7]
    got = gpu.barycentric_base_multi(h, g, np.array([z], dtype=np.uint64))
    wn = ora.root_of_unity(5)
    ninv = pow(n, workload.P - 2, workload.P)
    for c in range(m):
        coeffs = [v * ninv % workload.P for v in pyref.dft([int(v) for v in tr[:, c]], pow(wn, workload.P - 2, workload.P))]
        assert [int(v) % workload.P for v in got[0, c]] == pyref.horner3([[cf, 0, 0] for cf in coeffs], z)


@pytest.mark.parametrize("with_aux,with_parts,K,geo", [(False, False, 2, False), (True, False, 1, False),
                                                       (True, True, 2, False), (False, True, 4, False),
                                                       (True, True, 3, True), (False, True, 4, True),
                                                       (True, False, 3, False)])
def test_deep_composition(gpu, ora, with_aux, with_parts, K, geo):
    """geo: trace-term weights geometric in k past k = 1 (the reference's
    gamma layout), which the library folds into two accumulated sums."""
    log_n, m, ma, np_, blowup, g = 9, 6, 2, 2, 4, 3
    n, N = 1 << log_n, (1 << log_n) * blowup
    w = ora.lde_weights(n, g)
    tr = workload.trace_table(10, 0, n, m)
    hm, lm = gpu.coset_lde_row_major_with_merkle_tree_keep(tr, n, m, blowup, w)
    ha = hp = la = lp = None
    if with_aux:
        ta = workload.trace_table(10, 1, n, 3 * ma)
        ha, la = gpu.coset_lde_ext3_row_major_with_merkle_tree_keep(ta, n, ma, blowup, w)
    if with_parts:
        tp = workload.trace_table(10, 2, n, 3 * np_)
        hp, lp = gpu.coset_lde_ext3_row_major_with_merkle_tree_keep(tp, n, np_, blowup, w)
    cols = m + (ma if with_aux else 0)
    zp = workload.fp3_values(91, K)
    tood = workload.fp3_values(92, cols * K)
    gtr = workload.fp3_values(93, cols * K)
    if geo:
        r = [int(v) for v in workload.fp3_values(97, 1)[0]]
        g3 = gtr.reshape(cols, K, 3)
        for j in range(cols):
            for k in range(2, K):
                g3[j, k] = pyref.f3mul([int(v) for v in g3[j, k - 1]], r)
        gtr = g3.reshape(-1, 3)
    zpa = workload.fp3_values(94, 1)[0]
    pood = workload.fp3_values(95, np_)
    gh = workload.fp3_values(96, np_)
    kw = dict(aux=ha, parts=hp, z_parts=zpa if with_parts else None, parts_ood=pood if with_parts else None,
              gammas_h=gh if with_parts else None)
    buf = gpu.deep_composition(hm, g, zp, tood, gtr, **kw)
    got = buf.to_array(np.uint64, (N, 3))
    ref = ora.deep_composition(lm, g, zp, tood, gtr, aux=la, parts=lp, z_parts=kw["z_parts"],
                               parts_ood=kw["parts_ood"], gammas_h=kw["gammas_h"])
    assert np.array_equal(canon(got), canon(ref[bitrev_perm(N.bit_length() - 1)]))


def test_deep_composition_large_domain(gpu, ora):
    """2^20 LDE points: the pole factors take the sixteen-rows-per-thread path."""
    log_n, m, blowup, g, K = 18, 3, 4, 3, 2
    n, N = 1 << log_n, (1 << log_n) * blowup
    tr = workload.trace_table(11, 0, n, m)
    hm, lm = gpu.coset_lde_row_major_with_merkle_tree_keep(tr, n, m, blowup, ora.lde_weights(n, g))
    zp = workload.fp3_values(81, K)
    tood = workload.fp3_values(82, m * K)
    gtr = workload.fp3_values(83, m * K)
    got = gpu.deep_composition(hm, g, zp, tood, gtr).to_array(np.uint64, (N, 3))
    ref = ora.deep_composition(lm, g, zp, tood, gtr)
    perm = bitrev_perm(N.bit_length() - 1)
    assert np.array_equal(canon(got), canon(ref[perm]))


def test_deep_quotient_identity(gpu, ora):
    # deep(x) * (x - z) = sum_j gamma_j (t_j(x) - t_j(z)) for one pole: check
    # at a few rows with the exact OOD values, using big-integer arithmetic
    log_n, m, blowup, g = 6, 3, 4, 7
    n, N = 1 << log_n, (1 << log_n) * blowup
    tr = workload.trace_table(12, 0, n, m)
    h, lde = gpu.coset_lde_row_major_with_merkle_tree_keep(tr, n, m, blowup, ora.lde_weights(n, g))
    z = workload.fp3_values(3, 1)
    ood = gpu.barycentric_base_multi(h, g, z)[0]  # (m, 3)
    gam = workload.fp3_values(4, m)
    got = gpu.deep_composition(h, g, z, ood, gam).to_array(np.uint64, (N, 3))
    nat = got[bitrev_perm(N.bit_length() - 1)]
    wN = ora.root_of_unity(N.bit_length() - 1)
    for i in (0, 1, 17, N - 1):
        x = g * pow(wN, i, workload.P) % workload.P
        lhs = pyref.f3mul(nat[i].tolist(), pyref.f3sub([x, 0, 0], z[0].tolist()))
        rhs = [0, 0, 0]
        for j in range(m):
            rhs = pyref.f3add(rhs, pyref.f3mul(gam[j].tolist(), pyref.f3sub([int(lde[i, j]), 0, 0], ood[j].tolist())))
        assert lhs == rhs
