"""Device buffer cache and per-thread streams: commitments are unchanged after
the cache is handed back (lvm_trim) and between calls that reuse released
blocks, and tables committed from several host threads at once (one library
stream each) give the roots of the same tables committed one after another."""
from concurrent.futures import ThreadPoolExecutor

import pytest

from lambda_vm_amd import _native, workload

pytestmark = pytest.mark.gpu

SHAPES = [(1 << 10, 8), (1 << 12, 13), (1 << 9, 40), (1 << 11, 21), (1 << 8, 64), (1 << 13, 9)]


def _root(gpu, n, m, table):
    tr = gpu.splitmix_table_dev(n, m, 7, table)
    w = gpu.coset_weights_dev(n, 3)
    h = gpu.coset_lde_dev(tr, n, m, 4, w)
    root = h.root
    h.close()
    tr.free()
    w.free()
    return root


def test_roots_stable_across_reuse_and_trim(gpu):
    first = [_root(gpu, n, m, t) for t, (n, m) in enumerate(SHAPES)]
    again = [_root(gpu, n, m, t) for t, (n, m) in enumerate(SHAPES)]  # reuses released blocks
    gpu.trim()
    after = [_root(gpu, n, m, t) for t, (n, m) in enumerate(SHAPES)]
    assert first == again == after


def test_threads_match_sequential(gpu, ora):
    seq = [_root(gpu, n, m, t) for t, (n, m) in enumerate(SHAPES)]
    with ThreadPoolExecutor(max_workers=4, initializer=_native.ensure_device_this_thread, initargs=(0,)) as pool:
        for _ in range(3):
            par = list(pool.map(lambda a: _root(gpu, a[1][0], a[1][1], a[0]), enumerate(SHAPES)))
            assert par == seq
    n, m = SHAPES[1]
    _, nodes = ora.lde_commit(workload.trace_table_modp(7, 1, n, m), 4, 3, want_lde=False)
    assert seq[1] == bytes(nodes[0])
