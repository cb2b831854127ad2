"""Batched ragged commitment (lvm_batch_*) against the oracle and against the
per-table path on the same seeded inputs: roots, LDE tables and trees of
every table, with and without graph replay, after rebinding inputs and
after handing outputs over."""
import numpy as np
import pytest

from lambda_vm_amd import workload

pytestmark = pytest.mark.gpu
P = workload.P

# ragged shapes (n, m, blowup): tiny and non-multiple-of-16 widths, blowups
# 1..8, one table with more than 2^14 leaves (wide Merkle levels)
SHAPES = [(1 << 4, 3, 2), (1 << 10, 17, 4), (1 << 7, 40, 2), (1 << 12, 8, 4), (1 << 9, 1, 8), (1 << 11, 33, 1),
          (4, 5, 4), (2, 2, 2), (1 << 13, 16, 2), (1 << 14, 9, 4), (1 << 8, 120, 4)]


def canon(a):
    a = np.asarray(a, dtype=np.uint64)
    return np.where(a >= np.uint64(P), a - np.uint64(P), a)


def _inputs(gpu, shapes, cfg, offset=7):
    host, tables = [], []
    for t, (n, m, b) in enumerate(shapes):
        tr = workload.trace_table(cfg, t, n, m)
        host.append(tr)
        tables.append((gpu.DeviceBuffer.from_array(tr), gpu.coset_weights_dev(n, offset), n, m, b))
    return host, tables


def _per_table_roots(gpu, tables):
    roots = []
    for tr, w, n, m, b in tables:
        h = gpu.coset_lde_dev(tr, n, m, b, w)
        roots.append(h.root)
        h.close()
    return roots


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_roots_match_oracle_and_per_table(gpu, ora, use_graph):
    host, tables = _inputs(gpu, SHAPES, cfg=3)
    plan = gpu.BatchPlan(tables, use_graph=use_graph)
    roots = plan.run()
    for t, (n, m, b) in enumerate(SHAPES):
        _, ref_nodes = ora.lde_commit(host[t], b, 7, want_lde=False)
        assert roots[t] == bytes(ref_nodes[0]), (t, n, m, b)
    assert roots == _per_table_roots(gpu, tables)
    # replay: same inputs, same roots
    assert plan.run() == roots
    plan.close()


def test_batch_rebind_and_take(gpu, ora):
    _, tables = _inputs(gpu, SHAPES, cfg=3)
    plan = gpu.BatchPlan(tables, use_graph=True)
    first = plan.run()
    host2, tables2 = _inputs(gpu, SHAPES, cfg=5)
    plan.bind(tables2)
    second = plan.run()
    assert second != first
    assert second == _per_table_roots(gpu, tables2)
    # hand two tables over; their handles hold the full LDE and tree
    for t in (1, 9):
        n, m, b = SHAPES[t]
        h = plan.take(t)
        ref_lde, ref_nodes = ora.lde_commit(host2[t], b, 7)
        assert h.root == second[t]
        assert np.array_equal(canon(h.download()), canon(ref_lde))
        assert np.array_equal(h.nodes(), ref_nodes)
        h.close()
    # the plan gets fresh outputs for the taken tables and replays unchanged
    assert plan.run() == second
    plan.bind(tables)
    assert plan.run() == first
    plan.close()


def test_batch_one_shot(gpu, ora):
    host, tables = _inputs(gpu, SHAPES[:6], cfg=7)
    roots, handles = gpu.batch_lde_commit(tables)
    for t, (n, m, b) in enumerate(SHAPES[:6]):
        ref_lde, ref_nodes = ora.lde_commit(host[t], b, 7)
        assert roots[t] == bytes(ref_nodes[0])
        assert handles[t].root == roots[t]
        paths, rows = handles[t].open([0, (n * b) // 2 - 1])
        assert rows.shape[0] == 2
        handles[t].close()


def test_batch_single_table_and_errors(gpu):
    _, tables = _inputs(gpu, [(1 << 10, 4, 4)], cfg=3)
    plan = gpu.BatchPlan(tables)
    assert plan.run() == _per_table_roots(gpu, tables)
    _, other = _inputs(gpu, [(1 << 9, 4, 4)], cfg=3)
    with pytest.raises(gpu.LvmError):
        plan.bind(other)  # another shape
    with pytest.raises(gpu.LvmError):
        plan.take(1)  # out of range
    plan.close()
    with pytest.raises(gpu.LvmError):
        gpu.BatchPlan([(tables[0][0], tables[0][1], 1 << 10, 4, 3)])  # blowup not a power of two


def test_ntt_plan_override_keeps_results(gpu):
    _, tables = _inputs(gpu, [(1 << 12, 8, 4), (1 << 13, 5, 2)], cfg=3)
    base = _per_table_roots(gpu, tables)
    keep12, keep13 = gpu.get_ntt_plan(12), gpu.get_ntt_plan(13)
    try:
        gpu.set_ntt_plan(12, [4, 4, 4])
        gpu.set_ntt_plan(13, [5, 8])
        assert gpu.get_ntt_plan(12) == [4, 4, 4]
        assert _per_table_roots(gpu, tables) == base
        plan = gpu.BatchPlan(tables)
        assert plan.run() == base
        plan.close()
        with pytest.raises(gpu.LvmError):
            gpu.set_ntt_plan(12, [8, 8])  # does not sum to 12
        gpu.set_ntt_plan(12, [])  # an empty split restores the built-in default
        assert sum(gpu.get_ntt_plan(12)) == 12
    finally:
        # back to whatever was installed before (the tuned splits, if any)
        gpu.set_ntt_plan(12, keep12)
        gpu.set_ntt_plan(13, keep13)
    assert gpu.get_ntt_plan(12) == keep12 and gpu.get_ntt_plan(13) == keep13
