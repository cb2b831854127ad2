"""Proving service on the GPU: each job's result digest is the same alone and
in a mix of concurrently admitted jobs, the same with tables proved on one
host thread and on a pool of threads (one library stream each), the same
under tick and work-conserving dispatch, and its committed roots equal the
oracle's commitment of the same synthetic tables."""
import numpy as np
import pytest

from lambda_vm_amd import serve, workload

pytestmark = pytest.mark.gpu


def _jobs():
    shapes = [
        [(1 << 8, 8), (1 << 10, 13), (1 << 9, 20), (1 << 8, 9)],
        [(1 << 11, 16), (1 << 8, 33), (1 << 10, 8), (1 << 12, 10), (1 << 9, 12)],
        [(1 << 10, 24), (1 << 10, 24), (1 << 8, 8), (1 << 11, 17)],
        [(1 << 12, 8), (1 << 9, 40), (1 << 8, 11), (1 << 8, 64), (1 << 10, 9), (1 << 9, 15)],
    ]
    return [serve.JobSpec(j, s) for j, s in enumerate(shapes)]


def test_results_same_alone_and_in_mix(gpu):
    ex = serve.GpuExecutor(offset=3, fri_k=3, num_queries=8)
    jobs = _jobs()
    mix = serve.ProvingService(1 << 40, ex)
    st = mix.run([(0.001 * j, spec) for j, spec in enumerate(jobs)], time_scale=1.0)
    assert st["jobs"] == len(jobs)
    mixed = {js.spec.job_id: (js.digest(), js.roots) for js in mix.done}
    for spec in jobs:
        alone = serve.ProvingService(1 << 40, ex)
        alone.run([(0.0, spec)])
        js = alone.done[0]
        assert js.digest() == mixed[spec.job_id][0], spec.job_id
        assert js.roots == mixed[spec.job_id][1]
    ex.close()


def test_thread_pool_matches_single_thread(gpu):
    jobs = _jobs()
    digests = {}
    for workers in (1, 4):
        ex = serve.GpuExecutor(offset=3, fri_k=3, num_queries=8, workers=workers)
        svc = serve.ProvingService(1 << 40, ex, tables_per_tick=6)
        svc.run([(0.0, spec) for spec in jobs])
        ex.close()
        digests[workers] = {js.spec.job_id: js.digest() for js in svc.done}
        assert all(all(d is not None for d in js.digests) for js in svc.done)
    assert digests[1] == digests[4]


def test_async_dispatch_matches_tick(gpu):
    """Work-conserving dispatch (commits and tables on the pool as threads free
    up) gives every job the digest the tick form gives it."""
    jobs = _jobs()
    ex = serve.GpuExecutor(offset=3, fri_k=3, num_queries=8, workers=4)
    tick = serve.ProvingService(1 << 40, ex)
    tick.run([(0.0, spec) for spec in jobs])
    asy = serve.ProvingService(1 << 40, ex)
    st = asy.run_async([(0.002 * j, spec) for j, spec in enumerate(jobs)])
    ex.close()
    assert st["jobs"] == len(jobs) and asy.in_use == 0
    assert {js.spec.job_id: js.digest() for js in asy.done} == {js.spec.job_id: js.digest() for js in tick.done}


def test_service_roots_match_oracle(gpu, ora):
    ex = serve.GpuExecutor(offset=3, fri_k=3, num_queries=8)
    spec = _jobs()[1]
    svc = serve.ProvingService(1 << 40, ex)
    svc.run([(0.0, spec)])
    js = svc.done[0]
    for ti, (n, m) in enumerate(spec.tables):
        tr = workload.trace_table_modp(spec.cfg, spec.job_id * 64 + ti, n, m)
        _, nodes = ora.lde_commit(tr, spec.blowup, 3, want_lde=False)
        assert js.roots[ti] == bytes(nodes[0]), ti
    ex.close()


def test_device_generator_matches_host(gpu):
    for n, m, cfg, table in [(1, 1, 0, 0), (300, 7, 5, 129), (4096, 33, 3, 2)]:
        d = gpu.splitmix_table_dev(n, m, cfg, table)
        got = d.to_array(np.uint64, (n, m))
        assert np.array_equal(got, workload.trace_table_modp(cfg, table, n, m))
