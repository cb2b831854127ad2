"""Proving service scheduling (CPU, simulated executor): HBM-budget admission,
continuous batching of arrivals into shared commit batches, FIFO order,
latency accounting, and rejection of jobs larger than the budget."""
import pytest

from lambda_vm_amd import serve


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


class FakeExecutor:
    """Advances a simulated clock by work proportional to LDE cells."""

    def __init__(self, clock, rate=1e9):
        self.clock, self.rate = clock, rate
        self.commits = []

    def commit(self, jobs):
        self.commits.append([js.spec.job_id for js in jobs])
        cells = sum(js.spec.lde_cells for js in jobs)
        self.clock.t += 20e-6 + cells / self.rate
        for js in jobs:
            js.handles = [object() for _ in js.spec.tables]
            js.roots = [b"r" * 32 for _ in js.spec.tables]

    def prove_table(self, js, ti):
        n, m = js.spec.tables[ti]
        self.clock.t += 50e-6 + 2 * n * m * js.spec.blowup / self.rate
        return bytes([js.spec.job_id % 256, ti])


def _service(budget, rate=1e9):
    clock = FakeClock()
    return serve.ProvingService(budget, FakeExecutor(clock, rate), clock=clock), clock


def _run_sim(svc, clock, arrivals):
    pending = sorted(arrivals, key=lambda a: a[0])
    total = len(pending)
    while len(svc.done) < total:
        while pending and pending[0][0] <= clock.t:
            t, spec = pending.pop(0)
            svc.submit(spec, t)
        if not svc.tick():
            clock.t = pending[0][0]
    return svc.stats(max(js.finished for js in svc.done))


def test_admission_respects_budget_and_batches_arrivals():
    arrivals = serve.synth_jobs(rate=400.0, count=60, seed=3, max_log_cells=24)
    budget = 3 * max(serve.job_hbm_bytes(s) for _, s in arrivals)
    svc, clock = _service(budget)
    st = _run_sim(svc, clock, arrivals)
    assert st["jobs"] == 60
    assert svc.peak_in_use <= budget
    assert svc.in_use == 0
    # under load, several jobs share one commit batch
    assert max(svc.batch_sizes) > 1
    # FIFO admission: commit batches list job ids in arrival order
    order = [j for batch in svc.ex.commits for j in batch]
    assert order == sorted(order)
    assert st["p99_ms"] >= st["p50_ms"] > 0
    for js in svc.done:
        assert js.finished >= js.admitted >= js.arrival
        assert len(js.digests) == len(js.spec.tables)


def test_tight_budget_serialises_jobs():
    arrivals = serve.synth_jobs(rate=1000.0, count=12, seed=4, max_log_cells=22)
    budget = max(serve.job_hbm_bytes(s) for _, s in arrivals)  # one job at a time at worst
    svc, clock = _service(budget)
    _run_sim(svc, clock, arrivals)
    assert svc.peak_in_use <= budget
    # the big jobs never overlapped
    spans = sorted((js.admitted, js.finished, js.hbm) for js in svc.done)
    for (a0, f0, h0), (a1, f1, h1) in zip(spans, spans[1:]):
        if h0 + h1 > budget:
            assert a1 >= f0


def test_oversized_job_rejected():
    svc, _ = _service(1 << 20)
    with pytest.raises(ValueError):
        svc.submit(serve.JobSpec(0, [(1 << 16, 32)]), 0.0)


def test_job_shapes():
    arrivals = serve.synth_jobs(rate=10.0, count=200, seed=7)
    for t, spec in arrivals:
        assert 4 <= len(spec.tables) <= 20
        for n, m in spec.tables:
            assert n & (n - 1) == 0 and 8 <= m <= 64
    ts = [t for t, _ in arrivals]
    assert ts == sorted(ts)
    # mean gap close to 1/rate
    assert 0.07 < ts[-1] / len(ts) < 0.13


def test_oldest_first_policy_lowers_median_latency():
    arrivals = serve.synth_jobs(rate=300.0, count=80, seed=9, max_log_cells=24)
    budget = 8 * max(serve.job_hbm_bytes(s) for _, s in arrivals)
    stats = {}
    for policy in ("round_robin", "oldest"):
        clock = FakeClock()
        svc = serve.ProvingService(budget, FakeExecutor(clock, rate=5e8), clock=clock, policy=policy)
        stats[policy] = _run_sim(svc, clock, arrivals)
        assert svc.peak_in_use <= budget
        # same work either way: every job proved every table
        assert all(len(js.digests) == len(js.spec.tables) for js in svc.done)
    assert stats["oldest"]["p50_ms"] <= stats["round_robin"]["p50_ms"]


class ThreadedFakeExecutor:
    """A worker pool doing token work (a short sleep) per commit and table."""

    def __init__(self, workers=3):
        from concurrent.futures import ThreadPoolExecutor
        self.workers = workers
        self.pool = ThreadPoolExecutor(max_workers=workers)
        self.commits = []

    def commit(self, jobs):
        import time
        self.commits.append([js.spec.job_id for js in jobs])
        time.sleep(0.0005)
        for js in jobs:
            js.handles = [object() for _ in js.spec.tables]
            js.roots = [b"r" * 32 for _ in js.spec.tables]

    def prove_table(self, js, ti):
        import time
        assert js.handles, "table proved before its job was committed"
        time.sleep(0.0002)
        return bytes([js.spec.job_id % 256, ti])


def test_async_dispatch_completes_every_job_within_budget():
    arrivals = serve.synth_jobs(rate=2000.0, count=40, seed=5, max_log_cells=22)
    budget = 2 * max(serve.job_hbm_bytes(s) for _, s in arrivals)
    ex = ThreadedFakeExecutor()
    svc = serve.ProvingService(budget, ex)
    st = svc.run_async(arrivals)
    ex.pool.shutdown()
    assert st["jobs"] == 40 and svc.in_use == 0 and svc.peak_in_use <= budget
    order = [j for batch in ex.commits for j in batch]
    assert order == sorted(order)  # FIFO admission
    for js in svc.done:
        assert js.finished >= js.admitted >= js.arrival
        assert js.digests == [bytes([js.spec.job_id % 256, ti]) for ti in range(len(js.spec.tables))]
    assert st["p99_ms"] >= st["p50_ms"] > 0


def test_async_dispatch_needs_a_pool():
    svc, _ = _service(1 << 40)
    with pytest.raises(ValueError):
        svc.run_async([(0.0, serve.JobSpec(0, [(256, 8)]))])


def test_async_dispatch_batches_commits_under_load():
    """Jobs admitted while a commit runs join the next one: at most one
    commit is in flight, so a burst of arrivals shares commits, while a job
    arriving at an idle service commits alone at once."""
    class SlowCommit(ThreadedFakeExecutor):
        def __init__(self):
            super().__init__(workers=4)
            self.active = 0
            self.max_active = 0

        def commit(self, jobs):
            import time
            self.active += 1
            self.max_active = max(self.max_active, self.active)
            time.sleep(0.01)
            super().commit(jobs)
            self.active -= 1

    arrivals = serve.synth_jobs(rate=3000.0, count=30, seed=8, max_log_cells=22)
    ex = SlowCommit()
    svc = serve.ProvingService(1 << 50, ex)
    st = svc.run_async(arrivals)
    ex.pool.shutdown()
    assert st["jobs"] == 30 and ex.max_active == 1
    assert st["mean_jobs_per_commit_batch"] > 1.5
    assert [j for b in ex.commits for j in b] == list(range(30))
    lone = SlowCommit()
    svc = serve.ProvingService(1 << 50, lone)
    svc.run_async([(0.0, serve.JobSpec(0, [(256, 8)] * 4)), (0.5, serve.JobSpec(1, [(256, 8)] * 4))])
    lone.pool.shutdown()
    assert lone.commits == [[0], [1]]
