"""Proving service: continuous batching of multi-table proving jobs under an
HBM budget.

The reference bounds concurrent table proving by device memory
(crypto/stark/src/prover.rs:682-693 `estimate_table_vram_bytes`, :703-789
`VramGate` / `run_admitted`, heaviest tables first) and commits all tables of
a proof before forking per-table transcripts (prover.rs:3426-3538). Here the
same ideas serve a stream of jobs:
  - a job (4-20 tables) arrives, waits in FIFO order and is admitted once its
    estimated HBM footprint fits the budget beside the jobs in flight;
  - every scheduler tick commits the tables of all jobs admitted since the
    previous tick in one batched launch sequence (lvm_batch_*), so jobs that
    arrive while others are on the GPU share launches with each other;
  - committed jobs then advance through the remaining rounds (OOD, DEEP,
    FRI, queries) a few tables per tick, oldest job first (or one table of
    every job per tick), and free each table's LDE, tree and FRI layers as
    soon as its rounds finish;
  - the GPU executor proves a tick's tables on a small pool of host threads,
    each with its own library stream (the reference drives tables from
    several OS threads with per-table streams, prover.rs:620-638, :749-782),
    and commits the newly admitted jobs concurrently with them;
  - per job: latency from arrival to its final digest; per run: p50/p99
    latency, jobs/s and LDE cells/s.
A job's result depends only on its own inputs and transcript: the digest of
a job is the same alone and in any mix.
"""
from __future__ import annotations

import hashlib
import math
import random
import time
from collections import deque
from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait
from dataclasses import dataclass, field

import numpy as np

P = 0xFFFFFFFF00000001


@dataclass
class JobSpec:
    job_id: int
    tables: list  # (n, m)
    blowup: int = 4
    cfg: int = 5  # synthetic data stream family

    @property
    def lde_cells(self) -> int:
        return sum(n * m * self.blowup for n, m in self.tables)


def table_hbm_bytes(n: int, m: int, blowup: int) -> int:
    """Device bytes of one table while it is proved: trace, coefficients,
    LDE, Merkle tree, weights, DEEP codeword and FRI layers with their trees
    (the shape of prover.rs:682-693 for this pipeline)."""
    N = n * blowup
    return 16 * n * m + 8 * n + 8 * N * m + 32 * N + 2 * 24 * N + 2 * 32 * N


def job_hbm_bytes(spec: JobSpec) -> int:
    return sum(table_hbm_bytes(n, m, spec.blowup) for n, m in spec.tables)


def synth_jobs(rate: float, count: int, seed: int = 1, min_log_cells: int = 20, max_log_cells: int = 28,
               blowup: int = 4) -> list:
    """Poisson arrivals (exponential gaps, mean 1/rate) of jobs with 4-20
    tables and total LDE cells log-uniform in [2^min, 2^max]."""
    rng = random.Random(seed)
    t = 0.0
    out = []
    for j in range(count):
        t += rng.expovariate(rate)
        k = rng.randint(4, 20)
        total = 2.0 ** rng.uniform(min_log_cells, max_log_cells)
        w = [rng.random() + 0.05 for _ in range(k)]
        s = sum(w)
        tables = []
        for x in w:
            m = rng.randint(8, 64)
            cells = total * x / s
            log_n = int(round(math.log2(max(cells / (blowup * m), 1.0))))
            tables.append((1 << min(max(log_n, 8), 22), m))
        out.append((t, JobSpec(j, tables, blowup)))
    return out


@dataclass
class JobState:
    spec: JobSpec
    arrival: float
    admitted: float = 0.0
    finished: float = 0.0
    hbm: int = 0
    handles: list = field(default_factory=list)
    transcripts: list = field(default_factory=list)
    next_table: int = 0   # next table to hand out
    tables_done: int = 0
    digests: list = field(default_factory=list)  # per table, in table order
    roots: list = field(default_factory=list)

    def __post_init__(self):
        if not self.digests:
            self.digests = [None] * len(self.spec.tables)

    def digest(self) -> str:
        h = hashlib.sha256()
        for d in self.digests:
            h.update(d)
        return h.hexdigest()


def proof_digest(proof) -> bytes:
    """Canonical digest of a table proof: every field the verifier reads."""
    h = hashlib.sha256()
    h.update(proof.root)
    for arr in (np.asarray(proof.z, dtype=np.uint64), proof.ood, proof.fri_coeffs, proof.iotas, proof.fri_sym,
                proof.trace_rows):
        a = np.asarray(arr, dtype=np.uint64)
        h.update(np.where(a >= np.uint64(P), a - np.uint64(P), a).tobytes())
    h.update(np.asarray(proof.fri_roots).tobytes())
    for p in proof.fri_paths:
        h.update(np.asarray(p).tobytes())
    h.update(np.asarray(proof.trace_paths).tobytes())
    h.update(proof.transcript_state)
    return h.digest()


class GpuExecutor:
    """Runs the service's work on the GPU through the C ABI."""

    def __init__(self, offset: int = 3, fri_k: int = 6, num_queries: int = 30, workers: int = 4, device: int = 0):
        from . import _native, backend, pipeline
        _native.ensure_device(device)
        self.B, self.pipeline = backend, pipeline
        self.offset, self.fri_k, self.num_queries = offset, fri_k, num_queries
        self.workers = workers
        self.pool = (ThreadPoolExecutor(max_workers=workers, initializer=_native.ensure_device_this_thread,
                                        initargs=(device,)) if workers > 1 else None)

    def run_tick(self, admitted: list, tasks: list) -> list:
        """Commit `admitted` and prove `tasks` [(job, table)] concurrently;
        the digests of the tasks in order."""
        if self.pool is None:
            if admitted:
                self.commit(admitted)
            return [self.prove_table(js, ti) for js, ti in tasks]
        fc = self.pool.submit(self.commit, admitted) if admitted else None
        fs = [self.pool.submit(self.prove_table, js, ti) for js, ti in tasks]
        out = [f.result() for f in fs]
        if fc is not None:
            fc.result()
        return out

    def close(self) -> None:
        if self.pool is not None:
            self.pool.shutdown()
            self.pool = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def commit(self, jobs: list) -> None:
        """One batched commitment over every table of the given jobs."""
        B = self.B
        tables, owners, inputs = [], [], []
        for js in jobs:
            for ti, (n, m) in enumerate(js.spec.tables):
                tr = B.splitmix_table_dev(n, m, js.spec.cfg, js.spec.job_id * 64 + ti)
                w = B.coset_weights_dev(n, self.offset)
                inputs.append((tr, w))
                tables.append((tr, w, n, m, js.spec.blowup))
                owners.append((js, ti))
        plan = B.BatchPlan(tables, use_graph=False)
        roots = plan.run()
        for i, (js, ti) in enumerate(owners):
            js.handles.append(plan.take(i))
            js.roots.append(roots[i])
        plan.close()
        for tr, w in inputs:
            tr.free()
            w.free()
        # shared transcript over all roots, then one fork per table (prover.rs:3457-3538)
        for js in jobs:
            t = B.Transcript(b"lambda-vm-job-%d" % js.spec.job_id)
            for r in js.roots:
                t.append_bytes(r)
            for ti in range(len(js.spec.tables)):
                f = t.clone()
                if len(js.spec.tables) > 1:
                    f.append_bytes(int(ti).to_bytes(8, "little"))
                js.transcripts.append(f)

    def prove_table(self, js: JobState, ti: int) -> bytes:
        n, m = js.spec.tables[ti]
        h = js.handles[ti]
        proof = self.pipeline.prove_committed_table(h, js.transcripts[ti], n, m, js.spec.blowup, self.offset,
                                                    self.fri_k, self.num_queries)
        h.close()
        js.handles[ti] = None
        return proof_digest(proof)


class ProvingService:
    def __init__(self, hbm_budget: int, executor=None, clock=time.perf_counter, policy: str = "oldest",
                 tables_per_tick: int = 8):
        """policy "oldest": each tick proves up to tables_per_tick tables, oldest
        admitted job first (jobs finish in admission order, which keeps the
        latency of early jobs low); "round_robin": one table of every running
        job per tick (processor sharing)."""
        if policy not in ("oldest", "round_robin"):
            raise ValueError(policy)
        self.policy, self.tables_per_tick = policy, tables_per_tick
        self.budget = hbm_budget
        self.ex = executor if executor is not None else GpuExecutor()
        self.clock = clock
        self.waiting: deque = deque()
        self.running: list = []
        self.done: list = []
        self.in_use = 0
        self.peak_in_use = 0
        self.ticks = 0
        self.batch_sizes: list = []

    def submit(self, spec: JobSpec, arrival: float) -> None:
        hbm = job_hbm_bytes(spec)
        if hbm > self.budget:
            raise ValueError(f"job {spec.job_id} needs {hbm} bytes, more than the whole budget")
        self.waiting.append(JobState(spec, arrival, hbm=hbm))

    def _admit(self, now: float) -> list:
        admitted = []
        while self.waiting and self.in_use + self.waiting[0].hbm <= self.budget:
            js = self.waiting.popleft()
            js.admitted = now
            self.in_use += js.hbm
            self.peak_in_use = max(self.peak_in_use, self.in_use)
            admitted.append(js)
        return admitted

    def tick(self) -> bool:
        """Admit what fits and commit it in one batch while the committed jobs
        advance by up to tables_per_tick tables (or one table of every job);
        returns whether any work was done."""
        self.ticks += 1
        admitted = self._admit(self.clock())
        tasks = []
        if self.policy == "round_robin":
            for js in self.running:
                if js.next_table < len(js.spec.tables):
                    tasks.append((js, js.next_table))
                    js.next_table += 1
        else:
            left = self.tables_per_tick
            for js in self.running:
                while left and js.next_table < len(js.spec.tables):
                    tasks.append((js, js.next_table))
                    js.next_table += 1
                    left -= 1
        if hasattr(self.ex, "run_tick"):
            digests = self.ex.run_tick(admitted, tasks)
        else:
            if admitted:
                self.ex.commit(admitted)
            digests = [self.ex.prove_table(js, ti) for js, ti in tasks]
        if admitted:
            self.batch_sizes.append(len(admitted))
            self.running.extend(admitted)
        now = self.clock()
        for (js, ti), d in zip(tasks, digests):
            js.digests[ti] = d
            js.tables_done += 1
            if js.tables_done == len(js.spec.tables):
                js.finished = now
                self.running.remove(js)
                self.in_use -= js.hbm
                self.done.append(js)
        return bool(admitted) or bool(tasks)

    def run(self, arrivals: list, time_scale: float = 1.0) -> dict:
        """Serve `arrivals` [(t, spec)] in real time (t in seconds after start,
        scaled by time_scale); returns the run's statistics."""
        t0 = self.clock()
        pending = deque(sorted(arrivals, key=lambda a: a[0]))
        total = len(pending)
        while len(self.done) < total:
            now = self.clock() - t0
            while pending and pending[0][0] * time_scale <= now:
                t, spec = pending.popleft()
                self.submit(spec, t0 + t * time_scale)
            if not self.tick():
                if pending:
                    wait = pending[0][0] * time_scale - (self.clock() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.05))
        makespan = max(js.finished for js in self.done) - t0
        return self.stats(makespan)

    def run_async(self, arrivals: list, time_scale: float = 1.0) -> dict:
        """Work-conserving form of run(): no tick barrier. The executor's worker
        pool (`ex.pool`, `ex.workers` threads) receives a commit as soon as no
        other commit is in flight: every job admitted while one runs waits for
        it and then commits with the others admitted meanwhile in one batched
        plan (continuous batching: a lone job commits at once, under load the
        commits grow). A table's remaining rounds go to the pool as soon as a
        worker is free, oldest job's tables first; at most workers - 1 table
        tasks are in flight so that a commit never queues behind them. A job's
        tables start when its commit has finished (they fork its transcript).
        Same admission rule and accounting as run()."""
        pool, workers = getattr(self.ex, "pool", None), getattr(self.ex, "workers", 1)
        if pool is None:
            raise ValueError("run_async needs an executor with a worker pool")
        t0 = self.clock()
        pending = deque(sorted(arrivals, key=lambda a: a[0]))
        total = len(pending)
        ready: deque = deque()
        inflight = {}
        tables_in_flight = 0
        table_slots = max(1, workers - 1)
        to_commit: list = []
        commit_busy = False
        while len(self.done) < total:
            now = self.clock() - t0
            while pending and pending[0][0] * time_scale <= now:
                t, spec = pending.popleft()
                self.submit(spec, t0 + t * time_scale)
            to_commit.extend(self._admit(self.clock()))
            if to_commit and not commit_busy:
                self.ticks += 1
                self.batch_sizes.append(len(to_commit))
                inflight[pool.submit(self.ex.commit, to_commit)] = ("commit", to_commit)
                to_commit, commit_busy = [], True
            while ready and tables_in_flight < table_slots:
                js, ti = ready.popleft()
                inflight[pool.submit(self.ex.prove_table, js, ti)] = ("table", (js, ti))
                tables_in_flight += 1
            wait_s = 0.05
            if pending:
                wait_s = min(wait_s, max(0.0, pending[0][0] * time_scale - (self.clock() - t0)))
            if not inflight:
                if pending and wait_s > 0:
                    time.sleep(wait_s)
                continue
            finished, _ = wait(list(inflight), timeout=wait_s, return_when=FIRST_COMPLETED)
            for f in finished:
                kind, payload = inflight.pop(f)
                res = f.result()
                if kind == "commit":
                    commit_busy = False
                    for js in payload:
                        self.running.append(js)
                        ready.extend((js, ti) for ti in range(len(js.spec.tables)))
                        js.next_table = len(js.spec.tables)
                    continue
                tables_in_flight -= 1
                js, ti = payload
                js.digests[ti] = res
                js.tables_done += 1
                if js.tables_done == len(js.spec.tables):
                    js.finished = self.clock()
                    self.running.remove(js)
                    self.in_use -= js.hbm
                    self.done.append(js)
        makespan = max(js.finished for js in self.done) - t0
        return self.stats(makespan)

    def stats(self, makespan: float) -> dict:
        lat = np.array([js.finished - js.arrival for js in self.done])
        cells = sum(js.spec.lde_cells for js in self.done)
        return {
            "jobs": len(self.done),
            "makespan_s": makespan,
            "jobs_per_s": len(self.done) / makespan,
            "gcells_per_s": cells / makespan / 1e9,
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3),
            "max_ms": float(lat.max() * 1e3),
            "peak_hbm_bytes": self.peak_in_use,
            "budget_bytes": self.budget,
            "ticks": self.ticks,
            "mean_jobs_per_commit_batch": float(np.mean(self.batch_sizes)) if self.batch_sizes else 0.0,
            "lde_cells": cells,
        }
