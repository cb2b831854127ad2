"""Per-proof table scheduling of the multi-table prover, after the reference's
(crypto/stark/src/prover.rs:677-789): an estimate of each table's peak device
bytes, a byte-budget admission gate, and driver threads that take tables
heaviest first, each admitted through the gate before it runs.

  estimate_table_vram_bytes  prover.rs:682-694
  VramGate / VramPermit      prover.rs:696-738 (an oversized request is
                             admitted alone, when nothing else holds bytes)
  run_admitted               prover.rs:740-779 (a shared cursor over `order`)
  heaviest_first             prover.rs:782-789 (stable: equal estimates keep
                             their index order)
"""
from __future__ import annotations

import threading


def estimate_table_vram_bytes(main_cols: int, aux_cols: int, lde_size: int) -> int:
    """Co-resident LDE columns (8 bytes per base column, 24 per Fp3 aux
    column) times a scratch factor of 2, plus 256 bytes per LDE row of
    resident Merkle trees; a deliberate over-estimate."""
    per_row = main_cols * 8 + aux_cols * 24
    return lde_size * per_row * 2 + lde_size * 256


class VramPermit:
    def __init__(self, gate: "VramGate", nbytes: int):
        self.gate, self.bytes = gate, nbytes

    def release(self) -> None:
        if self.gate is not None:
            self.gate._release(self.bytes)
            self.gate = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.release()
        return False


class VramGate:
    """Blocks acquire() until the bytes fit under the budget; a request larger
    than the whole budget is admitted when nothing else holds bytes."""

    def __init__(self, budget: int):
        self.budget = budget
        self.used = 0
        self.peak = 0
        self._cv = threading.Condition()

    def acquire(self, nbytes: int) -> VramPermit:
        with self._cv:
            while not (self.used == 0 or self.used + nbytes <= self.budget):
                self._cv.wait()
            self.used += nbytes
            self.peak = max(self.peak, self.used)
        return VramPermit(self, nbytes)

    def _release(self, nbytes: int) -> None:
        with self._cv:
            self.used = max(0, self.used - nbytes)
            self._cv.notify_all()


def heaviest_first(estimates) -> list:
    """Table indices sorted by estimate, largest first."""
    return sorted(range(len(estimates)), key=lambda i: -estimates[i])


def run_admitted(order, estimates, gate: VramGate, workers: int, task, initializer=None) -> list:
    """Run task(idx) once per index on `workers` threads, taking indices in
    `order` from a shared cursor and admitting each through `gate` with its
    estimate. Returns the results in index order; the first exception raised
    by a task is re-raised after every thread has stopped."""
    results = [None] * len(estimates)
    errors = []
    lock = threading.Lock()
    cursor = [0]

    def worker():
        if initializer is not None:
            initializer()
        while True:
            with lock:
                if errors or cursor[0] >= len(order):
                    return
                idx = order[cursor[0]]
                cursor[0] += 1
            try:
                with gate.acquire(estimates[idx]):
                    results[idx] = task(idx)
            except BaseException as e:  # noqa: BLE001 - handed to the caller
                with lock:
                    errors.append(e)
                return

    n = max(1, min(workers, len(order)))
    if n == 1:
        worker()
    else:
        threads = [threading.Thread(target=worker) for _ in range(n)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
    if errors:
        raise errors[0]
    return results
