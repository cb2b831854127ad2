"""Table scheduling of the multi-table prover (lambda_vm_amd.scheduler, after
prover.rs:677-789): the byte estimate, heaviest-first order, the admission
gate (never above budget, an oversized table admitted alone, waiting threads
woken on release) and run_admitted (every index once, results in index
order, start order followed, errors re-raised)."""
import threading
import time

import pytest

from lambda_vm_amd import scheduler as S


def test_estimate_matches_reference_formula():
    # (main*8 + aux*24) * lde * 2 + lde * 256
    assert S.estimate_table_vram_bytes(10, 2, 1 << 10) == ((80 + 48) * 1024 * 2 + 1024 * 256)
    assert S.estimate_table_vram_bytes(0, 0, 7) == 7 * 256


def test_heaviest_first_is_stable_descending():
    assert S.heaviest_first([3, 9, 3, 1, 9]) == [1, 4, 0, 2, 3]
    assert S.heaviest_first([]) == []


def test_gate_budget_and_oversized_request():
    g = S.VramGate(100)
    a = g.acquire(60)
    got = []

    def take(nb):
        p = g.acquire(nb)
        got.append(nb)
        time.sleep(0.01)
        p.release()

    t = threading.Thread(target=take, args=(50,))  # 60 + 50 > 100: waits
    t.start()
    time.sleep(0.05)
    assert got == [] and g.used == 60
    a.release()
    t.join()
    assert got == [50] and g.used == 0
    with g.acquire(500):  # larger than the whole budget: admitted alone
        assert g.used == 500
    assert g.used == 0 and g.peak == 500


def test_run_admitted_respects_budget_and_order():
    est = [40, 70, 10, 70, 30, 20]
    gate = S.VramGate(100)
    order = S.heaviest_first(est)
    started, lock, live = [], threading.Lock(), [0]
    peak = [0]

    def task(i):
        with lock:
            started.append(i)
            live[0] += est[i]
            peak[0] = max(peak[0], live[0])
        time.sleep(0.005)
        with lock:
            live[0] -= est[i]
        return i * i

    res = S.run_admitted(order, est, gate, 3, task)
    assert res == [i * i for i in range(len(est))]
    assert sorted(started) == list(range(len(est)))
    assert peak[0] <= 100 and gate.peak <= 100 and gate.used == 0
    # one worker follows the order exactly
    started.clear()
    S.run_admitted(order, est, S.VramGate(1000), 1, task)
    assert started == order


def test_run_admitted_reraises_task_errors():
    def task(i):
        if i == 2:
            raise ValueError("boom")
        return i

    with pytest.raises(ValueError):
        S.run_admitted([0, 1, 2, 3], [1, 1, 1, 1], S.VramGate(10), 2, task)
