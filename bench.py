"""Benchmark of the STARK proving hot path on one MI355X.

Default workload (--workload ragged): the batched ragged commitment of one
proof's tables: 14 tables with log2 rows uniform in [12, 22], widths uniform
in [8, 120] columns and blowup 2 or 4 by a coin flip (seeded), traces
uniform seeded Goldilocks values already resident in HBM. A step is the coset
LDE of every table, its row-pair Keccak-256 leaves and Merkle tree, and all
roots to the host. The batched path (one work-list launch per stage step
across all tables, replayed as a captured HIP graph) is the measured value;
the plain per-table path (one coset_lde_dev call per table) is timed on the
same inputs beside it.

--workload single: one table of 2^22 rows x 64 columns, blowup 4, with either
the LDE + Merkle commitment only (--commit-only) or the full single-table
rounds (OOD, DEEP, FRI, queries).

Prints one JSON line: value = LDE cells (rows x columns of the evaluation
domains) per second, roofline of the dominant kernel launch from HIP events
recorded on the stream it runs on, and the oracle (C restatement of the
reference CPU prover) timed on the host cores over a bounded sample. The
default workload's config also carries the PCIe-inclusive rate (every trace
uploaded from page-locked host memory before each batched step,
value_with_h2d) and the p50/p99 latency of a short proving-service run
(60 multi-table jobs at 20 jobs/s, --serve-jobs / --rate / --no-serve).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ragged|single]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(kernel: str, n: int, m: int, blowup: int) -> float:
    """Compulsory HBM bytes per launch of the named kernel: what it must read
    and write once (rows x columns x 8 bytes per u64 element)."""
    N = n * blowup
    if kernel == "keccak_leaves":
        return 8.0 * N * m + 32.0 * (N // 2)  # read the LDE once, write the leaves
    if kernel.startswith("ntt_inv_pass"):
        return 16.0 * n * m  # read + write the n-row table
    if kernel == "ntt_coset_pass0":
        return 8.0 * n * m + 8.0 * N * m  # read the coefficients, write all cosets
    if kernel.startswith("ntt_fwd_pass"):
        return 16.0 * N * m  # read + write the N-row LDE
    if kernel == "merkle_inner":
        return 96.0 * (N // 2 - 1)  # read two children, write the parent
    raise KeyError(kernel)


KERNELS = ["ntt_inv_pass0", "ntt_inv_pass1", "ntt_inv_pass2", "ntt_coset_pass0", "ntt_fwd_pass1",
           "ntt_fwd_pass2", "keccak_leaves", "merkle_inner"]
# stage groups reported beside the dominant kernel (not single kernels)
STAGES = ["bary_denoms", "bary_sum", "deep_denoms", "deep_rows", "fri_fold_leaves", "fri_merkle", "fri_fold"]


def measured_traffic(kernel: str, key: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel calibration),
    when one exists for this exact workload (key); otherwise None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    data = json.load(open(path))
    rec = data.get("workloads", {}).get(key, {}).get(kernel)
    return rec.get("bytes_per_launch") if rec else None


def measured_valu_issue(kernel: str, key: str):
    """VALU issue fraction of `kernel` (SQ_INSTS_VALU / duration / peak issue
    rate) from the committed counter summary for this workload, or None."""
    path = os.path.join(ROOT, "profiles", "valu_issue.json")
    if not os.path.exists(path):
        return None
    rec = json.load(open(path)).get("workloads", {}).get(key, {}).get(kernel)
    return rec.get("valu_issue_frac") if rec else None


def cpu_baseline(log_n: int, m: int, blowup: int, offset: int, commit_only: bool, fri_k: int, queries: int) -> dict:
    import oracle
    from lambda_vm_amd import workload
    from oracle.pipeline import replay_single_table

    oracle.set_threads(0)
    threads = oracle.get_threads()
    sample_log_n = log_n - 2
    n = 1 << sample_log_n
    tr = workload.trace_table(1, 0, n, m)
    oracle.lde_commit(tr[: 1 << 12], blowup, offset, want_lde=False)  # warm the code paths
    t0 = time.perf_counter()
    if commit_only:
        oracle.lde_commit(tr, blowup, offset, want_lde=False)
        what = "LDE+commit"
    else:
        replay_single_table(tr, blowup, offset, fri_k, queries, b"lambda-vm-single-table")
        what = "LDE+commit, OOD, DEEP, FRI and queries"
    dt = time.perf_counter() - t0
    cells = n * blowup * m
    return {
        "value": round(cells / dt / 1e9, 5),
        "unit": "Gcells/s",
        "cores": threads,
        "kind": "port",
        "sample": f"one {what} of 2^{sample_log_n} rows x {m} cols, blowup {blowup} "
                  f"({cells} LDE cells, {dt:.2f} s) by the oracle on {threads} host threads",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log-n", type=int, default=22)
    ap.add_argument("--cols", type=int, default=64)
    ap.add_argument("--blowup", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--commit-only", action="store_true", help="time only the LDE + Merkle commitment")
    ap.add_argument("--fri-k", type=int, default=6, help="log2 terminal degree (terminal 2^(log B + k))")
    ap.add_argument("--queries", type=int, default=30)
    ap.add_argument("--workload", choices=["ragged", "single", "serve", "constraints", "prove", "vm"],
                    default="ragged")
    ap.add_argument("--tables", type=int, default=14, help="ragged: number of tables")
    ap.add_argument("--rate", type=float, default=20.0, help="serve: job arrivals per second (Poisson)")
    ap.add_argument("--jobs", type=int, default=120, help="serve: number of jobs")
    ap.add_argument("--sweep-rates", default="",
                    help="serve: comma-separated arrival rates (jobs/s) of a saturation sweep, e.g. 20,60,120,240")
    ap.add_argument("--serve-jobs", type=int, default=60,
                    help="ragged: jobs of the short serving pass that reports p50/p99 latency")
    ap.add_argument("--no-serve", action="store_true", help="ragged: skip the short serving pass")
    ap.add_argument("--budget-gib", type=float, default=48.0, help="serve: HBM budget for jobs in flight")
    ap.add_argument("--vm-set", choices=["8", "all"], default="8",
                    help="vm: eight 2^20-row tables, or every table of a 2^20-cycle run")
    ap.add_argument("--workers", type=int, default=8, help="serve: host threads proving tables concurrently")
    ap.add_argument("--tables-per-tick", type=int, default=8, help="serve: tables advanced per scheduler tick")
    ap.add_argument("--serve-mode", choices=["tick", "async"], default="async",
                    help="serve: tick barriers, or work-conserving dispatch to the worker pool")
    ap.add_argument("--prove-workers", type=int, default=0,
                    help="prove: host threads proving tables concurrently (0: one per table)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: the hot path is sharded by table on a single device "
              "and no multi-GPU figure is produced (run with --gpus 1)", file=sys.stderr)
        return 2
    runner = {"ragged": run_ragged, "single": run_single, "serve": run_serve,
              "constraints": run_constraints, "prove": run_prove, "vm": run_vm}[args.workload]
    out = runner(args)
    print(json.dumps(out), flush=True)
    return 0


def copy_rate_gbs(backend) -> float:
    """Plain copy of 1 GiB (the calibration kernel for the traffic counters);
    its rate is reported beside the roofline as the achievable HBM rate."""
    cbytes = 1 << 30
    src, dst = backend.DeviceBuffer(cbytes), backend.DeviceBuffer(cbytes)
    backend.dev_copy(dst, src, cbytes)
    t0 = time.perf_counter()
    for _ in range(5):
        backend.dev_copy(dst, src, cbytes)
    rate = 5 * 2 * cbytes / (time.perf_counter() - t0) / 1e9
    src.free()
    dst.free()
    return rate


class PinnedHost:
    """Page-locked host memory from the HIP runtime (hipHostMalloc), as the
    reference stages host traces before upload (crypto/math-cuda/src/device.rs:148-156)."""

    def __init__(self, nbytes: int):
        import ctypes as C
        self._hip = C.CDLL("libamdhip64.so")
        self.ptr = C.c_void_p()
        if self._hip.hipHostMalloc(C.byref(self.ptr), C.c_size_t(nbytes), C.c_uint(0)) != 0:
            raise RuntimeError("hipHostMalloc failed")
        self.nbytes = nbytes

    def array(self, count: int, offset_words: int = 0):
        import ctypes as C

        import numpy as np
        buf = (C.c_uint64 * count).from_address(self.ptr.value + 8 * offset_words)
        return np.ctypeslib.as_array(buf)

    def free(self) -> None:
        if self.ptr:
            self._hip.hipHostFree(self.ptr)
            self.ptr = None


def h2d_inclusive(backend, plan, tables, hosts, steps: int) -> dict:
    """Batched steps that first upload every trace from page-locked host memory
    into its device buffer (the end-to-end rate with the PCIe transfer)."""
    import ctypes as C
    total = sum(h.nbytes for h in hosts)
    pin = PinnedHost(total)
    views, off = [], 0
    for h in hosts:
        v = pin.array(h.size, off)
        v[:] = h.reshape(-1)
        views.append(v)
        off += h.size
    ref = plan.run()

    def upload():
        for (tr, *_), v in zip(tables, views):
            backend.N.call("lvm_memcpy_htod", C.c_void_p(tr.ptr), C.c_void_p(v.ctypes.data), v.nbytes)

    upload()
    backend.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        upload()
    t_h2d = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        upload()
        if plan.run() != ref:
            raise RuntimeError("roots changed after the upload")
    t_all = (time.perf_counter() - t0) / steps
    pin.free()
    return {"h2d_ms_per_step": round(t_h2d * 1e3, 3), "h2d_GBps": round(total / t_h2d / 1e9, 2),
            "ms_per_step_with_h2d": round(t_all * 1e3, 3), "trace_bytes": total}


def short_serving_pass(jobs: int, rate: float, budget_gib: float, workers: int) -> dict:
    """A short run of the proving service (Poisson arrivals of multi-table
    jobs, HBM-budget admission, work-conserving dispatch): p50/p99 latency."""
    from lambda_vm_amd import serve

    budget = int(budget_gib * (1 << 30))
    ex = serve.GpuExecutor(offset=3, fri_k=6, num_queries=30, workers=workers)
    warm = serve.ProvingService(budget, ex)
    warm.run_async([(0.0, serve.JobSpec(10 ** 6, [(1 << 10, 8)] * max(4, workers)))])
    svc = serve.ProvingService(budget, ex)
    st = svc.run_async(serve.synth_jobs(rate, jobs, seed=11))
    ex.close()
    return {"p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
            "jobs_per_s": round(st["jobs_per_s"], 3), "serve_jobs": st["jobs"], "serve_rate_per_s": rate,
            "serve_gcells_per_s": round(st["gcells_per_s"], 4),
            "serve_mean_jobs_per_commit_batch": round(st["mean_jobs_per_commit_batch"], 3)}


def ragged_cpu_baseline(shapes, offset: int) -> dict:
    """The oracle's LDE + commit of every table of the workload at half its
    rows (bounded sample), on all host threads."""
    import oracle
    from lambda_vm_amd import workload

    oracle.set_threads(0)
    threads = oracle.get_threads()
    sample = [(n // 2, m, b) for n, m, b in shapes]
    traces = [workload.trace_table(3, t, n, m) for t, (n, m, b) in enumerate(sample)]
    oracle.lde_commit(traces[0][: 1 << 10], 2, offset, want_lde=False)  # warm the code paths
    t0 = time.perf_counter()
    for tr, (n, m, b) in zip(traces, sample):
        oracle.lde_commit(tr, b, offset, want_lde=False)
    dt = time.perf_counter() - t0
    cells = sum(n * m * b for n, m, b in sample)
    return {
        "value": round(cells / dt / 1e9, 5),
        "unit": "Gcells/s",
        "cores": threads,
        "kind": "port",
        "sample": f"LDE+commit of the {len(sample)} tables at half their rows ({cells} LDE cells, {dt:.2f} s) "
                  f"by the oracle on {threads} host threads",
    }


def run_ragged(args) -> dict:
    from lambda_vm_amd import _native, backend, workload

    _native.ensure_device(0)
    offset = workload.GENERATOR
    shapes = workload.ragged_shapes(cfg=3, count=args.tables)
    log(f"[bench] ragged workload: {len(shapes)} tables {shapes}")
    tables, hosts = [], []
    for t, (n, m, b) in enumerate(shapes):
        tr = workload.trace_table(3, t, n, m)
        tables.append((backend.DeviceBuffer.from_array(tr), backend.coset_weights_dev(n, offset), n, m, b))
        hosts.append(tr)
    cells = sum(n * m * b for n, m, b in shapes)
    copy_gbs = copy_rate_gbs(backend)

    def per_table():
        roots = []
        for tr, w, n, m, b in tables:
            h = backend.coset_lde_dev(tr, n, m, b, w)
            roots.append(h.root)
            h.close()
        return roots

    plan = backend.BatchPlan(tables, use_graph=True)
    log(f"[bench] warmup {args.warmup}")
    ref = per_table()
    roots = None
    for _ in range(max(1, args.warmup)):
        roots = plan.run()
    if roots != ref:
        raise RuntimeError("batched roots differ from the per-table roots")
    backend.sync()
    log(f"[bench] timing {args.steps} batched steps (graph replay)")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        if plan.run() != ref:
            raise RuntimeError("root changed between steps")
    dt = (time.perf_counter() - t0) / args.steps
    value = cells / dt / 1e9

    log(f"[bench] timing {args.steps} per-table steps")
    for _ in range(args.warmup):
        per_table()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        per_table()
    dt_pt = (time.perf_counter() - t0) / args.steps

    log(f"[bench] timing {args.steps} batched steps with the trace upload from pinned host memory")
    pcie = h2d_inclusive(backend, plan, tables, hosts, args.steps)
    del hosts

    # the same launches without the graph, each bracketed by HIP events on
    # the library stream, for the per-launch durations of the roofline
    log(f"[bench] timing {args.steps} batched steps with per-launch events")
    eager = backend.BatchPlan(tables, use_graph=False)
    eager.run()
    backend.sync()
    backend.prof_reset()
    backend.prof_enable(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eager.run()
    dt_eager = (time.perf_counter() - t0) / args.steps
    backend.prof_enable(False)
    per_launch = {}
    for name, nbytes, blocks in eager.launches():
        cnt, tot = backend.prof_query(name)
        if cnt:
            per_launch[name] = (tot / cnt, nbytes, blocks)
    eager.close()
    plan.close()
    dom = max(per_launch, key=lambda k: per_launch[k][0])
    avg_ms, nbytes, _ = per_launch[dom]
    achieved = nbytes / (avg_ms / 1e3) / 1e9
    stage = {"ntt": 0.0, "leaves": 0.0, "merkle": 0.0, "roots": 0.0}
    for name, (ms, _, _) in per_launch.items():
        key = name.split("_")[1]
        stage["ntt" if key == "ntt" else "leaves" if key == "leaves" else "merkle" if key == "merkle" else "roots"] += ms
    roofline = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(dom, "ragged14"),
        "valu_issue_frac": measured_valu_issue(dom, "ragged14"),
        "avg_ms": round(avg_ms, 4),
        "per_launch_ms": {k: round(v[0], 4) for k, v in per_launch.items()},
        "per_launch_blocks": {k: v[2] for k, v in per_launch.items()},
        "stage_ms": {k: round(v, 3) for k, v in stage.items()},
        "copy_kernel_GBps": round(copy_gbs, 1),
    }
    out = {
        "metric": "batched LDE+Keccak-Merkle Gcells/s vs per-table launch; p50/p99 proof latency",
        "value": round(value, 4),
        "unit": "Gcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (Goldilocks field)",
        "data": "synthetic: SplitMix64-seeded traces uniform in [0,p), seeded ragged shapes",
        "config": {
            "workload": (f"batched ragged commitment of {len(shapes)} tables (log2 rows uniform in [12,22], "
                         "8..120 columns, blowup 2 or 4), coset LDE (offset 7) + row-pair Keccak-256 Merkle "
                         "tree per table, traces resident in HBM, one launch per stage step over all tables "
                         "replayed as a HIP graph"),
            "tables": [list(s) for s in shapes],
            "lde_cells": cells,
            "per_table_value": round(cells / dt_pt / 1e9, 4),
            "per_table_ms_per_step": round(dt_pt * 1e3, 3),
            "batched_over_per_table": round(dt_pt / dt, 3),
            "batched_eager_ms_per_step": round(dt_eager * 1e3, 3),
            "roots_sha": __import__("hashlib").sha256(b"".join(ref)).hexdigest(),
            # PCIe-inclusive: every trace uploaded from page-locked host memory, then the batched step
            "value_with_h2d": round(cells / (pcie["ms_per_step_with_h2d"] / 1e3) / 1e9, 4),
            **pcie,
        },
        "roofline": roofline,
    }
    for tr, w, *_ in tables:
        tr.free()
        w.free()
    if not args.no_serve:
        log(f"[bench] serving pass: {args.serve_jobs} jobs at {args.rate} jobs/s")
        out["config"].update(short_serving_pass(args.serve_jobs, args.rate, args.budget_gib, args.workers))
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline")
        out["cpu_baseline"] = ragged_cpu_baseline(shapes, offset)
    return out


def run_serve(args) -> dict:
    """Proving service: Poisson arrivals of multi-table jobs, HBM-budget
    admission, continuous batching; throughput and p50/p99 latency."""
    from lambda_vm_amd import serve

    arrivals = serve.synth_jobs(args.rate, args.jobs, seed=11)
    budget = int(args.budget_gib * (1 << 30))
    ex = serve.GpuExecutor(offset=3, fri_k=args.fri_k, num_queries=args.queries, workers=args.workers)
    # warm-up: one small job through the same path (code objects, pools, every worker's stream)
    warm = serve.ProvingService(budget, ex, tables_per_tick=args.tables_per_tick)
    warm_jobs = [(0.0, serve.JobSpec(10 ** 6, [(1 << 10, 8)] * max(4, args.workers)))]
    warm.run_async(warm_jobs) if args.serve_mode == "async" else warm.run(warm_jobs)
    log(f"[bench] serving {args.jobs} jobs at {args.rate} jobs/s, budget {args.budget_gib} GiB, "
        f"{args.workers} proving threads, {args.serve_mode} dispatch")
    svc = serve.ProvingService(budget, ex, tables_per_tick=args.tables_per_tick)
    st = svc.run_async(arrivals) if args.serve_mode == "async" else svc.run(arrivals)
    sweep = []
    for rate in [float(r) for r in args.sweep_rates.split(",") if r]:
        log(f"[bench] sweep: {args.jobs} jobs at {rate} jobs/s")
        sv = serve.ProvingService(budget, ex, tables_per_tick=args.tables_per_tick)
        sw = sv.run_async(serve.synth_jobs(rate, args.jobs, seed=11)) if args.serve_mode == "async" else \
            sv.run(serve.synth_jobs(rate, args.jobs, seed=11))
        sweep.append({"offered_jobs_per_s": rate, "jobs_per_s": round(sw["jobs_per_s"], 3),
                      "p50_ms": round(sw["p50_ms"], 3), "p99_ms": round(sw["p99_ms"], 3),
                      "gcells_per_s": round(sw["gcells_per_s"], 4),
                      "mean_jobs_per_commit_batch": round(sw["mean_jobs_per_commit_batch"], 3)})
    ex.close()
    out = {
        "metric": "batched LDE+Keccak-Merkle Gcells/s vs per-table launch; p50/p99 proof latency",
        "value": round(st["gcells_per_s"], 4),
        "unit": "Gcells/s",
        "n_gpus": 1,
        "steps": st["jobs"],
        "warmup": 1,
        "ms_per_step": round(st["makespan_s"] * 1e3 / st["jobs"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (Goldilocks field)",
        "data": "synthetic: device SplitMix64 tables (draws mod p), Poisson arrivals",
        "config": {
            "workload": (f"proving service: {args.jobs} jobs, Poisson arrivals at {args.rate}/s, 4-20 tables "
                         "per job, total LDE cells log-uniform 2^20..2^28, blowup 4, HBM budget "
                         f"{args.budget_gib} GiB; per job: batched coset LDE + Keccak-Merkle commit, then per "
                         f"table OOD, DEEP, FRI to a 2^(2+{args.fri_k}) terminal and {args.queries} queries"),
            "p50_ms": round(st["p50_ms"], 3),
            "p99_ms": round(st["p99_ms"], 3),
            "mean_ms": round(st["mean_ms"], 3),
            "jobs_per_s": round(st["jobs_per_s"], 3),
            "peak_hbm_gib": round(st["peak_hbm_bytes"] / (1 << 30), 3),
            "mean_jobs_per_commit_batch": round(st["mean_jobs_per_commit_batch"], 3),
            "lde_cells": st["lde_cells"],
            "workers": args.workers,
            "dispatch": args.serve_mode,
            "tables_per_tick": args.tables_per_tick if args.serve_mode == "tick" else None,
            "saturation_sweep": sweep or None,
        },
    }
    return out


def run_constraints(args) -> dict:
    """Constraint + LogUp stage: 2^21 x 80 FibonacciMultiColumn, 2^20 x 48
    Quadratic, 2^18 x 24 four-bus LogUp table, blowup 4, seeded uniform traces
    resident in HBM. A step: batched LDE+commit of the three main traces,
    LogUp aux build + aux LDE+commit, composition polynomial of each table
    (seeded combination coefficients) and its degree-2 split + commit."""
    import random

    import numpy as np

    from lambda_vm_amd import _native, air as A, backend, workload

    _native.ensure_device(0)
    offset, B = workload.PROOF_COSET_OFFSET, 4
    specs = [("fib", 1 << 21, 80), ("quad", 1 << 20, 48), ("logup", 1 << 18, 24)]
    rng = random.Random(2)

    def r3():
        return (rng.randrange(workload.P), rng.randrange(workload.P), rng.randrange(workload.P))

    airs, traces, weights = [], [], []
    for t, (kind, n, m) in enumerate(specs):
        a = (A.fibonacci_multi_column(m) if kind == "fib" else A.quadratic_multi(m) if kind == "quad"
             else A.logup_air(m, A.four_bus_interactions()))
        airs.append(a)
        traces.append(backend.splitmix_table_dev(n, m, 2, t))
        weights.append(backend.coset_weights_dev(n, offset))
    logup_desc = A.descriptor(airs[2])
    z, al = r3(), r3()
    ap = []
    cur = (1, 0, 0)
    from lambda_vm_amd import field as Fh
    for _ in range(logup_desc.alpha_powers_len):
        ap.append(cur)
        cur = tuple(Fh.f3_mul(list(cur), list(al)))
    ap = np.array(ap, dtype=np.uint64)
    betas = [[r3() for _ in range(a.num_constraints)] for a in airs]
    cells = sum(n * m * B for _, n, m in specs)
    copy_gbs = copy_rate_gbs(backend)

    def step():
        plan = backend.BatchPlan([(traces[i], weights[i], n, m, B) for i, (_, n, m) in enumerate(specs)],
                                 use_graph=False)
        roots = plan.run()
        mains = [plan.take(i) for i in range(3)]
        plan.close()
        n_l = specs[2][1]
        aux, L, _ = backend.logup_aux_dev(traces[2], n_l, specs[2][2], logup_desc, ap, z)
        aux_h = backend.coset_lde_dev(aux, n_l, 3 * logup_desc.num_out_cols, B, weights[2])
        aux.free()
        ninv = pow(n_l, workload.P - 2, workload.P)
        toff = np.array([x * ninv % workload.P for x in L], dtype=np.uint64)
        out = list(roots) + [aux_h.root]
        # the three compositions in shared launches (one compiled kernel over all tables)
        comps = []
        for i, a in enumerate(airs):
            groups, ex = a.groups()
            bnd = [(b.col, b.is_aux, b.step, b.value) for b in a.boundary]
            comps.append(dict(dev=a.device, main=mains[i], aux=aux_h if i == 2 else None, next_step=B,
                              trace_rows=specs[i][1], coset_offset=offset, betas=betas[i], root_group=groups,
                              group_exemptions=ex, boundary=bnd, b_betas=[r3() for _ in bnd],
                              rap=[z, al] if i == 2 else None, alpha=ap if i == 2 else None,
                              table_offset=toff if i == 2 else None))
        Hs = backend.eval_composition_batch(comps)
        for i, H in enumerate(Hs):
            parts = backend.composition_commit_d2(H, specs[i][1] * B, specs[i][1], offset)
            out.append(parts.root)
            H.free()
            parts.close()
        for h in mains:
            h.close()
        aux_h.close()
        return out

    for _ in range(max(1, args.warmup)):
        step()
    backend.sync()
    backend.prof_reset()
    backend.prof_enable(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    backend.sync()
    dt = (time.perf_counter() - t0) / args.steps
    backend.prof_enable(False)
    names = ["logup_terms", "logup_accumulate", "composition_transpose", "composition_compiled", "composition",
             "d2_split", "ntt_lde", "keccak_leaves", "merkle_inner"]
    stage = {}
    for k in names:
        cnt, tot = backend.prof_query(k)
        if cnt:
            stage[k] = round(tot / args.steps, 4)
    # the compiled constraint kernel, or the interpreter where it ran instead
    comp_kernel = "composition_compiled" if "composition_compiled" in stage else "composition"
    cnt, tot = backend.prof_query(comp_kernel)
    # compulsory bytes of the three composition launches: each LDE row read
    # once (main base words + aux Fp3 words) and one Fp3 written per row
    comp_bytes = sum(n * B * (8 * m + 24) for _, n, m in specs) + specs[2][1] * B * 24 * logup_desc.num_out_cols
    achieved = comp_bytes / (tot / 1e3 / args.steps) / 1e9 if cnt else None
    return {
        "metric": "batched LDE+Keccak-Merkle Gcells/s vs per-table launch; p50/p99 proof latency",
        "value": round(cells / dt / 1e9, 4),
        "unit": "Gcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (Goldilocks field)",
        "data": "synthetic: device SplitMix64 traces (draws mod p), seeded challenges",
        "config": {
            "workload": ("constraint + LogUp stage: 2^21x80 FibonacciMultiColumn, 2^20x48 Quadratic, 2^18x24 "
                         "4-bus LogUp, blowup 4: batched main LDE+commit, LogUp aux build + aux LDE+commit, "
                         "the three compositions in shared launches, degree-2 split + parts commit"),
            "lde_cells": cells,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": comp_kernel,
            "achieved": round(achieved, 2) if achieved else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": measured_traffic(comp_kernel, "constraints3"),
            "valu_issue_frac": measured_valu_issue(comp_kernel, "constraints3"),
            "compulsory_bytes": comp_bytes,
            "stage_ms_per_step": stage,
            "copy_kernel_GBps": round(copy_gbs, 1),
        },
    }


# stage names the library's event profiler records in a proof
STAGE_NAMES = ["ntt_lde", "keccak_leaves", "merkle_inner", "logup_terms", "logup_accumulate", "composition_transpose",
               "composition_step_inverse", "composition_compiled", "composition", "d2_split", "bary_denoms", "bary_sum",
               "deep_denoms", "deep_rows", "fri_fold_leaves", "fri_merkle", "fri_fold"]


def vm_all_tables(V) -> list:
    """(name, AIR, rows) of every VM table the GPU prover takes, at the heights
    of a 2^20-cycle run with a 16K-instruction program, 2^10 Keccak
    permutations, 16 scalar multiplications and two touched 2^18-byte pages
    (heights chosen here as a representative mix; HALT's single row is below
    the prover's two-row minimum)."""
    entries = V.decode_entries_from_words([(0x1000, [0x00150513] * ((1 << 14) - 1))])  # addi a0, a0, 1
    init = V.register_init_from_entry_point(0x1000)
    page = 0x1_0000_0000
    return [
        ("CPU", V.cpu_table_air(1 << 20), 1 << 20), ("DECODE", V.decode_table_air(entries), 1 << 14),
        ("CPU32", V.cpu32_table_air(), 1 << 17), ("EQ", V.eq_table_air(), 1 << 18),
        ("BYTEWISE", V.bytewise_table_air(), 1 << 18), ("IS_WH", V.is_wh_table_air(), 1 << 18),
        ("LT", V.lt_table_air(), 1 << 19), ("STORE", V.store_table_air(), 1 << 18),
        ("BRANCH", V.branch_table_air(), 1 << 18), ("MEMW_REGISTER", V.memw_register_table_air(), 1 << 21),
        ("LOAD", V.load_table_air(), 1 << 18), ("MUL", V.mul_table_air(), 1 << 16),
        ("SHIFT", V.shift_table_air(), 1 << 17), ("DVRM", V.dvrm_table_air(), 1 << 14),
        ("MEMW_A", V.memw_aligned_table_air(), 1 << 19), ("MEMW", V.memw_table_air(), 1 << 16),
        ("HINT", V.hint_table_air(), 1 << 10), ("KECCAK", V.keccak_table_air(), 1 << 10),
        ("KECCAK_RND", V.keccak_rnd_table_air(), 1 << 15), ("ECDAS", V.ecdas_table_air(), 1 << 12),
        ("ECSM", V.ecsm_table_air(), 1 << 4), ("COPY", V.copy_table_air(), 1 << 14),
        ("L2G_MEMORY", V.l2g_memory_air(1), 1 << 14), ("L2G_GLOBAL", V.l2g_global_air(1), 1 << 14),
        ("KECCAK_RC", V.keccak_rc_table_air(), 32), ("BITWISE", V.bitwise_table_air(), 1 << 20),
        ("REGISTER", V.register_table_air(init, init, V.cpu_final_pc_timestamp(1 << 20)), 128),
        ("PAGE", V.page_table_air(page), V.PAGE_SIZE), ("PAGE", V.page_table_air(page + V.PAGE_SIZE), V.PAGE_SIZE),
        ("GLOBAL_MEMORY", V.global_memory_table_air(page), V.PAGE_SIZE),
    ]


def run_vm(args) -> dict:
    """Full proof of VM tables in one transcript, blowup 4, their constraints
    and bus interactions as captured from the VM's table definitions. The
    default set is eight tables at 2^20 rows (IS_WH 2^18): EQ (12 columns),
    BYTEWISE (26), IS_WH (4), LT (17), STORE (16), BRANCH (14), MEMW_REGISTER
    (10), LOAD (18); --vm-set all proves every table of a 2^20-cycle run
    (vm_all_tables). Traces are uniform draws of the tables' shapes (the
    prover's work does not depend on whether the constraints hold; valid
    traces are proved and verified in the tests)."""
    from lambda_vm_amd import _native, air as A, backend, prover, vm_tables as V

    _native.ensure_device(0)
    if args.vm_set == "all":
        named = vm_all_tables(V)
        airs, rows = [a for _, a, _ in named], [n for _, _, n in named]
    else:
        named = None
        airs = [V.eq_table_air(), V.bytewise_table_air(), V.is_wh_table_air(), V.lt_table_air(),
                V.store_table_air(), V.branch_table_air(), V.memw_register_table_air(), V.load_table_air()]
        rows = [1 << 20, 1 << 20, 1 << 18, 1 << 20, 1 << 20, 1 << 20, 1 << 20, 1 << 20]
    for a in airs:
        a.bus_descriptor = A.descriptor(a)
    traces = [backend.splitmix_table_dev(n, a.num_main, 2, 40 + t) for t, (a, n) in enumerate(zip(airs, rows))]
    opts = prover.ProofOptions(blowup=4, coset_offset=3, fri_final_log_degree=args.fri_k,
                               num_queries=args.queries, grinding_factor=0)
    tables = [(a, tr, n) for a, tr, n in zip(airs, traces, rows)]
    cells = sum(n * a.num_main * 4 for a, n in zip(airs, rows))
    for _ in range(max(1, args.warmup)):
        prover.prove_air_tables(tables, opts)
    backend.sync()
    backend.prof_reset()
    backend.prof_enable(True)
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        prover.prove_air_tables(tables, opts)
        backend.sync()
        times.append(time.perf_counter() - t0)
    backend.prof_enable(False)
    stage = {}
    for k in STAGE_NAMES:
        cnt, tot = backend.prof_query(k)
        if cnt:
            stage[k] = round(tot / args.steps, 3)
    dt = sum(times) / len(times)
    return {
        "metric": "batched LDE+Keccak-Merkle Gcells/s vs per-table launch; p50/p99 proof latency",
        "value": round(cells / dt / 1e9, 4),
        "unit": "Gcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (Goldilocks field)",
        "data": "synthetic: device SplitMix64 traces of the VM tables' shapes (draws mod p)",
        "config": {
            "workload": (("full proof of every VM table of a 2^20-cycle run in one transcript: "
                          + ", ".join("%s 2^%dx%d" % (nm, n.bit_length() - 1, a.num_main) for nm, a, n in named)
                          if named else
                          "full proof of eight VM tables in one transcript: EQ 2^20x12, BYTEWISE 2^20x26, "
                          "IS_WH 2^18x4, LT 2^20x17, STORE 2^20x16, BRANCH 2^20x14, MEMW_REGISTER 2^20x10, "
                          "LOAD 2^20x18")
                         + " (their constraints and bus interactions), blowup 4, "
                         "FRI terminal 2^(2+%d), %d queries" % (args.fri_k, args.queries)),
            "lde_cells": cells,
            "proof_ms": [round(x * 1e3, 3) for x in times],
            "p50_ms": round(sorted(times)[len(times) // 2] * 1e3, 3),
            "stage_ms_per_proof": stage,
        },
    }


def run_prove(args) -> dict:
    """Full AIR proofs of the constraint-stage tables in one transcript: 2^21 x 80
    FibonacciMultiColumn, 2^20 x 48 Quadratic, 2^18 x 24 four-bus LogUp, blowup
    4, FRI to a 2^(2+k) terminal, args.queries queries. A step is one complete
    proof of the three tables (R1 batched commitment, LogUp aux, composition and
    its parts commitment, OOD, DEEP, FRI, openings) from traces resident in HBM.
    The traces are uniform draws (the prover's work does not depend on whether
    the constraints hold; proofs of valid traces are checked by the verifier
    restatement in the tests)."""
    from lambda_vm_amd import _native, air as A, backend, prover

    _native.ensure_device(0)
    specs = [("fib", 1 << 21, 80), ("quad", 1 << 20, 48), ("logup", 1 << 18, 24)]
    airs = [A.fibonacci_multi_column(80), A.quadratic_multi(48), A.logup_air(24, A.four_bus_interactions())]
    airs[0].boundary = [A.Boundary(c, False, s, (1, 0, 0)) for c in range(80) for s in (0, 1)]
    airs[1].boundary = [A.Boundary(c, False, 0, (1, 0, 0)) for c in range(48)]
    traces = [backend.splitmix_table_dev(n, m, 2, t) for t, (_, n, m) in enumerate(specs)]
    opts = prover.ProofOptions(blowup=4, coset_offset=3, fri_final_log_degree=args.fri_k,
                               num_queries=args.queries, grinding_factor=0)
    tables = [(a, tr, n) for a, tr, (_, n, _) in zip(airs, traces, specs)]
    cells = sum(n * m * 4 for _, n, m in specs)
    workers = args.prove_workers or None
    for _ in range(max(1, args.warmup)):
        prover.prove_air_tables(tables, opts, workers=workers)
    backend.sync()
    backend.prof_reset()
    backend.prof_enable(True)
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        prover.prove_air_tables(tables, opts, workers=workers)
        backend.sync()
        times.append(time.perf_counter() - t0)
    backend.prof_enable(False)
    stage = {}
    for k in STAGE_NAMES:
        cnt, tot = backend.prof_query(k)
        if cnt:
            stage[k] = round(tot / args.steps, 3)
    dt = sum(times) / len(times)
    return {
        "metric": "batched LDE+Keccak-Merkle Gcells/s vs per-table launch; p50/p99 proof latency",
        "value": round(cells / dt / 1e9, 4),
        "unit": "Gcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (Goldilocks field)",
        "data": "synthetic: device SplitMix64 traces (draws mod p); challenges from the transcript",
        "config": {
            "workload": ("full AIR proof of three tables in one transcript: 2^21x80 FibonacciMultiColumn, "
                         "2^20x48 Quadratic, 2^18x24 4-bus LogUp, blowup 4, FRI terminal 2^(2+%d), %d queries"
                         % (args.fri_k, args.queries)),
            "lde_cells": cells,
            "proof_ms": [round(x * 1e3, 3) for x in times],
            "table_workers": workers or len(tables),
            "stage_ms_per_proof": stage,
        },
    }


def run_single(args) -> dict:
    import numpy as np  # noqa: F401

    from lambda_vm_amd import _native, backend, workload

    _native.ensure_device(0)
    n, m, B = 1 << args.log_n, args.cols, args.blowup
    offset = workload.GENERATOR
    N = n * B
    log(f"[bench] generating trace 2^{args.log_n} x {m}")
    tr = workload.trace_table(1, 0, n, m)
    d_tr = backend.DeviceBuffer.from_array(tr)
    d_w = backend.coset_weights_dev(n, offset)
    del tr

    from lambda_vm_amd import pipeline

    def step_once():
        if args.commit_only:
            h = backend.coset_lde_dev(d_tr, n, m, B, d_w)
            root = h.root
            h.close()
            return root
        proof = pipeline.prove_single_table(d_tr, n, m, B, offset, d_w, fri_final_log_degree=args.fri_k,
                                            num_queries=args.queries)
        return proof.root + proof.transcript_state

    copy_gbs = copy_rate_gbs(backend)

    log(f"[bench] warmup {args.warmup}")
    root0 = None
    for _ in range(args.warmup):
        root0 = step_once()
    backend.sync()
    backend.prof_reset()
    backend.prof_enable(True)
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = step_once()
        if root0 is not None and r != root0:
            raise RuntimeError("root changed between steps")
    backend.sync()
    dt = time.perf_counter() - t0
    backend.prof_enable(False)
    ms = dt / args.steps * 1e3
    cells = N * m
    value = cells / (dt / args.steps) / 1e9

    groups = {}
    for k in KERNELS:
        cnt, tot = backend.prof_query(k)
        if cnt:
            groups[k] = (cnt, tot)
    dom = max(groups, key=lambda k: groups[k][1])
    cnt, tot = groups[dom]
    avg_s = tot / max(cnt, 1) / 1e3
    achieved = algorithmic_bytes(dom, n, m, B) / avg_s / 1e9
    roofline = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "avg_ms": round(avg_s * 1e3, 4),
        "per_kernel_ms": {k: round(v[1] / max(v[0], 1), 4) for k, v in groups.items()},
        "copy_kernel_GBps": round(copy_gbs, 1),
    }
    stage_ms = {}
    for k in STAGES:
        cnt, tot = backend.prof_query(k)
        if cnt:
            stage_ms[k] = round(tot / args.steps, 4)
    if stage_ms:
        roofline["other_stage_ms_per_step"] = stage_ms
    traffic = measured_traffic(dom, f"{n}x{m}x{B}")
    if traffic is not None:
        roofline["traffic"] = traffic
    out = {
        "metric": "batched LDE+Keccak-Merkle Gcells/s vs per-table launch; p50/p99 proof latency",
        "value": round(value, 4),
        "unit": "Gcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (Goldilocks field)",
        "data": "synthetic: SplitMix64-seeded trace uniform in [0,p)",
        "config": {
            "workload": (f"single table 2^{args.log_n} rows x {m} cols, blowup {B} coset LDE (offset {offset}) "
                         "+ row-pair Keccak-256 Merkle tree, trace resident in HBM"
                         + ("" if args.commit_only else
                            f"; then OOD (barycentric, 2 points), DEEP, FRI fold-2 to a 2^{B.bit_length() - 1 + args.fri_k}"
                            f" terminal and {args.queries} queries with the Fiat-Shamir transcript")),
            "rows": n, "cols": m, "blowup": B, "lde_cells": cells,
            "root": root0.hex() if root0 else None,
        },
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline")
        out["cpu_baseline"] = cpu_baseline(args.log_n, m, B, offset, args.commit_only, args.fri_k, args.queries)
    return out


if __name__ == "__main__":
    sys.exit(main())
