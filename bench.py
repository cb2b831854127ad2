"""Benchmark of the STARK proving hot path on one MI355X.

Workload (N=1): one table of 2^22 rows x 64 Goldilocks columns, uniform
seeded values, blowup 4 coset LDE (2^24-row evaluation domain, 8 GiB) and the
row-pair Keccak-256 Merkle tree over it, with the trace already resident in
HBM (row-major, the layout the reference hands to its device path). A step is
one full pass: inverse NTT, coset expansion, leaves, inner tree, root to host.

Prints one JSON line: value = LDE cells (rows x columns of the evaluation
domain) per second, roofline of the dominant kernel from HIP events recorded
on the stream it runs on, and the oracle (C restatement of the reference CPU
prover) timed on the host cores over a bounded sample.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--log-n 22] [--cols 64]
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


def measured_traffic(kernel: str, n: int, m: int, blowup: int):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel calibration),
    when one exists for this exact workload; otherwise None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    data = json.load(open(path))
    key = f"{n}x{m}x{blowup}"
    rec = data.get("workloads", {}).get(key, {}).get(kernel)
    return rec.get("bytes_per_launch") if rec else None


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
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: the hot path is sharded by table on a single device "
              "and no multi-GPU figure is produced (run with --gpus 1)", file=sys.stderr)
        return 2

    import numpy as np

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

    # plain copy of 1 GiB (the calibration kernel for the traffic counters);
    # its rate is reported beside the roofline as the achievable HBM rate
    cbytes = 1 << 30
    src, dst = backend.DeviceBuffer(cbytes), backend.DeviceBuffer(cbytes)
    backend.dev_copy(dst, src, cbytes)
    t0 = time.perf_counter()
    for _ in range(5):
        backend.dev_copy(dst, src, cbytes)
    copy_gbs = 5 * 2 * cbytes / (time.perf_counter() - t0) / 1e9
    src.free()
    dst.free()

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
    traffic = measured_traffic(dom, n, m, B)
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
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
