"""Autotuner of the NTT pass splits.

A transform of size 2^log_n runs as a chain of passes of radix 2^r (r <= 8,
one LDS exchange each, csrc/ntt.hip); which chain is fastest depends on the
size (how the last pass's rows fall into tiles, how many post-twiddle
multiplies each split needs) and was picked by hand before. The tuner times
every candidate split of each size on the device, through the plain per-table
coset LDE (inverse transform of the trace + coset expansion, the launches the
batched commitment also runs), at several widths and blowups, keeps the split
with the least summed NTT time per size and writes the table with every
alternative it measured as JSON. `install()` loads that table into the library
(lvm_set_ntt_plan) when the device is initialised, so the batched and the
per-table paths both run the tuned splits.

The reference makes the same kind of measured choice for its GPU dispatch
threshold (crypto/stark/src/gpu_lde.rs:38-55, a sweep on the target GPU).

usage (on the GPU): python -m lambda_vm_amd.tune [--out PATH] [--reps K]
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import time

PKG = os.path.dirname(os.path.abspath(__file__))
ARCH = "gfx950"
PLAN_FILE = os.path.join(PKG, f"ntt_plans_{ARCH}.json")
NTT_EVENTS = ("ntt_inv_pass0", "ntt_inv_pass1", "ntt_inv_pass2", "ntt_inv_pass3", "ntt_coset_pass0",
              "ntt_fwd_pass1", "ntt_fwd_pass2", "ntt_fwd_pass3")


def candidate_splits(log_n: int) -> list[list[int]]:
    """Ordered splits of log_n into passes of at most 8 bits: every split with
    the fewest passes the size allows (parts below 4 bits only where the size
    leaves no other choice), and the balanced splits with one pass more."""
    lo = max(1, (log_n + 7) // 8)
    floor = 4 if log_n >= 4 * lo else 1
    out = [list(p) for p in itertools.product(range(floor, 9), repeat=lo) if sum(p) == log_n]
    k = lo + 1
    if k <= log_n:
        q, r = divmod(log_n, k)
        for p in sorted(set(itertools.permutations([q + 1] * r + [q] * (k - r)))):
            if max(p) <= 8 and list(p) not in out:
                out.append(list(p))
    return out


def _ntt_ms(backend, tr, w, n, m, blowup, reps):
    """NTT time (ms, summed over the passes' HIP events) of one coset LDE."""
    for _ in range(2):
        backend.coset_lde_dev(tr, n, m, blowup, w).close()
    backend.sync()
    backend.prof_reset()
    backend.prof_enable(True)
    for _ in range(reps):
        backend.coset_lde_dev(tr, n, m, blowup, w).close()
    backend.sync()
    backend.prof_enable(False)
    total = 0.0
    for name in NTT_EVENTS:
        cnt, tot = backend.prof_query(name)
        total += tot
    return total / reps


def tune(log_sizes, widths=(16, 64, 120), blowups=(2, 4), reps=5, log=print) -> dict:
    from . import _native, backend, workload

    _native.ensure_device(0, tuned=False)
    table = {"arch": ARCH, "method": ("per size, every candidate split timed by HIP events around the NTT passes "
                                      "of the per-table coset LDE (coset offset 7, SplitMix64 trace), summed over "
                                      "the widths and blowups listed; the least total is installed"),
             "widths": list(widths), "blowups": list(blowups), "reps": reps, "plans": {}}
    for log_n in log_sizes:
        n = 1 << log_n
        cands = candidate_splits(log_n)
        backend.set_ntt_plan(log_n, [])
        rec = {"candidates": {}, "default": backend.get_ntt_plan(log_n)}
        bufs = []
        for m in widths:
            tr = backend.splitmix_table_dev(n, m, 9, log_n)
            bufs.append((m, tr))
        w = backend.coset_weights_dev(n, workload.GENERATOR)
        for split in cands:
            backend.set_ntt_plan(log_n, split)
            per = {}
            for m, tr in bufs:
                for b in blowups:
                    per[f"m{m}_b{b}"] = round(_ntt_ms(backend, tr, w, n, m, b, reps), 4)
            rec["candidates"]["-".join(map(str, split))] = {"ms": per, "total_ms": round(sum(per.values()), 4)}
            log(f"[tune] 2^{log_n} split {split}: {sum(per.values()):.3f} ms")
        backend.set_ntt_plan(log_n, [])
        for _, tr in bufs:
            tr.free()
        w.free()
        best = min(rec["candidates"], key=lambda k: rec["candidates"][k]["total_ms"])
        rec["best"] = [int(x) for x in best.split("-")]
        table["plans"][str(log_n)] = rec
        log(f"[tune] 2^{log_n}: best {rec['best']} (default {rec['default']})")
    return table


def install(lib_call=None, path: str = PLAN_FILE) -> int:
    """Load the tuned splits into the library (one lvm_set_ntt_plan per size);
    returns how many were installed (0 without a table)."""
    if not os.path.exists(path):
        return 0
    import ctypes as C

    from . import _native

    call = lib_call or _native.call
    data = json.load(open(path))
    if data.get("arch") != ARCH:
        return 0
    count = 0
    for k, rec in data.get("plans", {}).items():
        split = rec["best"]
        arr = (C.c_int * len(split))(*split)
        call("lvm_set_ntt_plan", int(k), arr, len(split))
        count += 1
    return count


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=PLAN_FILE)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--min-log", type=int, default=12)
    ap.add_argument("--max-log", type=int, default=22)
    args = ap.parse_args()
    t0 = time.time()
    table = tune(range(args.min_log, args.max_log + 1), reps=args.reps)
    table["tuning_s"] = round(time.time() - t0, 1)
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    print(f"[tune] wrote {args.out}", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
