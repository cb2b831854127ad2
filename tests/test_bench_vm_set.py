"""The all-tables VM workload of bench.py builds on the CPU: every VM table
but the one-row HALT, each at a power-of-two height of at least two rows (the
GPU prover's minimum), with its bus descriptor."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from lambda_vm_amd import air as A  # noqa: E402
from lambda_vm_amd import vm_tables as V  # noqa: E402

TABLES = {"CPU", "DECODE", "CPU32", "EQ", "BYTEWISE", "IS_WH", "LT", "STORE", "BRANCH", "MEMW_REGISTER", "LOAD", "MUL",
          "SHIFT", "DVRM", "MEMW_A", "MEMW", "HINT", "KECCAK", "KECCAK_RND", "ECDAS", "ECSM", "COPY", "L2G_MEMORY",
          "L2G_GLOBAL", "KECCAK_RC", "BITWISE", "REGISTER", "PAGE", "GLOBAL_MEMORY"}


def test_vm_all_tables_builds():
    named = bench.vm_all_tables(V)
    assert {nm for nm, _, _ in named} == TABLES and len(named) == 30
    for nm, air, n in named:
        assert n >= 2 and n & (n - 1) == 0, nm
        if air.num_precomputed:
            pre = air.precomputed() if callable(air.precomputed) else air.precomputed
            assert pre.shape == (n, air.num_precomputed), nm
        air.bus_descriptor = A.descriptor(air)
        assert air.interactions, nm
