"""Seeded synthetic inputs for the hot-path workloads.

Values come from SplitMix64 (the generator the reference's own GPU tests
use, crypto/stark/src/gpu_lde.rs:3409-3418), one stream per
(workload, table, column), rejection-sampled into [0, p). The vectorised
form below reproduces the sequential stream exactly: draw k is
mix(seed + (k+1)*0x9E3779B97F4A7C15), and candidates >= p are skipped.
"""
from __future__ import annotations

import numpy as np

P = 0xFFFFFFFF00000001
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Draws start..start+count-1 of the SplitMix64 stream seeded with `seed`."""
    with np.errstate(over="ignore"):
        k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def field_stream(seed: int, count: int) -> np.ndarray:
    """First `count` accepted (< p) values of the stream."""
    out = np.empty(count, dtype=np.uint64)
    got, start = 0, 0
    while got < count:
        need = count - got
        draw = splitmix64(seed, need + 64, start)
        start += need + 64
        ok = draw[draw < np.uint64(P)]
        take = min(ok.size, need)
        out[got:got + take] = ok[:take]
        got += take
    return out


def column_seed(cfg: int, table: int, col: int) -> int:
    return (0x1A3BDA7A ^ (cfg << 48) ^ (table << 32) ^ col) & 0xFFFFFFFFFFFFFFFF


def trace_table(cfg: int, table: int, n: int, m: int) -> np.ndarray:
    """Row-major n x m table, column c drawn from its own stream."""
    t = np.empty((n, m), dtype=np.uint64)
    for c in range(m):
        t[:, c] = field_stream(column_seed(cfg, table, c), n)
    return t


def trace_table_modp(cfg: int, table: int, n: int, m: int) -> np.ndarray:
    """Row-major n x m table, column c the first n draws of its stream reduced
    mod p (no rejection): the device generator's definition."""
    t = np.empty((n, m), dtype=np.uint64)
    for c in range(m):
        d = splitmix64(column_seed(cfg, table, c), n)
        t[:, c] = np.where(d >= np.uint64(P), d - np.uint64(P), d)
    return t


def fp3_values(seed: int, count: int) -> np.ndarray:
    return field_stream(seed, 3 * count).reshape(count, 3)


# Shapes of the hot-path workloads (blowup B, coset offset g)
GENERATOR = 7  # multiplicative generator of the Goldilocks group
PROOF_COSET_OFFSET = 3  # coset_offset of every production ProofOptions (crypto/stark/src/proof/options.rs:70)

WORKLOADS = {
    "plain_commit_2^16x32": dict(cfg=0, n=1 << 16, m=32, blowup=4, offset=GENERATOR),
    "single_table_2^22x64": dict(cfg=1, n=1 << 22, m=64, blowup=4, offset=GENERATOR, fri_final_log=8,
                                 ood_points=2),
}


def ragged_shapes(cfg: int = 3, count: int = 14, log_lo: int = 12, log_hi: int = 22, m_lo: int = 8,
                  m_hi: int = 120, blowups=(2, 4)):
    """Seeded shapes (n, m, blowup) of a batched ragged proof: log2 n uniform
    in [log_lo, log_hi], m uniform in [m_lo, m_hi], blowup by a coin flip.
    Three draws per table from the stream of column_seed(cfg, 0xFFFF, 0)."""
    d = splitmix64(column_seed(cfg, 0xFFFF, 0), 3 * count)
    out = []
    for t in range(count):
        log_n = log_lo + int(d[3 * t]) % (log_hi - log_lo + 1)
        m = m_lo + int(d[3 * t + 1]) % (m_hi - m_lo + 1)
        b = blowups[int(d[3 * t + 2]) & 1]
        out.append((1 << log_n, m, b))
    return out
