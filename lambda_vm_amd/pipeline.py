"""Single-table proving rounds on the GPU path, ordered like the reference
prover (crypto/stark/src/prover.rs): R1 commit, out-of-domain point, OOD
values, DEEP coefficients, FRI commit phase, query sampling and openings.

The table here has no AIR constraints (the single-table workload commits,
opens and low-degree-tests the trace columns only), so round 2 is absent:
  R1   coset LDE + row-pair Keccak Merkle tree, root -> transcript
       (crypto/stark/src/prover.rs:3426-3471)
  R3   z <- transcript avoiding the trace domain and the LDE coset
       (crypto/crypto/src/fiat_shamir/is_transcript.rs:44-75); frame points
       z, z*w_n (trace.rs:1112-1132); barycentric OOD values of every column
       at both points appended point-major (prover.rs:4547-4560)
  R4   gamma <- transcript; trace-term coefficients gamma^(j*K + k)
       (prover.rs:2159-2182 takes successive powers of gamma); DEEP codeword;
       FRI commit phase to a 2^(log B + k) terminal (fri/mod.rs:27-151);
       query indexes sample_u64(N/2) (prover.rs:2324-2333); FRI query phase
       and trace openings.
Everything between the rounds stays in HBM; the host sees roots, OOD values,
terminal coefficients and the opened rows.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import backend as B
from . import field as F


@dataclass
class TableProof:
    root: bytes
    z: list
    ood: np.ndarray          # (K, m, 3)
    fri_roots: np.ndarray    # (layers, 32)
    fri_coeffs: np.ndarray   # (2^k, 3)
    iotas: np.ndarray
    fri_sym: np.ndarray      # (layers, queries, 3)
    fri_paths: list
    trace_paths: np.ndarray  # (queries, depth, 32)
    trace_rows: np.ndarray   # (queries, 2, m)
    transcript_state: bytes


def sample_z_ood(t: B.Transcript, n: int, N: int, offset: int):
    g_pow = pow(offset, N, F.P)
    blowup = N // n
    while True:
        z = [int(v) for v in t.sample_fp3()]
        zt = F.f3_pow(z, n)
        if F.f3_is_one(zt):
            continue
        if F.f3_pow(zt, blowup) == [g_pow, 0, 0]:
            continue
        return z


def prove_committed_table(h: B.GpuLde, t: B.Transcript, n: int, m: int, blowup: int, offset: int,
                          fri_final_log_degree: int = 6, num_queries: int = 30, keep: bool = False):
    """Rounds R3/R4 of a committed table whose root the transcript has absorbed,
    in one library call (lvm_prove_committed_table); keep=True runs the same
    steps one call at a time and also returns the FRI state."""
    if not keep:
        return B.prove_committed_table(h, t, n, m, blowup, offset, fri_final_log_degree, num_queries)
    return prove_committed_table_steps(h, t, n, m, blowup, offset, fri_final_log_degree, num_queries, keep=True)


def prove_committed_table_steps(h: B.GpuLde, t: B.Transcript, n: int, m: int, blowup: int, offset: int,
                                fri_final_log_degree: int = 6, num_queries: int = 30, keep: bool = False):
    """Rounds R3/R4 driven from the host one library call at a time."""
    N = n * blowup
    blowup_log = blowup.bit_length() - 1
    # R3
    z = sample_z_ood(t, n, N, offset)
    wn = F.root_of_unity(n.bit_length() - 1)
    points = [z, F.f3_scale(z, wn)]
    K = len(points)
    ood = B.barycentric_base_multi(h, offset, np.array(points, dtype=np.uint64))  # (K, m, 3)
    t.append_fp3_many(ood.reshape(K * m, 3))  # point-major, columns in order
    # R4
    gamma = [int(v) for v in t.sample_fp3()]
    # column j at point k weighted by gamma^j (k = 0) or gamma^(m + j(K-1) + k-1)
    # (crypto/stark/src/ood.rs:57-92, every column opened at the next row)
    gp = B.fp3_powers(gamma, m * K)
    idx = [j if k == 0 else m + j * (K - 1) + k - 1 for j in range(m) for k in range(K)]
    gammas = gp[idx]  # index j*K + k
    trace_ood = np.ascontiguousarray(ood.transpose(1, 0, 2).reshape(m * K, 3))
    deep = B.deep_composition(h, offset, np.array(points, dtype=np.uint64), trace_ood, gammas)
    fri = B.GpuFri(None, offset, on_device=deep, length=N, adopt=True)
    roots, coeffs = fri.commit_phase(t, blowup_log, fri_final_log_degree)
    iotas = np.array([t.sample_u64(N >> 1) for _ in range(num_queries)], dtype=np.uint64)
    sym, paths = fri.query(iotas)
    tpaths, trows = h.open(iotas)
    proof = TableProof(h.root, z, ood, roots, coeffs, iotas, sym, paths, tpaths, trows, t.state())
    if keep:
        return proof, fri
    fri.close()
    return proof


def prove_single_table(d_trace: B.DeviceBuffer, n: int, m: int, blowup: int, offset: int,
                       d_weights: B.DeviceBuffer, fri_final_log_degree: int = 6, num_queries: int = 30,
                       label: bytes = b"lambda-vm-single-table", keep: bool = False):
    t = B.Transcript(label)
    # R1
    h = B.coset_lde_dev(d_trace, n, m, blowup, d_weights)
    t.append_bytes(h.root)
    res = prove_committed_table(h, t, n, m, blowup, offset, fri_final_log_degree, num_queries, keep)
    if keep:
        proof, fri = res
        return proof, h, fri
    h.close()
    return res
