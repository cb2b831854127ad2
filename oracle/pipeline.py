"""ORACLE — TEST INFRASTRUCTURE ONLY.

The single-table proving rounds replayed with the oracle (the C restatement
of the reference CPU prover): R1 commit, out-of-domain point, OOD values,
DEEP coefficients, FRI commit phase and query sampling, in the order of
crypto/stark/src/prover.rs (R1 :3426-3471, z :4494 with the rejection rule
of crypto/crypto/src/fiat_shamir/is_transcript.rs:44-75, OOD appends
:4547-4560, gamma powers :2159-2182, FRI crypto/stark/src/fri/mod.rs:27-151,
queries prover.rs:2324-2333; trace-term weights in the order of
crypto/stark/src/ood.rs:57-92). Used as the parity checker of the GPU rounds
and as the timed CPU baseline.
"""
from __future__ import annotations

import numpy as np

from . import P, Transcript, barycentric_base, deep_composition, fri_commit, lde_commit, root_of_unity


def _f3mul(a, b):
    a0, a1, a2 = (int(x) % P for x in a)
    b0, b1, b2 = (int(x) % P for x in b)
    return [(a0 * b0 + 2 * (a1 * b2 + a2 * b1)) % P, (a0 * b1 + a1 * b0 + 2 * a2 * b2) % P,
            (a0 * b2 + a1 * b1 + a2 * b0) % P]


def _f3pow(a, e):
    r, b = [1, 0, 0], [int(x) % P for x in a]
    while e:
        if e & 1:
            r = _f3mul(r, b)
        b = _f3mul(b, b)
        e >>= 1
    return r


def replay_single_table(tr: np.ndarray, blowup: int, g: int, k: int, queries: int, label: bytes):
    t = Transcript(label)
    lde, nodes = lde_commit(tr, blowup, g)
    t.append_bytes(bytes(nodes[0]))
    return replay_committed_table(lde, nodes, t, tr.shape[0], blowup, g, k, queries)


def replay_committed_table(lde: np.ndarray, nodes: np.ndarray, t: Transcript, n: int, blowup: int, g: int, k: int,
                           queries: int):
    """Rounds R3/R4 of a committed table (natural-order LDE and its tree) whose
    root the transcript `t` has absorbed."""
    m = lde.shape[1]
    N = n * blowup
    while True:
        z = [int(v) for v in t.sample_fp3()]
        zt = _f3pow(z, n)
        if zt == [1, 0, 0] or _f3pow(zt, blowup) == [pow(g, N, P), 0, 0]:
            continue
        break
    wn = root_of_unity(n.bit_length() - 1)
    points = [z, [v * wn % P for v in z]]
    ood = np.array([barycentric_base(lde, n, blowup, g, np.array(p, dtype=np.uint64)) for p in points])
    for kk in range(len(points)):
        for c in range(m):
            t.append_fp3(ood[kk, c])
    gamma = [int(v) for v in t.sample_fp3()]
    pw, cur = [], [1, 0, 0]
    for _ in range(len(points) * m):
        pw.append(cur)
        cur = _f3mul(cur, gamma)
    K = len(points)
    # crypto/stark/src/ood.rs:57-92: column j at row 0 gets gamma^j, at row
    # k >= 1 gamma^(m + j(K-1) + k-1); stored here at index j*K + k
    gam = [pw[j if k == 0 else m + j * (K - 1) + k - 1] for j in range(m) for k in range(K)]
    tood = np.ascontiguousarray(ood.transpose(1, 0, 2).reshape(len(points) * m, 3))
    deep = deep_composition(lde, g, np.array(points, dtype=np.uint64), tood, np.array(gam, dtype=np.uint64))
    bits = N.bit_length() - 1
    roots, coeffs, layers = fri_commit(deep[_bitrev(bits)], t, g, blowup.bit_length() - 1, k, keep_layers=True)
    iotas = [t.sample_u64(N // 2) for _ in range(queries)]
    return dict(root=bytes(nodes[0]), z=z, ood=ood, roots=roots, coeffs=coeffs, iotas=iotas, state=t.state(),
                lde=lde, nodes=nodes, layers=layers)


def _bitrev(bits: int) -> np.ndarray:
    idx = np.arange(1 << bits, dtype=np.uint64)
    out = np.zeros_like(idx)
    for b in range(bits):
        out |= ((idx >> np.uint64(b)) & np.uint64(1)) << np.uint64(bits - 1 - b)
    return out.astype(np.int64)
