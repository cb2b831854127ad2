"""Pins the oracle (the C restatement of the reference CPU prover) against
the reference's own known-answer vectors (tests/golden/kats.json) and against
independent mathematics (tests/pyref.py), so that it can serve as the parity
checker for the GPU path."""
import hashlib
import json
import os
import random

import numpy as np
import pytest

import pyref

HERE = os.path.dirname(os.path.abspath(__file__))
KATS = json.load(open(os.path.join(HERE, "golden", "kats.json")))
P = pyref.P


# ---------------- Keccak ----------------
def test_keccak256_reference_kats(ora):
    for v in KATS["keccak256"]:
        assert ora.keccak256(v["input_ascii"].encode()).hex() == v["digest_hex"]


@pytest.mark.parametrize("length", [0, 1, 8, 48, 64, 135, 136, 137, 200, 271, 272, 273, 512, 1024, 1100])
def test_keccak_permutation_vs_hashlib_sha3(ora, length):
    # same permutation + SHA3 domain byte 0x06 must equal hashlib.sha3_256
    data = bytes(random.Random(length).getrandbits(8) for _ in range(length))
    assert ora.keccak_pad(data, 0x06) == hashlib.sha3_256(data).digest()


# ---------------- transcript ----------------
def test_transcript_basic_challenge_kat(ora):
    k = KATS["transcript_basic_challenge"]
    t = ora.Transcript()
    for a in k["appends_before_challenge1"]:
        t.append_bytes(bytes(a))
    assert list(t.sample()) == k["challenge1"]
    for a in k["appends_before_challenge2"]:
        t.append_bytes(bytes(a))
    assert list(t.sample()) == k["challenge2"]


def test_transcript_field_sampling_rules(ora):
    # sample(): digest of everything absorbed, reversed, re-absorbed
    t = ora.Transcript(b"abc")
    d = ora.keccak256(b"abc")
    s = t.sample()
    assert s == d[::-1]
    # finalize_reset restarts the sponge, so the next challenge hashes the
    # reversed digest alone
    assert t.sample() == ora.keccak256(s)[::-1]
    # base sample = first 8 bytes of the squeeze, big-endian, rejected if >= p
    t2 = ora.Transcript(b"xyz")
    sq = ora.keccak256(b"xyz")[::-1]
    v = t2.sample_base()
    cands = [int.from_bytes(sq[i:i + 8], "big") for i in range(0, 32, 8)]
    assert v == next(c for c in cands if c < P)
    # distinct field elements give distinct challenges; equal ones equal
    a = ora.Transcript(); a.append_fp3([1, 0, 0])
    b = ora.Transcript(); b.append_fp3([0, 0, 1])
    c = ora.Transcript(); c.append_fp3([1, 0, 0])
    sa, sb, sc = a.sample_fp3(), b.sample_fp3(), c.sample_fp3()
    assert list(sa) != list(sb) and list(sa) == list(sc)


def test_transcript_sample_u64_rejection(ora):
    t = ora.Transcript(b"q")
    ub = 3
    vals = [t.sample_u64(ub) for _ in range(50)]
    assert all(0 <= v < ub for v in vals)
    # state() does not advance the sponge
    t3 = ora.Transcript(b"q")
    s1 = t3.state()
    assert t3.state() == s1 == ora.keccak256(b"q")


# ---------------- grinding ----------------
def test_grinding_reference_kats(ora):
    for v in KATS["grinding"]:
        assert ora.is_valid_nonce(bytes(v["seed"]), v["nonce"], v["grinding_factor"]) == v["valid"], v["name"]


def test_grinding_smallest_nonce(ora):
    seed = bytes(range(32))
    n = ora.generate_nonce(seed, 8)
    assert ora.is_valid_nonce(seed, n, 8)
    assert not any(ora.is_valid_nonce(seed, k, 8) for k in range(n))


# ---------------- field ----------------
def _edge_u64(rng, k):
    edges = [0, 1, 2, P - 1, P, P + 1, 2**64 - 1, 2**64 - 2, 0xFFFFFFFF, 0x100000000, P - 2**32]
    vals = [rng.choice(edges) if rng.random() < 0.3 else rng.getrandbits(64) for _ in range(k)]
    return np.array(vals, dtype=np.uint64)


def test_goldilocks_ops_vs_bigint(ora):
    rng = random.Random(1)
    a, b = _edge_u64(rng, 4000), _edge_u64(rng, 4000)
    for op, f in [("add", lambda x, y: x + y), ("sub", lambda x, y: x - y), ("mul", lambda x, y: x * y)]:
        got = ora.gl_op(op, a, b)
        for x, y, g in zip(a.tolist(), b.tolist(), got.tolist()):
            assert g % P == f(x, y) % P
    neg = ora.gl_op("neg", a)
    assert all((int(x) + int(y)) % P == 0 for x, y in zip(a, neg))
    nz = np.array([x for x in a.tolist() if x % P], dtype=np.uint64)
    inv = ora.gl_op("inv", nz)
    assert all(int(x) * int(y) % P == 1 for x, y in zip(nz, inv))


def test_fp3_ops_vs_schoolbook(ora):
    rng = random.Random(2)
    a = _edge_u64(rng, 900).reshape(-1, 3)
    b = _edge_u64(rng, 900).reshape(-1, 3)
    mul = ora.fp3_op("mul", a, b)
    sqr = ora.fp3_op("sqr", a)
    for x, y, m, s in zip(a.tolist(), b.tolist(), mul.tolist(), sqr.tolist()):
        assert [v % P for v in m] == pyref.f3mul(x, y)
        assert [v % P for v in s] == pyref.f3mul(x, x)
    nz = np.array([x for x in a.tolist() if any(v % P for v in x)], dtype=np.uint64)
    inv = ora.fp3_op("inv", nz)
    for x, y in zip(nz.tolist(), inv.tolist()):
        assert pyref.f3mul(x, y) == [1, 0, 0]


def test_roots_of_unity(ora):
    g = KATS["goldilocks"]
    assert g["two_adicity"] == 32
    r = g["root_of_unity_2_32"]
    assert pow(r, 1 << 32, P) == 1 and pow(r, 1 << 31, P) == P - 1  # order exactly 2^32
    for k in range(0, 33):
        assert ora.root_of_unity(k) == pow(r, 1 << (32 - k), P)


# ---------------- FFT / LDE ----------------
def test_fft_closed_form_x_plus_one(ora):
    # p(x) = 1 + x at every 256th root of unity (math-cuda/tests/ntt_known.rs:17-40)
    n = 256
    c = np.zeros((n, 1), dtype=np.uint64)
    c[0, 0] = c[1, 0] = 1
    got = ora.fft_two_half(c)[:, 0]
    w = ora.root_of_unity(8)
    assert [int(v) % P for v in got] == [(1 + pow(w, i, P)) % P for i in range(n)]


@pytest.mark.parametrize("log_n,m", [(1, 1), (2, 3), (5, 2), (8, 1), (9, 3)])
def test_fft_vs_quadratic_dft(ora, log_n, m):
    n = 1 << log_n
    rng = np.random.default_rng(log_n)
    x = rng.integers(0, 2**63, size=(n, m), dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    fwd = ora.fft_two_half(x)
    inv = ora.fft_two_half(x, inverse=True)
    w = pow(KATS["goldilocks"]["root_of_unity_2_32"], 1 << (32 - log_n), P)
    wi = pow(w, P - 2, P)
    for c in range(m):
        col = [int(v) for v in x[:, c]]
        assert [int(v) % P for v in fwd[:, c]] == pyref.dft(col, w)
        assert [int(v) % P for v in inv[:, c]] == pyref.dft(col, wi)


@pytest.mark.parametrize("log_n,m,blowup,offset", [(2, 2, 2, 7), (4, 3, 4, 3), (6, 1, 8, 7), (7, 2, 1, 5)])
def test_lde_is_polynomial_evaluation_on_coset(ora, log_n, m, blowup, offset):
    n = 1 << log_n
    N = n * blowup
    rng = np.random.default_rng(100 + log_n)
    tr = rng.integers(0, P, size=(n, m), dtype=np.uint64)
    lde = ora.coset_lde_row_major(tr, blowup, offset)
    r = KATS["goldilocks"]["root_of_unity_2_32"]
    wn = pow(r, 1 << (32 - log_n), P)
    wN = pow(r, 1 << (32 - (N.bit_length() - 1)), P)
    ninv = pow(n, P - 2, P)
    for c in range(m):
        col = [int(v) for v in tr[:, c]]
        coeffs = [v * ninv % P for v in pyref.dft(col, pow(wn, P - 2, P))]  # interpolant on <w_n>
        want = [pyref.horner(coeffs, offset * pow(wN, i, P) % P) for i in range(N)]
        assert [int(v) % P for v in lde[:, c]] == want


# ---------------- Merkle ----------------
def test_commit_leaf_layout_and_tree(ora):
    # leaf i = Keccak(BE(row[br(2i)]) || BE(row[br(2i+1)])), parents Keccak(l || r)
    rows, cols = 16, 3
    rng = np.random.default_rng(5)
    data = rng.integers(0, 2**64, size=(rows, cols), dtype=np.uint64)
    nodes = ora.commit_rows_bit_reversed(data)
    L = rows // 2
    for i in range(L):
        msg = b"".join((int(v) % P).to_bytes(8, "big") for k in (0, 1) for v in data[pyref.bitrev(2 * i + k, 4)])
        assert bytes(nodes[L - 1 + i]) == ora.keccak256(msg)
    for p in range(L - 1):
        assert bytes(nodes[p]) == ora.keccak256(bytes(nodes[2 * p + 1]) + bytes(nodes[2 * p + 2]))


def test_commit_ext3_and_column_subset(ora):
    rows, cols = 8, 4
    rng = np.random.default_rng(6)
    data = rng.integers(0, 2**64, size=(rows, 3 * cols), dtype=np.uint64)
    nodes = ora.commit_rows_bit_reversed(data, elem_words=3, c0=1, c1=3)
    for i in range(rows // 2):
        msg = b"".join((int(v) % P).to_bytes(8, "big")
                       for k in (0, 1) for v in data[pyref.bitrev(2 * i + k, 3), 3:9])
        assert bytes(nodes[rows // 2 - 1 + i]) == ora.keccak256(msg)


def test_merkle_padding_paths(ora):
    leaves = np.frombuffer(b"".join(ora.keccak256(bytes([i])) for i in range(5)), dtype=np.uint8).reshape(5, 32)
    nodes = ora.merkle_build(leaves)
    assert nodes.shape[0] == 15  # padded to 8 leaves by repeating the last
    assert all(bytes(nodes[7 + k]) == bytes(leaves[4]) for k in range(4, 8))
    for pos in range(5):
        path = ora.merkle_path(nodes, pos)
        assert len(path) == 3
        assert ora.merkle_verify(bytes(nodes[0]), bytes(leaves[pos]), pos, path)
        assert not ora.merkle_verify(bytes(nodes[0]), bytes(leaves[(pos + 1) % 5]), pos, path)


# ---------------- FRI ----------------
def _bitrev_codeword(coeffs3, offset, N, r):
    wN = pow(r, 1 << (32 - (N.bit_length() - 1)), P)
    nat = [pyref.horner3(coeffs3, [offset * pow(wN, i, P) % P, 0, 0]) for i in range(N)]
    bits = N.bit_length() - 1
    return [nat[pyref.bitrev(i, bits)] for i in range(N)]


def test_fri_fold_is_even_odd_combination(ora):
    # folding p = pe(x^2) + x po(x^2) gives 2(pe + zeta po) on the squared coset
    r = KATS["goldilocks"]["root_of_unity_2_32"]
    N, deg, g = 32, 8, 7
    rng = random.Random(9)
    coeffs = [[rng.randrange(P) for _ in range(3)] for _ in range(deg)]
    cw = np.array(_bitrev_codeword(coeffs, g, N, r), dtype=np.uint64)
    zeta = [rng.randrange(P) for _ in range(3)]
    tw = ora.fri_inv_twiddles(g, N)
    folded = ora.fri_fold(cw, np.array(zeta, dtype=np.uint64), tw)
    pe, po = coeffs[0::2], coeffs[1::2]
    comb = [pyref.f3scale(pyref.f3add(a, pyref.f3mul(zeta, b)), 2) for a, b in zip(pe, po)]
    want = _bitrev_codeword(comb, g * g % P, N // 2, r)
    assert [[int(v) % P for v in row] for row in folded] == want


def test_fri_layout_reference_values(ora):
    # terminal.rs:45-55 at the single-table workload: N = 2^24, blowup 4, k = 6
    assert ora.fri_layout(24, 2, 6) == (16, 15, 256, 6)
    assert ora.fri_layout(24, 2, 7) == (15, 14, 512, 7)
    assert ora.fri_layout(3, 2, 7) == (0, 0, 8, 1)  # clamped for tiny inputs
    assert ora.fri_layout(4, 2, 1) == (1, 0, 8, 1)


def test_fri_commit_terminal_recovers_folded_polynomial(ora):
    r = KATS["goldilocks"]["root_of_unity_2_32"]
    N, deg, g, blowup_log, k = 64, 16, 3, 2, 2
    rng = random.Random(10)
    coeffs = [[rng.randrange(P) for _ in range(3)] for _ in range(deg)]
    cw = np.array(_bitrev_codeword(coeffs, g, N, r), dtype=np.uint64)
    t = ora.Transcript(b"fri")
    roots, final, layers = ora.fri_commit(cw, t, g, blowup_log, k, keep_layers=True)
    # replay: same challenges from an identical transcript, fold the coefficients
    t2 = ora.Transcript(b"fri")
    cur = coeffs
    tf, nc, tl, ek = ora.fri_layout(6, blowup_log, k)
    for layer in range(tf):
        z = [int(v) for v in t2.sample_fp3()]
        cur = [pyref.f3scale(pyref.f3add(a, pyref.f3mul(z, b)), 2) for a, b in zip(cur[0::2], cur[1::2])]
        if layer < nc:
            t2.append_bytes(bytes(roots[layer]))
    want = (cur + [[0, 0, 0]] * (1 << ek))[: 1 << ek]
    assert [[int(v) % P for v in row] for row in final] == want
    # each committed root is the pair-leaf tree of that layer
    off = 0
    for layer in range(nc):
        size = N >> (layer + 1)
        ev = layers[off:off + size]
        off += size
        assert bytes(ora.merkle_build(ora.fri_layer_leaves(ev))[0]) == bytes(roots[layer])


# ---------------- barycentric / batch inverse ----------------
def test_barycentric_equals_horner(ora):
    r = KATS["goldilocks"]["root_of_unity_2_32"]
    n, m, B, g = 16, 3, 4, 7
    rng = np.random.default_rng(11)
    tr = rng.integers(0, P, size=(n, m), dtype=np.uint64)
    lde = ora.coset_lde_row_major(tr, B, g)
    z = [123456789, 987654321, 55555]
    got = ora.barycentric_base(lde, n, B, g, np.array(z, dtype=np.uint64))
    wn = pow(r, 1 << (32 - 4), P)
    ninv = pow(n, P - 2, P)
    for c in range(m):
        coeffs = [v * ninv % P for v in pyref.dft([int(v) for v in tr[:, c]], pow(wn, P - 2, P))]
        want = pyref.horner3([[cf, 0, 0] for cf in coeffs], z)
        assert [int(v) % P for v in got[c]] == want


def test_deep_composition_quotient_identity(ora):
    # deep(x) = sum_k [sum_j g_jk (t_j(x) - t_j(z_k))] / (x - z_k) + sum_j gh_j (H_j(x) - H_j(zp)) / (x - zp):
    # checked at every point with big-integer arithmetic
    n, m, B, g = 8, 2, 4, 7
    N = n * B
    rng = np.random.default_rng(13)
    tr = rng.integers(0, P, size=(n, m), dtype=np.uint64)
    lde = ora.coset_lde_row_major(tr, B, g)
    parts = rng.integers(0, P, size=(N, 3), dtype=np.uint64)  # one Fp3 part, arbitrary values
    K = 2
    zp = rng.integers(0, P, size=(K, 3), dtype=np.uint64)
    tood = rng.integers(0, P, size=(m * K, 3), dtype=np.uint64)
    gtr = rng.integers(0, P, size=(m * K, 3), dtype=np.uint64)
    zpa = rng.integers(0, P, size=3, dtype=np.uint64)
    pood = rng.integers(0, P, size=(1, 3), dtype=np.uint64)
    gh = rng.integers(0, P, size=(1, 3), dtype=np.uint64)
    out = ora.deep_composition(lde, g, zp, tood, gtr, parts=parts, z_parts=zpa, parts_ood=pood, gammas_h=gh)
    r = KATS["goldilocks"]["root_of_unity_2_32"]
    wN = pow(r, 1 << (32 - 5), P)
    for i in range(N):
        x = [g * pow(wN, i, P) % P, 0, 0]
        want = pyref.f3mul(pyref.f3mul(gh[0].tolist(), pyref.f3sub(parts[i].tolist(), pood[0].tolist())),
                           pyref.f3inv(pyref.f3sub(x, zpa.tolist())))
        for k in range(K):
            acc = [0, 0, 0]
            for j in range(m):
                acc = pyref.f3add(acc, pyref.f3mul(gtr[j * K + k].tolist(),
                                                   pyref.f3sub([int(lde[i, j]), 0, 0], tood[j * K + k].tolist())))
            want = pyref.f3add(want, pyref.f3mul(acc, pyref.f3inv(pyref.f3sub(x, zp[k].tolist()))))
        assert [int(v) % P for v in out[i]] == want


def test_batch_inverse(ora):
    rng = np.random.default_rng(12)
    for n in (1, 2, 7, 1000, 70000):
        v = rng.integers(1, P, size=(n, 3), dtype=np.uint64)
        inv = ora.batch_inverse_fp3(v)
        idx = range(n) if n < 2000 else rng.integers(0, n, 200)
        for i in idx:
            assert pyref.f3mul(v[i].tolist(), inv[i].tolist()) == [1, 0, 0]
