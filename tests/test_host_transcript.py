"""The library's host transcript (C ABI, no GPU needed) against the oracle
and the reference's transcript known-answer vectors."""
import json
import os

import numpy as np

from lambda_vm_amd import backend

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kats.json")))


def test_basic_challenge_kat():
    k = KATS["transcript_basic_challenge"]
    t = backend.Transcript()
    for a in k["appends_before_challenge1"]:
        t.append_bytes(bytes(a))
    assert list(t.sample()) == k["challenge1"]
    for a in k["appends_before_challenge2"]:
        t.append_bytes(bytes(a))
    assert list(t.sample()) == k["challenge2"]


def test_matches_oracle_on_mixed_sequences(ora):
    rng = np.random.default_rng(3)
    for trial in range(20):
        seed = bytes(rng.integers(0, 256, rng.integers(0, 300), dtype=np.uint8))
        a, b = backend.Transcript(seed), ora.Transcript(seed)
        for step in range(30):
            op = rng.integers(0, 7)
            if op == 0:
                d = bytes(rng.integers(0, 256, rng.integers(0, 200), dtype=np.uint8))
                a.append_bytes(d); b.append_bytes(d)
            elif op == 1:
                v = int(rng.integers(0, 2**63)) * 2 + 1
                a.append_base(v); b.append_base(v)
            elif op == 2:
                v = rng.integers(0, 2**63, 3, dtype=np.uint64)
                a.append_fp3(v); b.append_fp3(v)
            elif op == 3:
                assert a.sample() == b.sample()
            elif op == 4:
                assert a.sample_base() == b.sample_base()
            elif op == 5:
                assert list(a.sample_fp3()) == list(b.sample_fp3())
            else:
                ub = int(rng.integers(1, 2**40))
                assert a.sample_u64(ub) == b.sample_u64(ub)
            assert a.state() == b.state()
        c = a.clone()
        assert c.sample() == a.sample()


def test_append_fp3_many_equals_single_appends():
    rng = np.random.default_rng(5)
    P = 0xFFFFFFFF00000001
    for count in (0, 1, 7, 130):
        vals = rng.integers(0, 2**63, (count, 3), dtype=np.uint64) * np.uint64(2)  # some at or above p
        vals[: min(count, 2)] = np.uint64(P)  # non-canonical limbs are canonicalised
        a, b = backend.Transcript(b"x"), backend.Transcript(b"x")
        a.append_fp3_many(vals)
        for v in vals:
            b.append_fp3(v)
        assert a.sample() == b.sample()


def test_fp3_powers_match_schoolbook():
    from lambda_vm_amd import field as F
    P = 0xFFFFFFFF00000001
    for g in ([0, 0, 0], [1, 0, 0], [5, 7, 11], [P - 1, P - 2, 3]):
        got = backend.fp3_powers(g, 20)
        want = [[int(x) % P for x in v] for v in F.powers(g, 20)]
        assert got.tolist() == want
    # w^3 = 2: (0, 1, 0)^3 = 2
    assert backend.fp3_powers([0, 1, 0], 4)[3].tolist() == [2, 0, 0]
