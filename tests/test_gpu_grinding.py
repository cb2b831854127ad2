"""GPU proof-of-work nonce search (lvm_generate_nonce) against the oracle's
sequential search (smallest valid nonce) and the reference's grinding
vectors (crypto/stark/src/tests/grinding_tests.rs)."""
import json
import os
import random

import pytest

pytestmark = pytest.mark.gpu
KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kats.json")))


@pytest.mark.parametrize("g", [1, 3, 8, 12, 16])
def test_smallest_nonce_matches_oracle(gpu, ora, g):
    rng = random.Random(g)
    for _ in range(3):
        seed = bytes(rng.getrandbits(8) for _ in range(32))
        want = ora.generate_nonce(seed, g)
        got = gpu.generate_nonce(seed, g)
        assert got == want, (seed.hex(), g)
        assert ora.is_valid_nonce(seed, got, g)


def test_reference_grinding_vectors(gpu, ora):
    for case in KATS["grinding"]:
        if not case["valid"]:
            continue
        seed = bytes(case["seed"])
        got = gpu.generate_nonce(seed, case["grinding_factor"])
        # the vector's nonce is valid, so the smallest one is at most it
        assert got <= case["nonce"]
        assert ora.is_valid_nonce(seed, got, case["grinding_factor"])
        if case["nonce"] < (1 << 22):  # the sequential CPU search finishes in seconds
            assert got == ora.generate_nonce(seed, case["grinding_factor"])


def test_search_limit_and_argument_errors(gpu):
    seed = bytes(32)
    with pytest.raises(gpu.LvmError):
        gpu.generate_nonce(seed, 40, max_nonce=1000)  # 2^-40 odds: nothing below 1000
    with pytest.raises(gpu.LvmError):
        gpu.generate_nonce(seed, 0)
