"""lvm_generate_nonce_noalloc (block minima in a fixed device array, reduced
on the device, no allocation) against lvm_generate_nonce, both called
through the C ABI: the same nonce or the same status on consecutive calls,
the second with a smaller search."""
import ctypes as C
import random

import pytest

pytestmark = pytest.mark.gpu


def _nonce(name, seed, g, max_nonce):
    from lambda_vm_amd import _native as N

    sd = (C.c_uint8 * 32)(*seed)
    out = C.c_uint64(0)
    rc = getattr(N.lib(), name)(sd, C.c_uint8(g), C.c_uint64(max_nonce), C.byref(out))
    return rc, out.value


def _both(seed, g, max_nonce):
    print(f"[no-alloc] seed {seed.hex()[:16]} g={g} max_nonce={max_nonce}", flush=True)
    return _nonce("lvm_generate_nonce", seed, g, max_nonce), _nonce("lvm_generate_nonce_noalloc", seed, g, max_nonce)


def test_noalloc_nonce_equals_allocating_search(gpu, ora):
    rng = random.Random(2024)
    seeds = [bytes(rng.getrandbits(8) for _ in range(32)) for _ in range(3)]
    # a search long enough to run over several batches of 2^24 candidates, then
    # a smaller one: the second call must not see the first call's block minima
    (rc0, want), (rc1, got) = _both(seeds[0], 26, 1 << 40)
    assert rc0 == rc1 == 0
    assert got == want
    assert ora.is_valid_nonce(seeds[0], got, 26)
    (rc0, want), (rc1, got) = _both(seeds[1], 6, 1 << 12)
    assert rc0 == rc1 == 0
    assert got == want == ora.generate_nonce(seeds[1], 6)


def test_noalloc_nonce_status_matches(gpu):
    # nothing below the limit: both report LVM_ERR_STATE; bad grinding factor: LVM_ERR_ARG
    seed = bytes(range(32))
    (rc0, _), (rc1, _) = _both(seed, 40, 1000)
    assert rc0 == rc1 != 0
    (rc0, _), (rc1, _) = _both(seed, 0, 1000)
    assert rc0 == rc1 != 0
    # and a valid search right after the failures
    (rc0, want), (rc1, got) = _both(seed, 4, 1 << 10)
    assert rc0 == rc1 == 0 and got == want
