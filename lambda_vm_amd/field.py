"""Host-side Goldilocks / Fp3 scalars for orchestration (challenges, powers,
frame points). Plain Python integers; the bulk arithmetic runs on the GPU.

Fp3 = Fp[w]/(w^3 - 2) (crypto/math/src/field/extensions_goldilocks.rs:305-321).
"""
from __future__ import annotations

P = 0xFFFFFFFF00000001
ROOT_2_32 = 1753635133440165772  # crypto/math/src/field/goldilocks.rs:540


def root_of_unity(order: int) -> int:
    """Primitive 2^order-th root (crypto/math/src/field/traits.rs:83-95)."""
    return pow(ROOT_2_32, 1 << (32 - order), P) if order else 1


def f3(v) -> list[int]:
    return [int(x) % P for x in v]


def f3_add(a, b):
    return [(int(x) + int(y)) % P for x, y in zip(a, b)]


def f3_sub(a, b):
    return [(int(x) - int(y)) % P for x, y in zip(a, b)]


def f3_mul(a, b):
    a0, a1, a2 = (int(x) % P for x in a)
    b0, b1, b2 = (int(x) % P for x in b)
    return [(a0 * b0 + 2 * (a1 * b2 + a2 * b1)) % P, (a0 * b1 + a1 * b0 + 2 * a2 * b2) % P,
            (a0 * b2 + a1 * b1 + a2 * b0) % P]


def f3_scale(a, s: int):
    return [int(x) * int(s) % P for x in a]


def f3_pow(a, e: int):
    r, b = [1, 0, 0], f3(a)
    while e:
        if e & 1:
            r = f3_mul(r, b)
        b = f3_mul(b, b)
        e >>= 1
    return r


def f3_is_one(a) -> bool:
    return f3(a) == [1, 0, 0]


def powers(g, count: int):
    out, cur = [], [1, 0, 0]
    for _ in range(count):
        out.append(cur)
        cur = f3_mul(cur, g)
    return out
