"""Pins of the oracle's constraint walker and composition polynomial (CPU).

- The walk over the lowered flat program equals a big-integer evaluation of
  the captured (unlowered) program and, for the base AIRs, the constraint
  formulas written out directly: pins the lowering (slot reuse, uniform
  operands, mixed base/extension operations) and the walker.
- Composition of a trace that satisfies its AIR is a low-degree polynomial on
  the LDE coset (degree <= max_c(d_c (n-1) + e_c - n), boundary terms n - 2):
  a wrong zerofier (exemption count or root), frame row, sign or boundary
  term leaves high coefficients, which the negative controls show.
- Composition values at sampled rows equal the closed form
  sum_c beta_c C_c(x) prod_j (x - w^(n-j)) / (x^n - 1) + sum_b beta_b (t_b(x) - v_b)/(x - w^s_b).
"""
import random

import numpy as np
import pytest

from lambda_vm_amd import air as A
from lambda_vm_amd import workload
from pyref import P, f3add as _add, f3inv as _inv, f3mul as _mul, f3scale as _scale, f3sub as _sub

G = 7  # coset offset


def T(x):
    return tuple(int(v) % P for v in x)


def add3(a, b):
    return T(_add(a, b))


def sub3(a, b):
    return T(_sub(a, b))


def mul3(a, b):
    return T(_mul(a, b))


def scale3(a, s):
    return T(_scale(a, s))


def eval_ir(prog, main_at, aux_at, rap, alpha, toff):
    """Big-integer evaluation of the captured program (every value as Fp3)."""
    vals = []
    for op in prog.nodes:
        k = op[0]
        if k == "const_base":
            v = (prog.base_consts[op[1]], 0, 0)
        elif k == "const_ext":
            v = T(prog.ext_consts[op[1]])
        elif k == "var":
            v = (main_at(op[2], op[4]), 0, 0) if op[1] else T(aux_at(op[2], op[4]))
        elif k == "rap":
            v = T(rap[op[1]])
        elif k == "alpha":
            v = T(alpha[op[1]])
        elif k == "offset":
            v = T(toff)
        elif k == "add":
            v = add3(vals[op[1]], vals[op[2]])
        elif k == "sub":
            v = sub3(vals[op[1]], vals[op[2]])
        elif k == "mul":
            v = mul3(vals[op[1]], vals[op[2]])
        elif k == "neg":
            v = sub3((0, 0, 0), vals[op[1]])
        else:
            v = vals[op[1]]
        vals.append(v)
    return [vals[r] for r in prog.roots]


def rand3(rng):
    return (rng.randrange(P), rng.randrange(P), rng.randrange(P))


def _logup_setup(n, rng, cfg=5):
    its = A.four_bus_interactions()
    a = A.logup_air(24, its)
    d = A.descriptor(a)
    main = workload.trace_table(cfg, 0, n, 24)
    main[:, 16:20] %= np.uint64(5)  # small multiplicities
    z, alpha = rand3(rng), rand3(rng)
    from test_oracle_logup import alpha_powers
    ap = alpha_powers(alpha, max(d.alpha_powers_len, 8))
    return a, d, main, z, alpha, ap


@pytest.mark.parametrize("which", ["fib", "quad", "logup"])
def test_walker_matches_captured_program(ora, which):
    rng = random.Random(hash(which) & 0xFFFF)
    N, step = 64, 4
    if which == "fib":
        air, ma = A.fibonacci_multi_column(3), 0
    elif which == "quad":
        air, ma = A.quadratic_multi(2), 0
    else:
        air = A.logup_air(24, A.four_bus_interactions())
        ma = air.num_aux
    main = np.array([[rng.getrandbits(64) for _ in range(air.num_main)] for _ in range(N)], dtype=np.uint64)
    aux = np.array([[rng.getrandbits(64) for _ in range(3 * ma)] for _ in range(N)], dtype=np.uint64) if ma else None
    rap = [rand3(rng), rand3(rng)]
    alpha = [rand3(rng) for _ in range(10)]
    toff = rand3(rng)
    got = ora.eval_constraints(air.device, main, aux, step, rap, alpha, toff)
    for i in range(0, N, 7):
        main_at = lambda off, c: int(main[(i + off * step) % N, c]) % P  # noqa: E731
        aux_at = lambda off, c: aux[(i + off * step) % N, 3 * c:3 * c + 3]  # noqa: E731
        want = eval_ir(air.program, main_at, aux_at, rap, alpha, toff)
        for c, w in enumerate(want):
            assert T(got[c, i]) == w, (which, i, c)
        if which == "fib":
            for c in range(3):
                direct = (main_at(2, c) - main_at(1, c) - main_at(0, c)) % P
                assert T(got[c, i]) == (direct, 0, 0)
        if which == "quad":
            for c in range(2):
                direct = (main_at(1, c) - main_at(0, c) ** 2) % P
                assert T(got[c, i]) == (direct, 0, 0)


def _lde(ora, trace, B):
    return ora.coset_lde_row_major(trace, B, G)


def _composition_case(ora, which, n, B, rng, exemptions_override=None, boundary_shift=0):
    if which == "fib":
        air = A.fibonacci_multi_column(3)
        seeds = np.array([[rng.randrange(P), rng.randrange(P)] for _ in range(3)], dtype=np.uint64)
        trace = A.fibonacci_trace(n, 3, seeds)
        A.with_boundaries_from_trace(air, trace)
        aux_lde, rap, alpha, toff = None, None, None, None
    elif which == "quad":
        air = A.quadratic_multi(2)
        trace = A.quadratic_trace(n, 2, np.array([rng.randrange(P) for _ in range(2)], dtype=np.uint64))
        A.with_boundaries_from_trace(air, trace)
        aux_lde, rap, alpha, toff = None, None, None, None
    else:
        air, d, trace, z, alpha_c, ap = _logup_setup(n, rng)
        aux, L, _ = ora.logup_aux(d, trace, ap, z)
        aux_lde = _lde(ora, aux.reshape(n, -1), B)
        rap, alpha = [z, alpha_c], ap
        toff = scale3(T(L), pow(n, P - 2, P))
    main_lde = _lde(ora, trace, B)
    groups, ex = air.groups()
    if exemptions_override is not None:
        ex = exemptions_override
    betas = [rand3(rng) for _ in range(air.num_constraints)]
    bnd = [(b.col, b.is_aux, b.step + boundary_shift, b.value) for b in air.boundary]
    b_betas = [rand3(rng) for _ in bnd]
    H = ora.eval_composition(air.device, main_lde, aux_lde, B, n, G, betas, groups, ex, bnd, b_betas, rap, alpha,
                             toff)
    return air, H, main_lde, aux_lde, betas, bnd, b_betas, groups, ex, rap, alpha, toff


def _high_coeffs(ora, H, bound):
    coeffs = ora.fft_two_half(np.ascontiguousarray(H, dtype=np.uint64), inverse=True)
    c = np.where(coeffs >= np.uint64(P), coeffs - np.uint64(P), coeffs)
    return c[bound + 1:]


@pytest.mark.parametrize("which", ["fib", "quad", "logup"])
def test_composition_of_valid_trace_is_low_degree(ora, which):
    rng = random.Random(100 + len(which))
    n, B = 32, 4
    air, H, *_ = _composition_case(ora, which, n, B, rng)
    deg = max(max(air.max_degree * (n - 1) + e - n for e in air.groups()[1]), n - 2)
    assert not _high_coeffs(ora, H, deg).any(), which
    # the bound is tight enough to mean something: H is not of degree < n - 2
    assert _high_coeffs(ora, H, n - 3).any()


def test_low_degree_negative_controls(ora):
    n, B = 32, 4
    # Fibonacci with one end exemption too few: the last-but-one row's
    # constraint (which wraps around) is not divided out
    air, H, *_ = _composition_case(ora, "fib", n, B, random.Random(1), exemptions_override=[1])
    assert _high_coeffs(ora, H, 2 * n).any()
    # boundary constraint at the wrong step
    air, H, *_ = _composition_case(ora, "quad", n, B, random.Random(2), boundary_shift=1)
    assert _high_coeffs(ora, H, 2 * n).any()
    # a random (non-satisfying) main trace
    rng = random.Random(3)
    air = A.fibonacci_multi_column(2)
    tr = workload.trace_table(9, 1, n, 2)
    H = ora.eval_composition(air.device, _lde(ora, tr, B), None, B, n, G, [rand3(rng) for _ in range(2)], [0, 0],
                             [2])
    assert _high_coeffs(ora, H, 2 * n).any()


def test_composition_closed_form(ora):
    rng = random.Random(11)
    n, B = 16, 4
    N = n * B
    for which in ("fib", "logup"):
        air, H, main_lde, aux_lde, betas, bnd, b_betas, groups, ex, rap, alpha, toff = _composition_case(
            ora, which, n, B, rng)
        ev = ora.eval_constraints(air.device, main_lde, aux_lde, B, rap, alpha, toff)
        wn = ora.root_of_unity(n.bit_length() - 1)
        wN = ora.root_of_unity(N.bit_length() - 1)
        for i in (0, 5, N - 1):
            x = G * pow(wN, i, P) % P
            h = (0, 0, 0)
            for c in range(air.num_constraints):
                z = pow((pow(x, n, P) - 1) % P, P - 2, P)
                for j in range(1, ex[groups[c]] + 1):
                    z = z * (x - pow(wn, n - j, P)) % P
                h = add3(h, mul3(scale3(T(ev[c, i]), z), betas[c]))
            for (col, is_aux, step, val), bb in zip(bnd, b_betas):
                t = T(aux_lde[i, 3 * col:3 * col + 3]) if is_aux else (int(main_lde[i, col]) % P, 0, 0)
                zb = pow((x - pow(wn, step, P)) % P, P - 2, P)
                h = add3(h, mul3(scale3(bb, zb), sub3(t, val)))
            assert T(H[i]) == h, (which, i)


def _coset_coeffs(ora, vals, offset):
    """Coefficients (mod p, Python ints) of the polynomial taking vals[i] at offset*w_N^i."""
    N = vals.shape[0]
    raw = ora.fft_two_half(np.ascontiguousarray(vals, dtype=np.uint64), inverse=True)
    ninv = pow(N, P - 2, P)
    ginv = pow(offset, P - 2, P)
    out = []
    for j in range(N):
        s = ninv * pow(ginv, j, P) % P
        out.append(tuple(int(v) * s % P for v in raw[j]))
    return out


def _horner3(coeffs, x):
    acc = (0, 0, 0)
    for c in reversed(coeffs):
        acc = add3(scale3(acc, x), c)
    return acc


def test_d2_decomposition_reconstructs(ora):
    rng = random.Random(31)
    N = 64
    H = np.array([[rng.randrange(P) for _ in range(3)] for _ in range(N)], dtype=np.uint64)
    halves = ora.decompose_d2(H, G)
    ginv = pow(G, P - 2, P)
    parts = ora.coset_lde_row_major(halves, 2, ginv)  # weights (N/2)^-1 g^-j
    p0 = _coset_coeffs(ora, parts[:, 0:3], G)
    p1 = _coset_coeffs(ora, parts[:, 3:6], G)
    assert all(c == (0, 0, 0) for c in p0[N // 2:]) and all(c == (0, 0, 0) for c in p1[N // 2:])
    hc = _coset_coeffs(ora, H, G)
    for _ in range(4):
        x = rng.randrange(P)
        want = _horner3(hc, x)
        got = add3(_horner3(p0[:N // 2], x * x % P), scale3(_horner3(p1[:N // 2], x * x % P), x))
        assert got == want
