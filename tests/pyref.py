"""Independent big-integer arithmetic used to pin the oracle on small inputs:
plain `% p` modular arithmetic, schoolbook polynomial products mod w^3 - 2,
O(n^2) DFTs and Horner evaluation. None of it shares code with the oracle's
reduction tricks, so a wrong reduction, twiddle or index there shows up here.
"""
P = 0xFFFFFFFF00000001


def canon(x):
    return int(x) % P


def f3mul(a, b):
    a = [int(v) % P for v in a]
    b = [int(v) % P for v in b]
    c = [0] * 5
    for i in range(3):
        for j in range(3):
            c[i + j] += a[i] * b[j]
    # w^3 = 2, w^4 = 2w
    return [(c[0] + 2 * c[3]) % P, (c[1] + 2 * c[4]) % P, c[2] % P]


def f3add(a, b):
    return [(int(x) + int(y)) % P for x, y in zip(a, b)]


def f3sub(a, b):
    return [(int(x) - int(y)) % P for x, y in zip(a, b)]


def f3scale(a, s):
    return [int(x) * int(s) % P for x in a]


def f3pow(a, e):
    r = [1, 0, 0]
    b = [int(v) % P for v in a]
    while e:
        if e & 1:
            r = f3mul(r, b)
        b = f3mul(b, b)
        e >>= 1
    return r


def f3inv(a):
    # a^(p^3 - 2) in the field of order p^3
    return f3pow(a, P ** 3 - 2)


def root_of_unity(log_n, root_2_32):
    return pow(root_2_32, 1 << (32 - log_n), P)


def dft(vals, omega):
    n = len(vals)
    return [sum(int(v) * pow(omega, i * k, P) for i, v in enumerate(vals)) % P for k in range(n)]


def horner(coeffs, x):
    acc = 0
    for c in reversed(coeffs):
        acc = (acc * x + int(c)) % P
    return acc


def horner3(coeffs3, x3):
    acc = [0, 0, 0]
    for c in reversed(coeffs3):
        acc = f3add(f3mul(acc, x3), c)
    return acc


def bitrev(i, bits):
    return int(format(i, f"0{bits}b")[::-1], 2) if bits else 0
