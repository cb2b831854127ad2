"""ORACLE — TEST INFRASTRUCTURE ONLY.

AIR-table proofs on the CPU: a replay of every proving round with the oracle
primitives (the checker of the GPU prover), and a restatement of the
reference verifier that accepts or rejects a proof from its transcript alone.

Round order and formulas follow crypto/stark/src/prover.rs (R1 :3426-3538,
:3867, :3928-3931; R2 :4453-4482 with :1683-1760; R3 :4494-4554; R4
:2143-2299) and crypto/stark/src/verifier.rs:
  - step 2 (:266-424): H(z) = sum_i H_i(z^P) z^i equals the constraint
    combination at the OOD frame (transition values times their zerofier
    inverses, boundary quotients (t(z) - v)/(z - g^step)), with challenges
    replayed from the transcript;
  - step 3 (:454-555, :740-838, :839-1157): the DEEP value at every query
    point and its symmetric point rebuilt from the openings and the OOD
    values, then folded through every committed FRI layer (Merkle-checked)
    down to the terminal codeword rebuilt from the final coefficients
    (fri/terminal.rs:113-155);
  - step 4 (:597-704): the main, aux and composition row pairs open against
    their roots.
Shape metadata, restated here from the reference rules (the AIR description
only says which kind of AIR it is):
  - composition parts = composition_poly_degree_bound(n) / n: AirWithBuses
    (lookup.rs:1063-1078) (max_degree - 1).max(1); an example AIR states its
    own bound (`example_parts`: fibonacci_multi_column.rs:93 -> 1,
    quadratic_air.rs:149 -> 2). One part is H itself, two the degree-2 split
    (prover.rs:1605-1680), three or more H's interpolant broken by stride P
    and each part evaluated on the LDE coset (prover.rs:1858-1900,
    polynomial.rs:137-150 break_in_parts);
  - frame rows: AirWithBuses [0, 1] (lookup.rs:950), example AIRs theirs;
  - next-row openings: AirWithBuses only the accumulated LogUp column
    (lookup.rs:1042-1053), none without interactions; example AIRs every
    column (traits.rs:217-220). The trace-term coefficients follow
    ood.rs:57-92 build_pruned_trace_term_coeffs (pruned openings weigh zero),
    the proof's next-row block holds the opened columns only (ood.rs:94-126).

Used only by tests. AIR descriptions come in as data: an object with the
lowered constraint program (`device`), `num_main`, `num_aux`, `max_degree`,
`bus` (an AirWithBuses table), `example_parts` (an example AIR's degree
bound / n), `num_constraints`, `groups()`, `boundary`, `interactions` and, for
tables with bus interactions, the flat bus layout `bus_descriptor`.
"""
from __future__ import annotations

import numpy as np

from . import (P, Transcript, barycentric_base, commit_rows_bit_reversed, decompose_d2, deep_composition,
               eval_composition, fft_two_half, fri_commit, fri_layout, generate_nonce, gl_op, is_valid_nonce,
               keccak256, lde_commit, lde_weights, logup_aux, merkle_build, merkle_path, merkle_verify,
               root_of_unity)

# ---------------- Fp3 = Fp[w]/(w^3 - 2) on Python ints ----------------


def f3(v):
    return tuple(int(x) % P for x in v)


def f3add(a, b):
    return tuple((x + y) % P for x, y in zip(a, b))


def f3sub(a, b):
    return tuple((x - y) % P for x, y in zip(a, b))


def f3mul(a, b):
    a0, a1, a2 = a
    b0, b1, b2 = b
    return ((a0 * b0 + 2 * (a1 * b2 + a2 * b1)) % P, (a0 * b1 + a1 * b0 + 2 * a2 * b2) % P,
            (a0 * b2 + a1 * b1 + a2 * b0) % P)


def f3scale(a, s):
    return tuple(x * s % P for x in a)


def f3pow(a, e):
    r, b = (1, 0, 0), a
    while e:
        if e & 1:
            r = f3mul(r, b)
        b = f3mul(b, b)
        e >>= 1
    return r


def f3inv(a):
    if a == (0, 0, 0):
        raise ZeroDivisionError("Fp3 zero")
    return f3pow(a, P ** 3 - 2)


def f3powers(g, count):
    out, cur = [], (1, 0, 0)
    for _ in range(count):
        out.append(cur)
        cur = f3mul(cur, g)
    return out


ONE, ZERO = (1, 0, 0), (0, 0, 0)


def bitrev(i, bits):
    return int(format(i, f"0{bits}b")[::-1], 2) if bits else 0


def fp3_of_components(c0, c1, c2):
    """a0 + a1 w + a2 w^2 for Fp3 values a_i of the component columns."""
    return f3add(f3add(c0, f3mul((0, 1, 0), c1)), f3mul((0, 0, 1), c2))


def be_words(words) -> bytes:
    return b"".join((int(x) % P).to_bytes(8, "big") for x in words)


def row_pair_leaf(r0, r1) -> bytes:
    """Leaf of a row pair (crypto/stark/src/commitment.rs:81-91)."""
    return keccak256(be_words(r0) + be_words(r1))


def fri_pair_leaf(e0, e1) -> bytes:
    return keccak256(be_words(e0) + be_words(e1))


def commit_rows(rows_natural: np.ndarray):
    """Row-pair Merkle tree of an LDE given in natural row order."""
    N = rows_natural.shape[0]
    bits = N.bit_length() - 1
    leaves = np.array([np.frombuffer(row_pair_leaf(rows_natural[bitrev(2 * i, bits)],
                                                   rows_natural[bitrev(2 * i + 1, bits)]), dtype=np.uint8)
                       for i in range(N // 2)])
    return merkle_build(leaves)


def sample_z_ood(t: Transcript, n: int, N: int, offset: int):
    """is_transcript.rs:44-75: resample while z^n = 1 or z^N = offset^N."""
    gN = pow(offset, N, P)
    while True:
        z = f3(t.sample_fp3())
        zn = f3pow(z, n)
        if zn == ONE or f3pow(zn, N // n) == (gN, 0, 0):
            continue
        return z


def frame_rows(air) -> int:
    """Transition offsets: [0, 1] for AirWithBuses (lookup.rs:950)."""
    return 2 if air.bus else int(air.device.max_offset) + 1


def num_parts(air) -> int:
    """composition_poly_degree_bound(n) / n (prover.rs:1705)."""
    if air.bus:
        return max(air.max_degree - 1, 1)  # lookup.rs:1063-1078
    return int(air.example_parts)


def next_row_cols(air) -> list:
    """Columns opened at z*g ([main | aux] indices): lookup.rs:1042-1053 for
    AirWithBuses (the accumulated column sits after the committed term
    columns, lookup.rs:1946), every column for the example AIRs."""
    W = air.num_main + air.num_aux
    if air.bus:
        return [air.num_main + air.num_aux - 1] if air.interactions else []
    return list(range(W))


def gamma_exponents(W, K, next_cols):
    """ood.rs:57-92: exponents of gamma of the surviving trace openings
    {(column, frame row): e}: every column's current row, then the opened
    columns (ascending) at rows 1..K-1."""
    e = {(j, 0): j for j in range(W)}
    p = W
    for j in sorted(next_cols):
        for k in range(1, K):
            e[(j, k)] = p
            p += 1
    return e


def break_in_parts_lde(H: np.ndarray, num_parts_: int, g: int) -> np.ndarray:
    """prover.rs:1858-1900 for three or more parts: H (N, 3) natural values on
    g*<w_N>; interpolate_offset_fft (polynomial.rs:359-372: inverse FFT, 1/N,
    scale by g^-k), break_in_parts (polynomial.rs:137-150: part i takes
    coefficients i, i+P, i+2P, ...), then each part's evaluate_offset_fft at
    the N coset points (polynomial.rs:316-327). Returns (N, 3P) natural rows."""
    N = H.shape[0]
    a = fft_two_half(np.ascontiguousarray(H, dtype=np.uint64), inverse=True)  # N * c_k * g^k
    ginv = pow(g, P - 2, P)
    sc = lde_weights(N, ginv)  # N^-1 g^-k
    coef = np.stack([gl_op("mul", a[:, c].copy(), sc) for c in range(3)], axis=1)
    gpow = gl_op("mul", lde_weights(N, g), np.full(N, N, dtype=np.uint64))  # g^t
    out = np.zeros((N, 3 * num_parts_), dtype=np.uint64)
    for i in range(num_parts_):
        part = coef[i::num_parts_]
        L = part.shape[0]
        d = np.zeros((N, 3), dtype=np.uint64)
        for c in range(3):
            d[:L, c] = gl_op("mul", part[:, c].copy(), gpow[:L].copy())
        out[:, 3 * i:3 * i + 3] = fft_two_half(d)
    return out


# ---------------- constraint program at the OOD frame ----------------

def eval_program_ext(dev, frame, rap, alpha, toff):
    """Constraint values of a lowered program (constraint_ir/device.rs) on an
    Fp3 frame: frame[k][("main"|"aux", col)]. Slot and operand semantics of
    the reference walker (device.rs:390); values taken at their emitting node
    when the program lists them, else from the pinned root slots."""
    nodes = np.asarray(dev.nodes, dtype=np.int64)
    bc = [int(x) for x in np.asarray(dev.base_consts).reshape(-1)]
    ec = [f3(v) for v in np.asarray(dev.ext_consts).reshape(-1, 3)]
    base = [ZERO] * max(dev.num_base_slots, 1)
    ext = [ZERO] * max(dev.num_ext_slots, 1)
    roots = [int(r) for r in dev.roots]
    emit = {}
    if getattr(dev, "root_nodes", None) is not None:
        for c, k in enumerate(dev.root_nodes):
            emit.setdefault(int(k), []).append(c)
    out = [None] * len(roots)

    def operand(enc):
        kind, x = enc >> 29, enc & 0x1FFFFFFF
        return [lambda: base[x], lambda: ext[x], lambda: (bc[x], 0, 0), lambda: ec[x], lambda: rap[x],
                lambda: alpha[x], lambda: toff][kind]()

    for k, (op, a, b, res) in enumerate(nodes.tolist()):
        if op == 0:
            v = (bc[a], 0, 0)
        elif op == 1:
            v = ec[a]
        elif op == 2:
            col, is_main, off = a & 0xFFFF, (b >> 16) & 1, (b >> 8) & 0xFF
            v = frame[off][("main" if is_main else "aux", col)]
        elif op == 3:
            v = rap[a]
        elif op == 4:
            v = alpha[a]
        elif op == 5:
            v = toff
        elif op in (6, 7, 8):
            p, q = operand(a), operand(b)
            v = f3add(p, q) if op == 6 else f3sub(p, q) if op == 7 else f3mul(p, q)
        elif op == 9:
            v = f3sub(ZERO, operand(a))
        else:
            v = operand(a)
        slot = res & 0x7FFFFFFF
        (ext if res >> 31 else base)[slot] = v
        for c in emit.get(k, []):
            out[c] = v
    for c, r in enumerate(roots):
        if out[c] is None:
            out[c] = (ext if r >> 31 else base)[r & 0x7FFFFFFF]
    return out


# ---------------- the prover replay ----------------

def replay_air_tables(tables, blowup, offset, fri_k, num_queries, grinding, label):
    """tables: [(air, host trace (n, m) uint64)]. Proof dicts in table order,
    field for field what lambda_vm_amd.prover returns."""
    Bf, g = blowup, offset
    t = Transcript(label)
    commits = [_commit_main(air, tr, Bf, g) for air, tr in tables]
    for c in commits:
        if len(c) == 3:  # a preprocessed table's fixed-columns root first (prover.rs:3457-3466)
            t.append_bytes(bytes(c[2][0]))
        t.append_bytes(bytes(c[1][0]))
    z_logup = alpha = None
    if any(air.interactions for air, _ in tables):
        z_logup, alpha = f3(t.sample_fp3()), f3(t.sample_fp3())
    forks = []
    for idx in range(len(tables)):
        f = t.clone()
        if len(tables) > 1:
            f.append_bytes(int(idx).to_bytes(8, "little"))
        forks.append(f)
    out = []
    for idx, (air, tr) in enumerate(tables):
        out.append(_replay_table(air, tr, commits[idx], forks[idx], Bf, g, fri_k, num_queries, grinding,
                                 z_logup, alpha))
    return out


def _commit_main(air, tr, Bf, g):
    """(LDE, main tree) of a table; for a preprocessed table (first k columns
    fixed) the main tree covers the multiplicity columns [k, m) and a third
    entry is the fixed columns' tree (prover.rs:1194-1290, lde.rs:726)."""
    k = getattr(air, "num_precomputed", 0)
    if not k:
        return lde_commit(tr, Bf, g)
    lde, _ = lde_commit(tr, Bf, g)
    m = tr.shape[1]
    return lde, commit_rows_bit_reversed(lde, 1, k, m), commit_rows_bit_reversed(lde, 1, 0, k)


def precomputed_commitment(air, Bf, g) -> bytes:
    """The verifier's own commitment of a preprocessed table's fixed columns:
    coset LDE of air.precomputed() and its row-pair Keccak tree (the static
    commitments of prover/src/tables/keccak_rc.rs:97-116, bitwise.rs:199-218
    are this value for coset offset 3)."""
    _, nodes = lde_commit(np.ascontiguousarray(air.precomputed(), dtype=np.uint64), Bf, g, want_lde=False)
    return bytes(nodes[0])


def _replay_table(air, tr, commit, t, Bf, g, fri_k, num_queries, grinding, z_logup, alpha):
    n, m = tr.shape
    N = n * Bf
    bits = N.bit_length() - 1
    main_lde, main_nodes = commit[0], commit[1]
    pre_nodes = commit[2] if len(commit) == 3 else None
    kpre = air.num_precomputed if pre_nodes is not None else 0
    rap = ap = toff = None
    aux_lde = aux_nodes = L = None
    if air.interactions:
        desc = air.bus_descriptor  # the flat bus layout, data of the AIR under test
        ap = f3powers(alpha, desc.alpha_powers_len)
        aux, L, _ = logup_aux(desc, tr, np.array(ap, dtype=np.uint64), np.array(z_logup, dtype=np.uint64))
        aux_lde, aux_nodes = lde_commit(aux.reshape(n, -1), Bf, g)
        t.append_bytes(bytes(aux_nodes[0]))
        L = f3(L)
        t.append_fp3(np.array(L, dtype=np.uint64))
        rap = [z_logup, alpha]
        toff = f3scale(L, pow(n, P - 2, P))
    beta = f3(t.sample_fp3())
    T, nb = air.num_constraints, len(air.boundary)
    coeffs = f3powers(beta, T + nb)
    groups, ex = air.groups()
    bnd = [(b.col, b.is_aux, b.step, b.value) for b in air.boundary]
    H = eval_composition(air.device, main_lde, aux_lde, Bf, n, g, coeffs[:T], groups, ex, bnd, coeffs[T:],
                         rap, ap, None if toff is None else np.array(toff, dtype=np.uint64))
    Pn = num_parts(air)
    if Pn == 2:
        halves = decompose_d2(H, g)
        parts_lde, parts_nodes = lde_commit(halves, 2, pow(g, P - 2, P))
    elif Pn == 1:
        parts_lde, parts_nodes = H, commit_rows(H)
    else:
        parts_lde = break_in_parts_lde(H, Pn, g)
        parts_nodes = commit_rows_bit_reversed(parts_lde)
    t.append_bytes(bytes(parts_nodes[0]))
    z = sample_z_ood(t, n, N, g)
    K = frame_rows(air)
    wn = root_of_unity(n.bit_length() - 1)
    points = [f3scale(z, pow(wn, k, P)) for k in range(K)]
    ood = []
    for p in points:
        row = [f3(v) for v in barycentric_base(main_lde, n, Bf, g, np.array(p, dtype=np.uint64))]
        if aux_lde is not None:
            comp = [f3(v) for v in barycentric_base(aux_lde, n, Bf, g, np.array(p, dtype=np.uint64))]
            row += [fp3_of_components(*comp[3 * c:3 * c + 3]) for c in range(len(comp) // 3)]
        ood.append(row)
    W = len(ood[0])
    ncols = sorted(next_row_cols(air))
    for j in range(W):
        t.append_fp3(np.array(ood[0][j], dtype=np.uint64))
    for j in ncols:
        for k in range(1, K):
            t.append_fp3(np.array(ood[k][j], dtype=np.uint64))
    zP = f3pow(z, Pn)
    # each part's value at z^P by the barycentric formula over the n
    # trace-coset points (prover.rs:2012-2100: domain_size n, stride blowup)
    comp = [f3(v) for v in barycentric_base(parts_lde, n, Bf, g, np.array(zP, dtype=np.uint64))]
    parts_ood = [fp3_of_components(*comp[3 * i:3 * i + 3]) for i in range(Pn)]
    for v in parts_ood:
        t.append_fp3(np.array(v, dtype=np.uint64))
    gamma = f3(t.sample_fp3())
    expo = gamma_exponents(W, K, ncols)
    S = len(expo)
    gp = f3powers(gamma, S + Pn)
    trace_ood = np.array([ood[k][j] for j in range(W) for k in range(K)], dtype=np.uint64)
    gammas_tr = np.array([gp[expo[(j, k)]] if (j, k) in expo else ZERO for j in range(W) for k in range(K)],
                         dtype=np.uint64)
    deep = deep_composition(main_lde, g, np.array(points, dtype=np.uint64), trace_ood, gammas_tr, aux=aux_lde,
                            parts=parts_lde, z_parts=np.array(zP, dtype=np.uint64),
                            parts_ood=np.array(parts_ood, dtype=np.uint64), gammas_h=np.array(gp[S:],
                                                                                               dtype=np.uint64))
    perm = np.array([bitrev(i, bits) for i in range(N)], dtype=np.int64)
    roots, fcoeffs, layers = fri_commit(deep[perm], t, g, Bf.bit_length() - 1, fri_k, keep_layers=True)
    nonce = None
    if grinding > 0:
        nonce = generate_nonce(t.state(), grinding)
        t.append_bytes(int(nonce).to_bytes(8, "big"))
    iotas = [t.sample_u64(N // 2) for _ in range(num_queries)]
    # openings
    def open_rows(lde, nodes):
        rows = np.array([[lde[bitrev(2 * i, bits)], lde[bitrev(2 * i + 1, bits)]] for i in iotas], dtype=np.uint64)
        paths = [merkle_path(nodes, i) for i in iotas]
        return rows, paths
    main_rows, main_paths = open_rows(main_lde, main_nodes)
    pre_rows = pre_paths = None
    if pre_nodes is not None:
        pre_rows, main_rows = main_rows[:, :, :kpre].copy(), main_rows[:, :, kpre:].copy()
        pre_paths = [merkle_path(pre_nodes, i) for i in iotas]
    aux_rows = aux_paths = None
    if aux_lde is not None:
        aux_rows, aux_paths = open_rows(aux_lde, aux_nodes)
    parts_rows, parts_paths = open_rows(parts_lde, parts_nodes)
    fri_sym, fri_paths, off = [], [], 0
    for li in range(len(roots)):
        size = N >> (li + 1)
        ev = layers[off:off + size]
        off += size
        lnodes = merkle_build(np.array([np.frombuffer(fri_pair_leaf(ev[2 * j], ev[2 * j + 1]), dtype=np.uint8)
                                        for j in range(size // 2)]))
        fri_sym.append([f3(ev[(i >> li) ^ 1]) for i in iotas])
        fri_paths.append([merkle_path(lnodes, (i >> li) >> 1) for i in iotas])
    return dict(trace_rows=n, main_root=bytes(main_nodes[0]), aux_root=None if aux_nodes is None else bytes(aux_nodes[0]),
                table_contribution=L, composition_root=bytes(parts_nodes[0]), ood_block0=[ood[0][j] for j in range(W)],
                ood_block1=np.array([[ood[k][j] for k in range(1, K)] for j in ncols],
                                    dtype=np.uint64).reshape(len(ncols), K - 1, 3), parts_ood=parts_ood,
                fri_roots=[bytes(r) for r in roots], fri_coeffs=[f3(c) for c in fcoeffs], nonce=nonce, iotas=iotas,
                fri_sym=fri_sym, fri_paths=fri_paths, main_rows=main_rows, main_paths=main_paths, aux_rows=aux_rows,
                aux_paths=aux_paths, parts_rows=parts_rows, parts_paths=parts_paths, transcript_state=t.state(),
                challenges=dict(z_logup=z_logup, alpha=alpha, beta=beta, z=z, gamma=gamma),
                precomputed_root=None if pre_nodes is None else bytes(pre_nodes[0]), precomputed_rows=pre_rows,
                precomputed_paths=pre_paths)


# ---------------- the verifier ----------------

def _get(proof, name):
    return proof[name] if isinstance(proof, dict) else getattr(proof, name)


def _f3_list(x):
    """Fp3 values of an array or nested lists of ints (never through floats)."""
    if isinstance(x, np.ndarray):
        return [f3(v) for v in x.astype(np.uint64).reshape(-1, 3)]
    flat = []

    def walk(v):
        if isinstance(v, (list, tuple)) and v and isinstance(v[0], (list, tuple, np.ndarray)):
            for w in v:
                walk(w)
        else:
            flat.append(f3(v))
    walk(x)
    return flat


def verify_air_tables(airs, proofs, blowup, offset, fri_k, num_queries, grinding, label):
    """airs: [AIR description] in table order; proofs: dicts or objects with
    the AirTableProof fields. Returns (accepted, reason)."""
    if len(airs) != len(proofs):
        return False, "table count"
    t = Transcript(label)
    for air, pr in zip(airs, proofs):
        if getattr(air, "num_precomputed", 0):
            # the fixed columns' root must be the verifier's own commitment of them
            pre = _get(pr, "precomputed_root")
            if pre is None or bytes(pre) != precomputed_commitment(air, blowup, offset):
                return False, "precomputed commitment"
            t.append_bytes(bytes(pre))
        elif _get(pr, "precomputed_root") is not None:
            return False, "unexpected precomputed commitment"
        t.append_bytes(bytes(_get(pr, "main_root")))
    z_logup = alpha = None
    if any(air.interactions for air in airs):
        z_logup, alpha = f3(t.sample_fp3()), f3(t.sample_fp3())
    for idx, (air, pr) in enumerate(zip(airs, proofs)):
        f = t.clone()
        if len(airs) > 1:
            f.append_bytes(int(idx).to_bytes(8, "little"))
        ok, why = _verify_table(air, pr, f, blowup, offset, fri_k, num_queries, grinding, z_logup, alpha)
        if not ok:
            return False, f"table {idx}: {why}"
    return True, ""


def _verify_table(air, pr, t, Bf, g, fri_k, num_queries, grinding, z_logup, alpha):
    G = lambda k: _get(pr, k)  # noqa: E731
    n = int(G("trace_rows"))
    N = n * Bf
    bits = N.bit_length() - 1
    m = air.num_main
    rap = ap = toff = None
    if air.interactions:
        if G("aux_root") is None or G("table_contribution") is None:
            return False, "missing aux commitment"
        t.append_bytes(bytes(G("aux_root")))
        L = f3(G("table_contribution"))
        t.append_fp3(np.array(L, dtype=np.uint64))
        rap = [z_logup, alpha]
        ap = f3powers(alpha, air.bus_descriptor.alpha_powers_len)
        toff = f3scale(L, pow(n, P - 2, P))
    beta = f3(t.sample_fp3())
    T, nb = air.num_constraints, len(air.boundary)
    coeffs = f3powers(beta, T + nb)
    t.append_bytes(bytes(G("composition_root")))
    z = sample_z_ood(t, n, N, g)
    K = frame_rows(air)
    Pn = num_parts(air)
    ma = air.num_aux
    W = m + ma
    ncols = sorted(next_row_cols(air))
    b0 = _f3_list(G("ood_block0"))
    b1 = _f3_list(G("ood_block1"))
    pood = _f3_list(G("parts_ood"))
    if len(b0) != W or len(b1) != len(ncols) * (K - 1) or len(pood) != Pn:
        return False, "OOD shape"
    b1 = [[f3(b1[i * (K - 1) + k]) for k in range(K - 1)] for i in range(len(ncols))]
    for v in b0:
        t.append_fp3(np.array(v, dtype=np.uint64))
    for i in range(len(ncols)):
        for k in range(K - 1):
            t.append_fp3(np.array(b1[i][k], dtype=np.uint64))
    for v in pood:
        t.append_fp3(np.array(v, dtype=np.uint64))
    # the full frame with pruned next-row entries zero (ood.rs reconstruct_ood_full)
    ood = [b0] + [[ZERO] * W for _ in range(K - 1)]  # ood[k][j]
    for i, j in enumerate(ncols):
        for k in range(K - 1):
            ood[k + 1][j] = b1[i][k]
    gamma = f3(t.sample_fp3())
    expo = gamma_exponents(W, K, ncols)
    S = len(expo)
    gp = f3powers(gamma, S + Pn)
    gw = lambda j, k: gp[expo[(j, k)]] if (j, k) in expo else ZERO  # noqa: E731
    # FRI transcript replay
    tf, nc, tl, ek = fri_layout(bits, Bf.bit_length() - 1, fri_k)
    roots = [bytes(r) for r in G("fri_roots")]
    fco = _f3_list(G("fri_coeffs"))
    if len(roots) != nc or len(fco) != (1 << ek):
        return False, "FRI shape"
    zetas = []
    for li in range(nc):
        zetas.append(f3(t.sample_fp3()))
        t.append_bytes(roots[li])
    if tf > 0:
        zetas.append(f3(t.sample_fp3()))
    for c in fco:
        t.append_fp3(np.array(c, dtype=np.uint64))
    if grinding > 0:
        nonce = G("nonce")
        if nonce is None or not is_valid_nonce(t.state(), int(nonce), grinding):
            return False, "proof of work"
        t.append_bytes(int(nonce).to_bytes(8, "big"))
    iotas = [t.sample_u64(N // 2) for _ in range(num_queries)]
    if [int(i) for i in G("iotas")] != iotas:
        return False, "query indices"
    # step 2: the composition at z
    wn = root_of_unity(n.bit_length() - 1)
    frame = [{("main", j): ood[k][j] for j in range(m)} | {("aux", j): ood[k][m + j] for j in range(ma)}
             for k in range(K)]
    vals = eval_program_ext(air.device, frame, rap, ap, toff)
    groups, ex = air.groups()
    zn_minus_1 = f3sub(f3pow(z, n), ONE)
    acc = ZERO
    for c, v in enumerate(vals):
        e = ex[groups[c]]
        num = ONE
        for j in range(1, e + 1):
            num = f3mul(num, f3sub(z, (pow(wn, n - j, P), 0, 0)))
        acc = f3add(acc, f3mul(f3mul(coeffs[c], v), f3mul(num, f3inv(zn_minus_1))))
    for bi, b in enumerate(air.boundary):
        tv = ood[0][(m + b.col) if b.is_aux else b.col]
        den = f3sub(z, (pow(wn, b.step, P), 0, 0))
        acc = f3add(acc, f3mul(coeffs[T + bi], f3mul(f3sub(tv, f3(b.value)), f3inv(den))))
    claimed = ZERO
    for v in reversed(pood):
        claimed = f3add(f3mul(claimed, z), v)
    if claimed != acc:
        return False, "composition polynomial at z"
    # step 4: openings
    main_rows = np.asarray(G("main_rows"))
    parts_rows = np.asarray(G("parts_rows"))
    aux_rows = G("aux_rows")
    main_root, comp_root = bytes(G("main_root")), bytes(G("composition_root"))
    kpre = getattr(air, "num_precomputed", 0)
    if kpre:
        pre_rows = np.asarray(G("precomputed_rows"))
        if pre_rows.shape != (len(iotas), 2, kpre) or main_rows.shape[2] != m - kpre:
            return False, "preprocessed opening shape"
    for q, i in enumerate(iotas):
        if not merkle_verify(main_root, row_pair_leaf(main_rows[q, 0], main_rows[q, 1]), i, G("main_paths")[q]):
            return False, "main opening"
        if kpre and not merkle_verify(bytes(G("precomputed_root")), row_pair_leaf(pre_rows[q, 0], pre_rows[q, 1]),
                                      i, G("precomputed_paths")[q]):
            return False, "precomputed opening"
        if not merkle_verify(comp_root, row_pair_leaf(parts_rows[q, 0], parts_rows[q, 1]), i,
                             G("parts_paths")[q]):
            return False, "composition opening"
        if ma and not merkle_verify(bytes(G("aux_root")), row_pair_leaf(aux_rows[q][0], aux_rows[q][1]), i,
                                    G("aux_paths")[q]):
            return False, "aux opening"
    # step 3: DEEP values at the query points, FRI folding, terminal
    row_sum = []
    for k in range(K):
        s = ZERO
        for j in range(W):
            s = f3add(s, f3mul(gw(j, k), ood[k][j]))
        row_sum.append(s)
    zP = f3pow(z, Pn)
    h_sum_z = ZERO
    for i in range(Pn):
        h_sum_z = f3add(h_sum_z, f3mul(gp[S + i], pood[i]))
    wN = root_of_unity(bits)
    toff_fri = pow(g, 1 << tf, P)
    wT = root_of_unity(tl.bit_length() - 1)
    tbits = tl.bit_length() - 1

    def terminal(idx):
        x = toff_fri * pow(wT, bitrev(idx, tbits), P) % P
        acc_ = ZERO
        for c in reversed(fco):
            acc_ = f3add(f3scale(acc_, x), c)
        return acc_

    fri_sym = G("fri_sym")
    fri_paths = G("fri_paths")
    for q, i in enumerate(iotas):
        deep_pair = []
        for s in (0, 1):
            x = g * pow(wN, bitrev(2 * i + s, bits), P) % P
            base = list(pre_rows[q, s]) + list(main_rows[q, s]) if kpre else list(main_rows[q, s])
            tvals = [(int(base[j]), 0, 0) for j in range(m)]
            if ma:
                aw = [int(v) for v in np.asarray(aux_rows[q][s]).reshape(-1)]
                tvals += [(aw[3 * c] % P, aw[3 * c + 1] % P, aw[3 * c + 2] % P) for c in range(ma)]
            val = ZERO
            for k in range(K):
                bs = ZERO
                for j in range(W):
                    bs = f3add(bs, f3mul(gw(j, k), tvals[j]))
                den = f3sub((x, 0, 0), f3scale(z, pow(wn, k, P)))
                val = f3add(val, f3mul(f3sub(bs, row_sum[k]), f3inv(den)))
            pw = [int(v) for v in np.asarray(parts_rows[q, s]).reshape(-1)]
            hs = ZERO
            for pi in range(Pn):
                hs = f3add(hs, f3mul(gp[S + pi], (pw[3 * pi] % P, pw[3 * pi + 1] % P, pw[3 * pi + 2] % P)))
            val = f3add(val, f3mul(f3sub(hs, h_sum_z), f3inv(f3sub((x, 0, 0), zP))))
            deep_pair.append(val)
        p0, p0s = deep_pair
        if tf == 0:
            if p0 != terminal(2 * i) or p0s != terminal(2 * i + 1):
                return False, "terminal (no fold)"
            continue
        x_inv = pow(g * pow(wN, bitrev(2 * i, bits), P) % P, P - 2, P)
        v = f3add(f3add(p0, p0s), f3mul(f3scale(zetas[0], x_inv), f3sub(p0, p0s)))
        idx = i
        xi = x_inv
        for li in range(nc):
            sym = f3(fri_sym[li][q])
            pair = (sym, v) if idx % 2 == 1 else (v, sym)
            if not merkle_verify(roots[li], fri_pair_leaf(*pair), idx >> 1, fri_paths[li][q]):
                return False, f"FRI layer {li} opening"
            xi = xi * xi % P
            v = f3add(f3add(v, sym), f3mul(f3scale(zetas[li + 1], xi), f3sub(v, sym)))
            idx >>= 1
        if v != terminal(idx):
            return False, "FRI terminal"
    return True, ""
