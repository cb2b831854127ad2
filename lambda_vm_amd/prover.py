"""Proving AIR tables end to end on the GPU: every round of the reference's
multi-table prover for the synthetic AIRs, in its transcript order.

Order (crypto/stark/src/prover.rs):
  R1   every table's main LDE + Merkle root, absorbed in table order
       (:3426-3471); the shared LogUp challenges (z, alpha) when a table has
       bus interactions (:3476-3484, lookup.rs:105 LOGUP_NUM_CHALLENGES = 2);
       one transcript fork per table, domain-separated by the table index as
       u64 LE when there are several tables (:3529-3538); per table the aux
       (LogUp) trace, its LDE + root, and its table contribution L
       (:3867, :3928-3931).
  R2   beta; transition coefficients beta^0..beta^(T-1), then the boundary
       coefficients (:4453-4474); the composition polynomial on the LDE, split
       into P = degree bound / n parts (:1705; one part: H itself, two: the
       degree-2 split, three or more: interpolation, break_in_parts and each
       part evaluated on the LDE coset, :1683-1900) and committed (:4482).
  R3   z out of the trace domain and the LDE coset (is_transcript.rs:44-75);
       the trace at z * g^k for every frame offset k; the current-row block
       (all columns) then the pruned next-row block, only the columns the
       AIR reads at the next row, column by column (ood.rs:94-126,
       prover.rs:4543-4553); the parts at z^P, each by the barycentric formula
       over the n trace-coset points (:2012-2100).
  R4   gamma; the first S powers (S = surviving openings: W current-row
       values plus (K-1) per next-row column) weight the trace terms in the
       order of ood.rs:57-92 (current row: column j gets gamma^j; then each
       next-row column in ascending order, offsets 1..K-1), every pruned
       (column, offset) gets weight zero; the next P powers the parts
       (prover.rs:2155-2182); the DEEP codeword, FRI commit phase, grinding
       nonce (BE bytes, :2277-2286), queries sample_u64(N/2) (:2324-2333), FRI
       decommitments and the row-pair openings of the main, aux and parts
       trees (:2888+).
Bus AIRs (every VM table, the LogUp table) open only the accumulated column
at the next row (lookup.rs:1042-1053); the example AIRs open every column
(traits.rs:217-220).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import backend as B
from . import field as F
from .air import AirProgram
from .pipeline import sample_z_ood

P = F.P


@dataclass
class ProofOptions:
    blowup: int = 4
    coset_offset: int = 3
    fri_final_log_degree: int = 2
    num_queries: int = 16
    grinding_factor: int = 0
    vram_budget: int | None = None  # device bytes for tables in flight (None: no limit)
    # FRI commit phase of all tables in shared launches (lvm_fri_commit_phase_batch)
    # when the budget admits every table at once; else each table folds on its own thread
    batch_fri: bool = True


@dataclass
class AirTableProof:
    """One table's proof: the values a verifier reads (canonical limbs)."""
    trace_rows: int
    main_root: bytes
    aux_root: bytes | None
    table_contribution: tuple | None  # L (Fp3)
    composition_root: bytes
    ood_block0: np.ndarray   # (W, 3): every column at z
    ood_block1: np.ndarray   # (C, K-1, 3): the C next-row columns at z*g^k, k >= 1
    parts_ood: np.ndarray    # (P, 3): H_i(z^P)
    fri_roots: np.ndarray
    fri_coeffs: np.ndarray
    nonce: int | None
    iotas: np.ndarray
    fri_sym: np.ndarray      # (layers, Q, 3)
    fri_paths: list          # per layer (Q, depth, 32)
    main_rows: np.ndarray    # (Q, 2, m)
    main_paths: np.ndarray
    aux_rows: np.ndarray | None
    aux_paths: np.ndarray | None
    parts_rows: np.ndarray   # (Q, 2, 3P)
    parts_paths: np.ndarray
    transcript_state: bytes = b""
    challenges: dict = field(default_factory=dict)
    # preprocessed tables (StarkProof::lde_trace_precomputed_merkle_root and
    # DeepPolynomialOpening::precomputed_trace_polys, proof/stark.rs:65-90):
    # main_root / main_rows / main_paths then cover the multiplicity columns only
    precomputed_root: bytes | None = None
    precomputed_rows: np.ndarray | None = None   # (Q, 2, k)
    precomputed_paths: np.ndarray | None = None


def frame_offsets(air: AirProgram) -> int:
    """K: rows of the OOD frame (the AIR's transition offsets)."""
    return air.frame_rows


def num_parts(air: AirProgram) -> int:
    """Composition parts: degree bound / trace length (prover.rs:1705)."""
    return air.num_parts


def fp3_columns(vals: np.ndarray) -> np.ndarray:
    """(K, 3C, 3) component-column values -> (K, C, 3) Fp3-column values
    a = a0 + a1*w + a2*w^2 (w^3 = 2)."""
    K, w3, _ = vals.shape
    out = np.zeros((K, w3 // 3, 3), dtype=np.uint64)
    for k in range(K):
        for c in range(w3 // 3):
            a0, a1, a2 = ([int(x) for x in vals[k, 3 * c + i]] for i in range(3))
            out[k, c] = [(a0[0] + 2 * a1[2] + 2 * a2[1]) % P, (a0[1] + a1[0] + 2 * a2[2]) % P,
                         (a0[2] + a1[1] + a2[0]) % P]
    return out


def trace_term_gamma_exponents(W: int, K: int, next_cols) -> dict:
    """{(column j, frame row k): exponent of gamma} for every surviving OOD
    opening (crypto/stark/src/ood.rs:57-92 build_pruned_trace_term_coeffs with
    step size 1): the current row of every column first, then each next-row
    column (ascending) at rows 1..K-1. Pruned openings are absent (weight 0)."""
    e = {(j, 0): j for j in range(W)}
    p = W
    for j in sorted(next_cols):
        for k in range(1, K):
            e[(j, k)] = p
            p += 1
    return e


def prove_air_tables(tables, opts: ProofOptions = ProofOptions(), label: bytes = b"lambda-vm-air-tables",
                     workers: int | None = None):
    """tables: [(AirProgram, device trace DeviceBuffer (n, m) row-major, n)].
    Returns [AirTableProof] in table order. After the shared R1 every table's
    remaining rounds run on its own host thread and library stream (the
    reference chains aux build, aux commit and rounds 2-4 per table on driver
    threads, prover.rs:3590-3600); workers=1 proves them one after another."""
    t = B.Transcript(label)
    Bf, g = opts.blowup, opts.coset_offset
    # ---- R1: main commitments of every table, one batched plan; a
    # preprocessed table commits its fixed and its multiplicity columns to two
    # trees over one LDE, and its fixed-columns root goes first (prover.rs:3457-3466) ----
    weights = [B.coset_weights_dev(n, g) for _, _, n in tables]
    plain = [i for i, (air, _, _) in enumerate(tables) if not air.num_precomputed]
    mains, pres = [None] * len(tables), [None] * len(tables)
    if plain:
        plan = B.BatchPlan([(tables[i][1], weights[i], tables[i][2], tables[i][0].num_main, Bf) for i in plain],
                           use_graph=False)
        plan.run()
        for k, i in enumerate(plain):
            mains[i] = plan.take(k)
        plan.close()
    for i, (air, tr, n) in enumerate(tables):
        if air.num_precomputed:
            h, pre_root, pre_nodes = B.coset_lde_split_trees_dev(tr, n, air.num_main, Bf, weights[i],
                                                                 air.num_precomputed)
            mains[i], pres[i] = h, (pre_root, pre_nodes, air.num_precomputed)
    for i in range(len(tables)):
        if pres[i] is not None:
            t.append_bytes(pres[i][0])
        t.append_bytes(mains[i].root)
    needs_lookup = any(air.interactions for air, _, _ in tables)
    z_logup = alpha = None
    if needs_lookup:
        z_logup = [int(v) for v in t.sample_fp3()]
        alpha = [int(v) for v in t.sample_fp3()]
    forks = []
    for idx in range(len(tables)):
        f = t.clone()
        if len(tables) > 1:
            f.append_bytes(int(idx).to_bytes(8, "little"))
        forks.append(f)
    # ---- R1 (per table): LogUp aux traces of every table, committed in one
    # batched plan (each aux root then goes into its own table's fork, so the
    # transcripts are those of committing the tables one by one) ----
    auxs = [None] * len(tables)
    built = []
    for idx, (air, tr, n) in enumerate(tables):
        if air.interactions:
            desc = air_descriptor(air)
            ap = B.fp3_powers(alpha, desc.alpha_powers_len)
            aux_dev, L, _ = B.logup_aux_dev(tr, n, air.num_main, desc, ap, z_logup)
            built.append((idx, aux_dev, desc, ap, L))
    if built:
        aplan = B.BatchPlan([(aux_dev, weights[idx], tables[idx][2], 3 * desc.num_out_cols, Bf)
                             for idx, aux_dev, desc, ap, L in built], use_graph=False)
        aplan.run()
        for i, (idx, aux_dev, desc, ap, L) in enumerate(built):
            auxs[idx] = (aplan.take(i), desc, ap, L)
        aplan.close()
        for _, aux_dev, _, _, _ in built:
            aux_dev.free()

    # rounds 2-4 per table: driver threads take tables heaviest first, each
    # admitted through the device-byte gate (prover.rs:3590-3600 with
    # run_admitted / VramGate / heaviest_first, prover.rs:677-789)
    from . import _native
    from .scheduler import VramGate, estimate_table_vram_bytes, heaviest_first, run_admitted
    est = [estimate_table_vram_bytes(air.num_main, air.num_aux, n * Bf) for air, _, n in tables]
    budget = opts.vram_budget if opts.vram_budget is not None else 1 << 62
    gate = VramGate(budget)
    w = len(tables) if workers is None else workers
    init = lambda: _native.ensure_device_this_thread(0)  # noqa: E731
    # all tables' FRI layers fold together when every table can be resident at once
    batched = opts.batch_fri and len(tables) > 1 and sum(est) <= budget

    def upto_deep(idx):
        air, tr, n = tables[idx]
        return _rounds_to_deep(air, tr, n, mains[idx], weights[idx], forks[idx], opts, z_logup, alpha, auxs[idx],
                               pres[idx])

    if batched:
        # R2: every table's composition in shared launches, then per table its
        # parts commitment, R3 and DEEP on the driver threads
        inputs = [_composition_inputs(air, tr, n, mains[i], weights[i], forks[i], opts, z_logup, alpha, auxs[i])
                  for i, (air, tr, n) in enumerate(tables)]
        Hs = B.eval_composition_batch([c for c, _ in inputs])

        def after(idx):
            air, _, n = tables[idx]
            return _rounds_after_composition(air, n, mains[idx], forks[idx], opts, inputs[idx][1], Hs[idx], pres[idx])

        states = run_admitted(heaviest_first(est), est, gate, w, after, initializer=init)
        fris = [B.GpuFri(None, g, on_device=st.deep, length=st.n * Bf, adopt=True) for st in states]
        fri_out = B.fri_commit_phase_batch(fris, [st.t for st in states], [Bf.bit_length() - 1] * len(states),
                                           opts.fri_final_log_degree)
        proofs = [_finish_table(st, opts, fri, roots, coeffs) for st, fri, (roots, coeffs) in zip(states, fris, fri_out)]
    else:
        def one(idx):
            st = upto_deep(idx)
            fri = B.GpuFri(None, g, on_device=st.deep, length=st.n * Bf, adopt=True)
            roots, coeffs = fri.commit_phase(st.t, Bf.bit_length() - 1, opts.fri_final_log_degree)
            return _finish_table(st, opts, fri, roots, coeffs)
        proofs = run_admitted(heaviest_first(est), est, gate, w, one, initializer=init)
    for h, wt in zip(mains, weights):
        h.close()
        wt.free()
    return proofs


@dataclass
class _TableState:
    """One table between its DEEP codeword and the FRI commit phase."""
    n: int
    t: B.Transcript
    main: B.GpuLde
    aux: B.GpuLde | None
    parts: B.GpuLde
    deep: B.DeviceBuffer
    pre: tuple | None  # (fixed-columns root, its host node array, k) of a preprocessed table
    L: tuple | None
    ood: np.ndarray
    block1: np.ndarray
    parts_ood: np.ndarray
    challenges: dict


def _prove_table(air: AirProgram, tr: B.DeviceBuffer, n: int, main: B.GpuLde, weights: B.DeviceBuffer,
                 t: B.Transcript, opts: ProofOptions, z_logup, alpha, committed_aux=None, pre=None) -> AirTableProof:
    """Rounds of one table after R1, its FRI commit phase on its own."""
    st = _rounds_to_deep(air, tr, n, main, weights, t, opts, z_logup, alpha, committed_aux, pre)
    fri = B.GpuFri(None, opts.coset_offset, on_device=st.deep, length=n * opts.blowup, adopt=True)
    roots, coeffs = fri.commit_phase(st.t, opts.blowup.bit_length() - 1, opts.fri_final_log_degree)
    return _finish_table(st, opts, fri, roots, coeffs)


def _rounds_to_deep(air: AirProgram, tr: B.DeviceBuffer, n: int, main: B.GpuLde, weights: B.DeviceBuffer,
                    t: B.Transcript, opts: ProofOptions, z_logup, alpha, committed_aux=None, pre=None) -> _TableState:
    """Rounds of one table after R1 up to its DEEP codeword, its composition
    evaluated on its own."""
    comp, ctx = _composition_inputs(air, tr, n, main, weights, t, opts, z_logup, alpha, committed_aux)
    H = B.eval_composition(**comp)
    return _rounds_after_composition(air, n, main, t, opts, ctx, H, pre)


def _composition_inputs(air: AirProgram, tr: B.DeviceBuffer, n: int, main: B.GpuLde, weights: B.DeviceBuffer,
                        t: B.Transcript, opts: ProofOptions, z_logup, alpha, committed_aux=None):
    """R1's per-table part (aux trace, its root and contribution absorbed) and
    R2's challenge: the arguments of the table's composition (a dict for
    eval_composition / eval_composition_batch) and what the later rounds need.
    committed_aux: (aux LDE handle, LogUp descriptor, alpha powers, L) when the
    table's aux trace was committed with the other tables' (prove_air_tables),
    else it is built and committed here."""
    Bf, g = opts.blowup, opts.coset_offset
    m = air.num_main
    rap = ap = toff = None
    aux = None
    L = None
    # ---- R1 (per table): LogUp aux trace, commitment, contribution ----
    if air.interactions:
        if committed_aux is not None:
            aux, desc, ap, L = committed_aux
        else:
            desc = air_descriptor(air)
            ap = B.fp3_powers(alpha, desc.alpha_powers_len)
            aux_dev, L, _ = B.logup_aux_dev(tr, n, m, desc, ap, z_logup)
            aux = B.coset_lde_dev(aux_dev, n, 3 * desc.num_out_cols, Bf, weights)
            aux_dev.free()
        t.append_bytes(aux.root)
        L = tuple(int(v) % P for v in L)
        t.append_fp3(L)
        rap = [z_logup, alpha]
        ninv = pow(n, P - 2, P)
        toff = np.array([x * ninv % P for x in L], dtype=np.uint64)
    # ---- R2: composition ----
    beta = [int(v) for v in t.sample_fp3()]
    T, nb = air.num_constraints, len(air.boundary)
    coeffs = B.fp3_powers(beta, T + nb)
    groups, ex = air.groups()
    bnd = [(b.col, b.is_aux, b.step, b.value) for b in air.boundary]
    comp = dict(dev=air.device, main=main, aux=aux, next_step=Bf, trace_rows=n, coset_offset=g, betas=coeffs[:T],
                root_group=groups, group_exemptions=ex, boundary=bnd, b_betas=coeffs[T:], rap=rap, alpha=ap,
                table_offset=toff)
    return comp, dict(aux=aux, L=L, beta=beta, z_logup=z_logup, alpha=alpha)


def _rounds_after_composition(air: AirProgram, n: int, main: B.GpuLde, t: B.Transcript, opts: ProofOptions, ctx,
                              H: B.DeviceBuffer, pre=None) -> _TableState:
    """R2's commitment of the composition parts, R3 and R4 up to the DEEP
    codeword."""
    Bf, g = opts.blowup, opts.coset_offset
    N = n * Bf
    aux, L, beta, z_logup, alpha = ctx["aux"], ctx["L"], ctx["beta"], ctx["z_logup"], ctx["alpha"]
    Pn = num_parts(air)
    parts = B.composition_commit_parts(H, N, n, Pn, g)
    H.free()
    t.append_bytes(parts.root)
    # ---- R3: out-of-domain evaluations ----
    z = sample_z_ood(t, n, N, g)
    K = frame_offsets(air)
    wn = F.root_of_unity(n.bit_length() - 1)
    points, cur = [], list(z)
    for _ in range(K):
        points.append(cur)
        cur = F.f3_scale(cur, wn)
    pts = np.array(points, dtype=np.uint64)
    ood = B.barycentric_base_multi(main, g, pts)  # (K, m, 3)
    if aux is not None:
        ood = np.concatenate([ood, fp3_columns(B.barycentric_base_multi(aux, g, pts))], axis=1)
    W = ood.shape[1]
    next_cols = sorted(air.next_row_cols())
    t.append_fp3_many(ood[0])                                  # current-row block
    # pruned next-row block, column by column: (C, K-1, 3)
    block1 = np.ascontiguousarray(ood[1:, next_cols].transpose(1, 0, 2)).reshape(len(next_cols), K - 1, 3)
    t.append_fp3_many(block1.reshape(-1, 3))
    zP = F.f3_pow(z, Pn)
    parts_ood = fp3_columns(B.barycentric_base_multi(parts, g, np.array([zP], dtype=np.uint64)))[0]  # (P, 3)
    t.append_fp3_many(parts_ood)
    # ---- R4: DEEP, FRI, grinding, queries ----
    gamma = [int(v) for v in t.sample_fp3()]
    expo = trace_term_gamma_exponents(W, K, next_cols)
    S = len(expo)
    gp = B.fp3_powers(gamma, S + Pn)
    # the rectangular (column, frame row) grid; pruned openings weigh zero
    gammas_tr = np.zeros((W * K, 3), dtype=np.uint64)
    trace_ood = np.zeros((W * K, 3), dtype=np.uint64)
    for j in range(W):
        for k in range(K):
            if (j, k) in expo:
                gammas_tr[j * K + k] = gp[expo[(j, k)]]
            trace_ood[j * K + k] = ood[k, j]
    deep = B.deep_composition(main, g, pts, trace_ood, gammas_tr, aux=aux, parts=parts, z_parts=zP,
                              parts_ood=parts_ood, gammas_h=gp[S:])
    return _TableState(n=n, t=t, main=main, aux=aux, parts=parts, deep=deep, pre=pre, L=L, ood=ood, block1=block1,
                       parts_ood=parts_ood,
                       challenges=dict(z_logup=z_logup, alpha=alpha, beta=beta, z=z, gamma=gamma))


def _finish_table(st: _TableState, opts: ProofOptions, fri, fri_roots, fri_coeffs) -> AirTableProof:
    """After the FRI commit phase: grinding, queries, FRI decommitment and
    the row-pair openings of the main, aux and parts trees."""
    t, n, main, aux, parts = st.t, st.n, st.main, st.aux, st.parts
    N = n * opts.blowup
    nonce = None
    if opts.grinding_factor > 0:
        nonce = B.generate_nonce(t.state(), opts.grinding_factor)
        t.append_bytes(int(nonce).to_bytes(8, "big"))
    iotas = np.array([t.sample_u64(N >> 1) for _ in range(opts.num_queries)], dtype=np.uint64)
    fri_sym, fri_paths = fri.query(iotas)
    fri.close()
    main_paths, main_rows = main.open(iotas)
    pre_root = pre_rows = pre_paths = None
    if st.pre is not None:
        # whole rows come back: the fixed columns open against their own tree
        pre_root, pre_nodes, k = st.pre
        pre_rows, main_rows = main_rows[:, :, :k], main_rows[:, :, k:]
        pre_paths = B.merkle_paths_from_nodes(pre_nodes, iotas)
    aux_paths = aux_rows = None
    if aux is not None:
        aux_paths, aux_rows = aux.open(iotas)
        aux.close()
    parts_paths, parts_rows = parts.open(iotas)
    parts.close()
    canon = lambda a: None if a is None else np.where(a >= np.uint64(P), a - np.uint64(P), a)  # noqa: E731
    return AirTableProof(
        trace_rows=n, main_root=main.root, aux_root=None if aux is None else bytes(aux.root),
        table_contribution=st.L, composition_root=parts.root,
        ood_block0=canon(st.ood[0].copy()), ood_block1=canon(st.block1), parts_ood=canon(st.parts_ood),
        fri_roots=fri_roots, fri_coeffs=canon(fri_coeffs), nonce=nonce, iotas=iotas, fri_sym=canon(fri_sym),
        fri_paths=fri_paths, main_rows=canon(main_rows), main_paths=main_paths, aux_rows=canon(aux_rows),
        aux_paths=aux_paths, parts_rows=canon(parts_rows), parts_paths=parts_paths,
        transcript_state=t.state(), challenges=st.challenges, precomputed_root=pre_root,
        precomputed_rows=canon(pre_rows), precomputed_paths=pre_paths)


def air_descriptor(air: AirProgram):
    from .air import descriptor
    return descriptor(air)
