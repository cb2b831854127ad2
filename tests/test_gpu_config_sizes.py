"""Parity at the full sizes of the benchmark workloads: the HIP path against
the oracle on the same seeded inputs (canonical values; roots byte for byte,
the comparison rule of crypto/math-cuda/tests/lde.rs:54-85 and
merkle_root_parity.rs:1-10).

  - single table 2^22 x 64, blowup 4, offset 7: commitment, OOD at z and z*w
    (barycentric), DEEP, FRI (fold 2) to a 2^8 terminal, 30 queries with
    their FRI and trace openings;
  - constraint + LogUp stage: 2^21 x 80 FibonacciMultiColumn, 2^20 x 48
    Quadratic, 2^18 x 24 four-bus LogUp table, blowup 4, offset 3: main roots,
    LogUp aux trace and contribution, aux root, every composition value and
    the degree-2 parts root of each table;
  - batched ragged commitment: the 14 seeded tables of the default benchmark
    (log2 rows 12..22, 8..120 columns, blowup 2 or 4), one batched plan
    replayed as a HIP graph, every root against the oracle's commitment;
  - proving service: a synthetic job of the serving workload (2^24..2^27 LDE
    cells) through the service; its roots and every table's rounds (digest of
    every field a verifier reads) against the oracle's replay.
Each case prints its name and shape before it touches the GPU."""
import random

import numpy as np
import pytest

from lambda_vm_amd import workload

pytestmark = [pytest.mark.gpu, pytest.mark.slow, pytest.mark.timeout(600)]
P = workload.P


def canon(a):
    a = np.asarray(a, dtype=np.uint64)
    return np.where(a >= np.uint64(P), a - np.uint64(P), a)


def _say(msg):
    print(f"[config-size] {msg}", flush=True)


def _bitrev(bits):
    idx = np.arange(1 << bits, dtype=np.uint64)
    out = np.zeros_like(idx)
    for b in range(bits):
        out |= ((idx >> np.uint64(b)) & np.uint64(1)) << np.uint64(bits - 1 - b)
    return out.astype(np.int64)


def test_single_table_2p22x64_full_rounds(gpu, ora):
    from lambda_vm_amd import pipeline
    from oracle.pipeline import replay_single_table

    log_n, m, B, g, k, q = 22, 64, 4, 7, 6, 30
    n, N = 1 << log_n, (1 << log_n) * B
    _say(f"single table 2^{log_n} x {m}, blowup {B}, offset {g}, terminal 2^{2 + k}, {q} queries")
    tr = workload.trace_table(1, 0, n, m)
    d_tr = gpu.DeviceBuffer.from_array(tr)
    d_w = gpu.coset_weights_dev(n, g)
    proof = pipeline.prove_single_table(d_tr, n, m, B, g, d_w, fri_final_log_degree=k, num_queries=q,
                                        label=b"config-size-single")
    d_tr.free()
    d_w.free()
    ref = replay_single_table(tr, B, g, k, q, b"config-size-single")
    del tr
    assert proof.root == ref["root"]
    assert proof.z == ref["z"]
    assert np.array_equal(canon(proof.ood), canon(ref["ood"]))
    assert np.array_equal(proof.fri_roots, ref["roots"])
    # folds from 2^24 down to the 2^(2+k) terminal; the last fold is not committed
    assert len(ref["roots"]) == 24 - (2 + k) - 1
    assert np.array_equal(canon(proof.fri_coeffs), canon(ref["coeffs"]))
    assert proof.iotas.tolist() == ref["iotas"]
    assert proof.transcript_state == ref["state"]
    bits = N.bit_length() - 1
    for qi, p in enumerate(ref["iotas"]):
        leaf = bytes(ref["nodes"][N // 2 - 1 + p])
        assert ora.merkle_verify(ref["root"], leaf, p, proof.trace_paths[qi])
        r0 = int(format(2 * p, f"0{bits}b")[::-1], 2)
        assert np.array_equal(canon(proof.trace_rows[qi, 0]), canon(ref["lde"][r0]))
    off = 0
    for li in range(len(ref["roots"])):
        size = N >> (li + 1)
        ev = ref["layers"][off:off + size]
        off += size
        for qi, iota in enumerate(ref["iotas"]):
            assert np.array_equal(canon(proof.fri_sym[li, qi]), canon(ev[(iota >> li) ^ 1]))


def test_constraint_logup_stage_three_tables(gpu, ora):
    from lambda_vm_amd import air as A
    from lambda_vm_amd import field as Fh

    g, B = workload.PROOF_COSET_OFFSET, 4
    specs = [("fib", 1 << 21, 80), ("quad", 1 << 20, 48), ("logup", 1 << 18, 24)]
    rng = random.Random(2)

    def r3():
        return (rng.randrange(P), rng.randrange(P), rng.randrange(P))

    airs = [A.fibonacci_multi_column(80), A.quadratic_multi(48), A.logup_air(24, A.four_bus_interactions())]
    airs[0].boundary = [A.Boundary(c, False, s, (c + s, 0, 0)) for c in range(0, 80, 9) for s in (0, 1)]
    airs[1].boundary = [A.Boundary(c, False, 0, (7, 0, 0)) for c in range(0, 48, 5)]
    desc = A.descriptor(airs[2])
    z, al = r3(), r3()
    ap, cur = [], (1, 0, 0)
    for _ in range(desc.alpha_powers_len):
        ap.append(cur)
        cur = tuple(Fh.f3_mul(list(cur), list(al)))
    ap = np.array(ap, dtype=np.uint64)
    # LogUp stage first (its LDE is the smallest): aux trace, contribution, aux root
    kind, n, m = specs[2]
    _say(f"LogUp aux {n} x {m}, {len(airs[2].interactions)} interactions")
    l_tr = gpu.splitmix_table_dev(n, m, 2, 2)
    l_host = workload.trace_table_modp(2, 2, n, m)
    aux_dev, L, _ = gpu.logup_aux_dev(l_tr, n, m, desc, ap, z)
    ref_aux, ref_L, _ = ora.logup_aux(desc, l_host, ap, np.array(z, dtype=np.uint64))
    ref_aux = ref_aux.reshape(n, -1)
    assert np.array_equal(canon(aux_dev.to_array(np.uint64, ref_aux.shape)), canon(ref_aux))
    assert tuple(int(x) % P for x in L) == tuple(int(x) % P for x in ref_L)
    toff = np.array([int(x) * pow(n, P - 2, P) % P for x in ref_L], dtype=np.uint64)
    w_l = gpu.coset_weights_dev(n, g)
    aux_h = gpu.coset_lde_dev(aux_dev, n, ref_aux.shape[1], B, w_l)
    ref_aux_lde, ref_aux_nodes = ora.lde_commit(ref_aux, B, g)
    assert aux_h.root == bytes(ref_aux_nodes[0])
    aux_dev.free()
    for i, (kind, n, m) in enumerate(specs):
        N = n * B
        _say(f"composition {kind} {n} x {m}, blowup {B}: main root, H, degree-2 parts root")
        d_tr = gpu.splitmix_table_dev(n, m, 2, i) if i < 2 else l_tr
        host_tr = workload.trace_table_modp(2, i, n, m) if i < 2 else l_host
        w = gpu.coset_weights_dev(n, g) if i < 2 else w_l
        main_h = gpu.coset_lde_dev(d_tr, n, m, B, w)
        ref_lde, ref_nodes = ora.lde_commit(host_tr, B, g)
        assert main_h.root == bytes(ref_nodes[0]), kind
        a = airs[i]
        groups, ex = a.groups()
        betas = [r3() for _ in range(a.num_constraints)]
        bnd = [(b.col, b.is_aux, b.step, b.value) for b in a.boundary]
        bb = [r3() for _ in bnd]
        extra = dict(rap=[z, al], alpha=ap, table_offset=toff) if kind == "logup" else {}
        aux = aux_h if kind == "logup" else None
        Hb = gpu.eval_composition(a.device, main_h, aux, B, n, g, betas, groups, ex, bnd, bb, **extra)
        ref_H = ora.eval_composition(a.device, ref_lde, ref_aux_lde if kind == "logup" else None, B, n, g, betas,
                                     groups, ex, bnd, bb, extra.get("rap"), extra.get("alpha"),
                                     extra.get("table_offset"))
        del ref_lde
        got_H = Hb.to_array(np.uint64, (N, 3))
        assert np.array_equal(canon(got_H), canon(ref_H)[_bitrev(N.bit_length() - 1)]), kind
        parts = gpu.composition_commit_d2(Hb, N, n, g)
        _, ref_parts_nodes = ora.lde_commit(ora.decompose_d2(ref_H, g), 2, pow(g, P - 2, P), want_lde=False)
        assert parts.root == bytes(ref_parts_nodes[0]), kind
        parts.close()
        Hb.free()
        main_h.close()
        d_tr.free()
        if i < 2:
            w.free()
    aux_h.close()
    w_l.free()


def test_batched_ragged_14_tables(gpu, ora):
    shapes = workload.ragged_shapes(cfg=3, count=14)
    _say(f"batched ragged commitment of {len(shapes)} tables {shapes}")
    tables, hosts = [], []
    for t, (n, m, b) in enumerate(shapes):
        tr = workload.trace_table(3, t, n, m)
        hosts.append(tr)
        tables.append((gpu.DeviceBuffer.from_array(tr), gpu.coset_weights_dev(n, workload.GENERATOR), n, m, b))
    plan = gpu.BatchPlan(tables, use_graph=True)
    roots = plan.run()
    again = plan.run()  # a second graph replay gives the same roots
    plan.close()
    assert roots == again
    for t, (n, m, b) in enumerate(shapes):
        _, nodes = ora.lde_commit(hosts[t], b, workload.GENERATOR, want_lde=False)
        assert roots[t] == bytes(nodes[0]), (t, n, m, b)
    for tr, w, *_ in tables:
        tr.free()
        w.free()


def _serving_job():
    """The largest job of the serving workload's arrivals (seed 11, as the
    benchmark draws them) with 2^24..2^27 LDE cells."""
    from lambda_vm_amd import serve
    jobs = [spec for _, spec in serve.synth_jobs(20.0, 120, seed=11)]
    fit = [j for j in jobs if (1 << 24) <= j.lde_cells <= (1 << 27)]
    return max(fit, key=lambda j: j.lde_cells)


def test_serving_job_matches_oracle(gpu, ora):
    from lambda_vm_amd import pipeline, serve
    from oracle.pipeline import replay_committed_table

    spec = _serving_job()
    offset, k, q = 3, 6, 30
    _say(f"serving job {spec.job_id}: {len(spec.tables)} tables, {spec.lde_cells} LDE cells, {spec.tables}")
    ex = serve.GpuExecutor(offset=offset, fri_k=k, num_queries=q, workers=4)
    svc = serve.ProvingService(1 << 40, ex)
    svc.run_async([(0.0, spec)])
    ex.close()
    js = svc.done[0]
    # the oracle: commit every table, absorb every root, fork per table, replay its rounds
    commits = []
    for ti, (n, m) in enumerate(spec.tables):
        tr = workload.trace_table_modp(spec.cfg, spec.job_id * 64 + ti, n, m)
        commits.append(ora.lde_commit(tr, spec.blowup, offset))
        assert js.roots[ti] == bytes(commits[-1][1][0]), ti
    t = ora.Transcript(b"lambda-vm-job-%d" % spec.job_id)
    for _, nodes in commits:
        t.append_bytes(bytes(nodes[0]))
    for ti, (n, m) in enumerate(spec.tables):
        f = t.clone()
        if len(spec.tables) > 1:
            f.append_bytes(int(ti).to_bytes(8, "little"))
        lde, nodes = commits[ti]
        ref = replay_committed_table(lde, nodes, f, n, spec.blowup, offset, k, q)
        N = n * spec.blowup
        bits = N.bit_length() - 1
        paths, rows = [], []
        for p in ref["iotas"]:
            paths.append(ora.merkle_path(nodes, p))
            rows.append([lde[int(format(2 * p, f"0{bits}b")[::-1], 2)], lde[int(format(2 * p + 1, f"0{bits}b")[::-1], 2)]])
        fri_sym, fri_paths, off = [], [], 0
        for li in range(len(ref["roots"])):
            size = N >> (li + 1)
            ev = ref["layers"][off:off + size]
            off += size
            lnodes = ora.merkle_build(ora.fri_layer_leaves(ev))
            fri_sym.append([ev[(i >> li) ^ 1] for i in ref["iotas"]])
            fri_paths.append(np.stack([ora.merkle_path(lnodes, (i >> li) >> 1) for i in ref["iotas"]]))
        proof = pipeline.TableProof(root=ref["root"], z=ref["z"],
                                    ood=ref["ood"], fri_roots=np.asarray(ref["roots"]),
                                    fri_coeffs=np.asarray(ref["coeffs"]),
                                    iotas=np.array(ref["iotas"], dtype=np.uint64),
                                    fri_sym=np.array(fri_sym, dtype=np.uint64).reshape(len(ref["roots"]), q, 3),
                                    fri_paths=fri_paths, trace_paths=np.stack(paths),
                                    trace_rows=np.array(rows, dtype=np.uint64), transcript_state=ref["state"])
        assert js.digests[ti] == serve.proof_digest(proof), ti
