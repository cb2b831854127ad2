"""AIR-table proofs on the GPU (lambda_vm_amd.prover) against the oracle's
replay of the same rounds on the same traces: every committed root, OOD
value, FRI layer root and coefficient, query index, opened row and path is
equal (canonical limbs), the canonical proof bytes are identical, and the
verifier restatement accepts the GPU proofs (also after a round trip through
the bytes).
Tables: FibonacciMultiColumn (one composition part), Quadratic, a 4-bus
LogUp table and the VM's EQ, BYTEWISE, IS_WH, LT, STORE, BRANCH, MEMW_REGISTER
and LOAD tables (two parts), the preprocessed KECCAK_RC, BITWISE, REGISTER,
PAGE and GLOBAL_MEMORY tables (fixed columns under their own tree), HINT (27
interactions), a one-interaction bus table (degree 2: one
part) and a degree-4 bus table with and without interactions (three parts),
alone and together in one transcript. Bus tables open only their
accumulated column at the next row."""
import numpy as np
import pytest

from lambda_vm_amd import air as A
from lambda_vm_amd import workload
from oracle.air_proof import replay_air_tables, verify_air_tables

pytestmark = pytest.mark.gpu
P = workload.P


def canon(a):
    a = np.asarray(a, dtype=np.uint64)
    return np.where(a >= np.uint64(P), a - np.uint64(P), a)


def _tables(n, which, seed):
    rng = np.random.default_rng(seed)
    out = []
    for kind in which:
        if kind == "fib":
            a = A.fibonacci_multi_column(6)
            tr = A.fibonacci_trace(n, 6, rng.integers(0, 2**62, (6, 2)).astype(np.uint64))
            A.with_boundaries_from_trace(a, tr)
        elif kind == "quad":
            a = A.quadratic_multi(5)
            tr = A.quadratic_trace(n, 5, rng.integers(0, 2**62, 5).astype(np.uint64))
            A.with_boundaries_from_trace(a, tr)
        elif kind == "eq":
            from lambda_vm_amd import vm_tables as V
            a = V.eq_table_air()
            a.bus_descriptor = A.descriptor(a)
            ops = [V.EqOperation(int(x), int(x) if i % 3 == 0 else int(y), bool(i % 2))
                   for i, (x, y) in enumerate(rng.integers(0, 2**63, (n - 3, 2)))]
            tr = V.generate_eq_trace(ops)
        elif kind == "bytewise":
            from lambda_vm_amd import vm_tables as V
            a = V.bytewise_table_air()
            a.bus_descriptor = A.descriptor(a)
            ops = [V.BytewiseOperation(int(x), int(y), int(o)) for x, y, o in
                   zip(rng.integers(0, 2**63, n - 3), rng.integers(0, 2**63, n - 3), rng.integers(0, 3, n - 3))]
            tr = V.generate_bytewise_trace(ops)
        elif kind == "lt":
            from lambda_vm_amd import vm_tables as V
            a = V.lt_table_air()
            a.bus_descriptor = A.descriptor(a)
            xs = rng.integers(0, 2**63, (n - 3, 2))
            ops = [V.LtOperation(int(x) * 2 + (i & 1), int(y) * 2 + 1, bool(i % 2), bool(i % 3 == 0))
                   for i, (x, y) in enumerate(xs)]
            tr = V.generate_lt_trace(ops)
        elif kind == "store":
            from lambda_vm_amd import vm_tables as V
            a = V.store_table_air()
            a.bus_descriptor = A.descriptor(a)
            ops = [V.StoreOperation.of_width(int(x), int(t), int(v), (1, 2, 4, 8)[i % 4])
                   for i, (x, t, v) in enumerate(rng.integers(0, 2**63, (n - 3, 3)))]
            tr = V.generate_store_trace(ops)
        elif kind == "branch":
            from lambda_vm_amd import vm_tables as V
            a = V.branch_table_air()
            a.bus_descriptor = A.descriptor(a)
            ops = [V.BranchOperation(int(x) * 4, int(o) & (2**64 - 1), int(r), bool(i % 3 == 0))
                   for i, (x, o, r) in enumerate(zip(rng.integers(0, 2**61, n - 3), rng.integers(-2**20, 2**20, n - 3),
                                                     rng.integers(0, 2**63, n - 3)))]
            tr = V.generate_branch_trace(ops)
        elif kind == "memw_register":
            from lambda_vm_amd import vm_tables as V
            a = V.memw_register_table_air()
            a.bus_descriptor = A.descriptor(a)
            acc, ts = [], 5 << 32
            for i, (r, dt, v) in enumerate(zip(rng.integers(0, 32, n - 3), rng.integers(1, 1 << 16, n - 3),
                                               rng.integers(0, 2**63, n - 3))):
                acc.append(V.RegisterAccess(2 * int(r), ts + int(dt), int(v), int(v) ^ (i & 7), ts, bool(i % 2)))
                ts += int(dt)
            tr = V.generate_memw_register_trace(acc)
        elif kind == "load":
            from lambda_vm_amd import vm_tables as V
            a = V.load_table_air()
            a.bus_descriptor = A.descriptor(a)
            ops = [V.LoadOperation.of_memory(int(x), 10 + i, (1, 2, 4, 8)[i % 4], bool(i % 3) and i % 4 != 3, int(v))
                   for i, (x, v) in enumerate(zip(rng.integers(0, 2**40, n - 3),
                                                  rng.integers(0, 2**64, n - 3, dtype=np.uint64)))]
            tr = V.generate_load_trace(ops)
        elif kind == "is_wh":
            from lambda_vm_amd import vm_tables as V
            a = V.is_wh_table_air()
            a.bus_descriptor = A.descriptor(a)
            keys = [V.is_wh_key(int(v) << 16) for v in rng.integers(0, 2**47, n - 3)]
            tr = V.generate_is_wh_trace(V.dedup_is_wh_rows(keys))
        elif kind == "bus_deg2":
            a = A.logup_air(24, A.four_bus_interactions()[:1])
            a.bus_descriptor = A.descriptor(a)
            tr = workload.trace_table(8, 3, n, 24)
        elif kind == "keccak_rc":
            from lambda_vm_amd import vm_tables as V
            a = V.keccak_rc_table_air()
            a.bus_descriptor = A.descriptor(a)
            tr = V.generate_keccak_rc_trace(num_keccak_ops=int(rng.integers(1, 1000)))
        elif kind == "bitwise":
            from lambda_vm_amd import vm_tables as V
            a = V.bitwise_table_air()
            a.bus_descriptor = A.descriptor(a)
            kinds = list(V.BITWISE_MU_COLUMN)
            tr = V.generate_bitwise_trace([(kinds[int(k)], int(x), int(y), int(z)) for k, x, y, z in zip(
                rng.integers(0, len(kinds), 5000), rng.integers(0, 256, 5000), rng.integers(0, 256, 5000),
                rng.integers(0, 16, 5000))])
        elif kind in ("product", "product_nobus"):
            a = A.product_bus_air(A.four_bus_interactions() if kind == "product" else [])
            a.bus_descriptor = A.descriptor(a) if a.interactions else None
            tr = A.product_trace(n, 4 if kind == "product" else 5)
        else:
            a = A.logup_air(24, A.four_bus_interactions())
            a.bus_descriptor = A.descriptor(a)
            tr = workload.trace_table(8, 2, n, 24)
        out.append((a, np.ascontiguousarray(tr, dtype=np.uint64)))
    return out


def _f3rows(x):
    """Fp3 values as canonical (k, 3) rows (k = 0 for an empty block)."""
    if not isinstance(x, np.ndarray):
        x = np.array([[int(c) for c in v] for v in x], dtype=np.uint64)
    return canon(x).reshape(-1, 3)


@pytest.mark.parametrize("which,n,blowup,k,q,grind", [(("fib",), 256, 4, 2, 8, 0), (("quad",), 128, 8, 3, 6, 0),
                                                     (("logup",), 256, 4, 2, 8, 0),
                                                     (("eq",), 256, 4, 2, 8, 0), (("eq", "fib"), 1024, 8, 2, 12, 2),
                                                     (("bytewise", "eq"), 512, 4, 3, 10, 0),
                                                     (("is_wh", "bytewise", "eq"), 256, 4, 2, 8, 0),
                                                     (("lt", "eq"), 256, 4, 2, 8, 2),
                                                     (("store", "lt"), 512, 4, 2, 8, 0),
                                                     (("branch", "store", "eq"), 256, 4, 2, 8, 2),
                                                     (("memw_register", "branch"), 512, 4, 3, 8, 0),
                                                     (("load", "memw_register"), 256, 4, 2, 8, 2),
                                                     (("fib", "quad", "logup"), 512, 4, 3, 10, 4),
                                                     (("bus_deg2", "product"), 256, 4, 2, 8, 0),
                                                     (("product_nobus", "fib"), 512, 8, 2, 8, 2),
                                                     (("keccak_rc",), 32, 4, 2, 8, 0),
                                                     (("keccak_rc", "eq", "fib"), 64, 8, 2, 8, 2)])
def test_gpu_proofs_match_oracle_and_verify(gpu, which, n, blowup, k, q, grind):
    from lambda_vm_amd import prover
    tables = _tables(n, which, seed=n + len(which))
    opts = prover.ProofOptions(blowup=blowup, coset_offset=3, fri_final_log_degree=k, num_queries=q,
                               grinding_factor=grind)
    dev = [(a, gpu.DeviceBuffer.from_array(tr), tr.shape[0]) for a, tr in tables]
    got = prover.prove_air_tables(dev, opts, label=b"gpu-air")
    for _, d, _ in dev:
        d.free()
    ref = replay_air_tables(tables, blowup, 3, k, q, grind, b"gpu-air")
    for t, (g, r) in enumerate(zip(got, ref)):
        assert g.main_root == r["main_root"], t
        assert g.aux_root == r["aux_root"], t
        if r["table_contribution"] is not None:
            assert tuple(g.table_contribution) == tuple(r["table_contribution"])
        assert g.composition_root == r["composition_root"], t
        assert np.array_equal(_f3rows(g.ood_block0), _f3rows(r["ood_block0"])), t
        assert np.array_equal(_f3rows(g.ood_block1.reshape(-1, 3)),
                              _f3rows([v for col in r["ood_block1"] for v in col])), t
        assert np.array_equal(_f3rows(g.parts_ood), _f3rows(r["parts_ood"])), t
        assert [bytes(x) for x in g.fri_roots] == r["fri_roots"], t
        assert np.array_equal(_f3rows(g.fri_coeffs), _f3rows(r["fri_coeffs"])), t
        assert g.nonce == r["nonce"]
        assert g.iotas.tolist() == r["iotas"], t
        assert np.array_equal(canon(g.main_rows), canon(r["main_rows"])), t
        assert np.array_equal(canon(g.parts_rows), canon(r["parts_rows"])), t
        if r["aux_rows"] is not None:
            assert np.array_equal(canon(g.aux_rows), canon(r["aux_rows"])), t
        for li in range(len(r["fri_roots"])):
            assert np.array_equal(_f3rows(g.fri_sym[li]), _f3rows(r["fri_sym"][li])), (t, li)
            for qi in range(q):
                assert np.array_equal(g.fri_paths[li][qi], r["fri_paths"][li][qi]), (t, li, qi)
        for qi in range(q):
            assert np.array_equal(g.main_paths[qi], r["main_paths"][qi])
        assert g.precomputed_root == r["precomputed_root"], t
        if r["precomputed_root"] is not None:
            assert np.array_equal(canon(g.precomputed_rows), canon(r["precomputed_rows"])), t
            for qi in range(q):
                assert np.array_equal(g.precomputed_paths[qi], r["precomputed_paths"][qi]), (t, qi)
        assert g.transcript_state == r["transcript_state"], t
    ok, why = verify_air_tables([a for a, _ in tables], got, blowup, 3, k, q, grind, b"gpu-air")
    assert ok, why
    # the canonical proof bytes of the GPU prover and of the replay are identical
    from lambda_vm_amd import proof_io
    from lambda_vm_amd.prover import AirTableProof
    data = proof_io.serialize_multi_proof(got)
    assert data == proof_io.serialize_multi_proof([AirTableProof(**r) for r in ref])
    assert verify_air_tables([a for a, _ in tables], proof_io.deserialize_multi_proof(data), blowup, 3, k, q, grind,
                             b"gpu-air")[0]


def test_table_scheduling_does_not_change_proofs(gpu):
    """Heaviest-first driver threads under a device-byte budget (one table at a
    time when the budget is tiny) give the same proof bytes as unlimited."""
    from lambda_vm_amd import proof_io, prover
    tables = _tables(256, ("fib", "quad", "eq", "bytewise"), seed=41)
    out = []
    # (unlimited: the FRI layers of all tables fold in shared launches; a tiny
    # budget: one table at a time, each folding on its own; per-table FRI forced)
    for workers, budget, batch_fri in ((None, None, True), (2, 1, True), (1, None, True), (None, None, False)):
        opts = prover.ProofOptions(blowup=4, coset_offset=3, fri_final_log_degree=2, num_queries=8,
                                   vram_budget=budget, batch_fri=batch_fri)
        dev = [(a, gpu.DeviceBuffer.from_array(tr), 256) for a, tr in tables]
        out.append(proof_io.serialize_multi_proof(prover.prove_air_tables(dev, opts, label=b"sched",
                                                                          workers=workers)))
        for _, d, _ in dev:
            d.free()
    assert out[0] == out[1] == out[2] == out[3]


@pytest.mark.parametrize("heights", [(1 << 8, 1 << 8), (1 << 10, 1 << 6, 1 << 8), (1 << 12, 1 << 4)])
def test_batched_fri_commit_equals_per_table(gpu, heights):
    """lvm_fri_commit_phase_batch over codewords of different lengths (tables
    that stop committing at different depths) gives every table the roots,
    terminal coefficients and transcript of its own commit phase."""
    rng = np.random.default_rng(len(heights) * 7 + heights[0])
    k = 2
    cws = [rng.integers(0, P, size=(4 * h, 3), dtype=np.uint64) for h in heights]
    labels = [b"fri-%d" % i for i in range(len(heights))]
    ref = []
    for cw, lab in zip(cws, labels):
        t = gpu.Transcript(lab)
        f = gpu.GpuFri(cw, 3)
        roots, coeffs = f.commit_phase(t, 2, k)
        ref.append((roots, coeffs, t.state()))
        f.close()
    ts = [gpu.Transcript(lab) for lab in labels]
    fs = [gpu.GpuFri(cw, 3) for cw in cws]
    got = gpu.fri_commit_phase_batch(fs, ts, [2] * len(fs), k)
    for (r_roots, r_coeffs, r_state), (g_roots, g_coeffs), t, f in zip(ref, got, ts, fs):
        assert np.array_equal(g_roots, r_roots)
        assert np.array_equal(canon(g_coeffs), canon(r_coeffs))
        assert t.state() == r_state
        f.close()


@pytest.mark.slow
def test_gpu_bitwise_table_proof(gpu):
    """The 2^20-row BITWISE table (eleven fixed columns, ten receivers) beside
    EQ: the GPU proof equals the oracle replay, its precomputed root is the
    reference's static commitment for blowup 2, and it verifies."""
    import json
    import os

    from lambda_vm_amd import prover
    print("[gpu-air] BITWISE 2^20 x 21 + EQ 256, blowup 2", flush=True)
    tables = _tables(256, ("bitwise", "eq"), seed=9)
    opts = prover.ProofOptions(blowup=2, coset_offset=3, fri_final_log_degree=4, num_queries=6)
    dev = [(a, gpu.DeviceBuffer.from_array(tr), tr.shape[0]) for a, tr in tables]
    got = prover.prove_air_tables(dev, opts, label=b"gpu-bitwise")
    for _, d, _ in dev:
        d.free()
    ref = replay_air_tables(tables, 2, 3, 4, 6, 0, b"gpu-bitwise")
    kats = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kats.json")))
    assert got[0].precomputed_root == bytes(kats["preprocessed_commitments"]["bitwise"]["2"])
    from lambda_vm_amd import proof_io
    from lambda_vm_amd.prover import AirTableProof
    assert proof_io.serialize_multi_proof(got) == proof_io.serialize_multi_proof([AirTableProof(**r) for r in ref])
    assert verify_air_tables([a for a, _ in tables], got, 2, 3, 4, 6, 0, b"gpu-bitwise")[0]


def test_gpu_register_table_proof(gpu):
    """REGISTER (four fixed columns, its final-state FINI column and the
    Memory-bus sender pair) with the MEMW_REGISTER chain whose tokens it
    closes, in one transcript: the GPU proofs equal the oracle replay byte
    for byte, REGISTER's precomputed root is the verifier's own commitment and
    both verify."""
    from lambda_vm_amd import proof_io, prover
    from lambda_vm_amd import vm_tables as V
    from lambda_vm_amd.prover import AirTableProof
    from oracle.air_proof import precomputed_commitment
    from tests.test_vm_table_register import register_chain

    print("[gpu-air] REGISTER 128 x 6 + MEMW_REGISTER, blowup 4", flush=True)
    acc, final, init = register_chain(200, 11)
    reg = V.register_table_air(init, V.fini_from_final_state(final, init), final_pc_ts=1)
    reg.bus_descriptor = A.descriptor(reg)
    mem = V.memw_register_table_air()
    mem.bus_descriptor = A.descriptor(mem)
    tables = [(reg, V.generate_register_trace(final, init)), (mem, V.generate_memw_register_trace(acc))]
    tables = [(a, np.ascontiguousarray(tr, dtype=np.uint64)) for a, tr in tables]
    opts = prover.ProofOptions(blowup=4, coset_offset=3, fri_final_log_degree=2, num_queries=8)
    dev = [(a, gpu.DeviceBuffer.from_array(tr), tr.shape[0]) for a, tr in tables]
    got = prover.prove_air_tables(dev, opts, label=b"gpu-register")
    for _, d, _ in dev:
        d.free()
    ref = replay_air_tables(tables, 4, 3, 2, 8, 0, b"gpu-register")
    assert got[0].precomputed_root == precomputed_commitment(reg, 4, 3)
    assert proof_io.serialize_multi_proof(got) == proof_io.serialize_multi_proof([AirTableProof(**r) for r in ref])
    assert verify_air_tables([reg, mem], got, 4, 3, 2, 8, 0, b"gpu-register")[0]


def test_gpu_page_tables_proof(gpu):
    """A zero-initialised 2^18-byte PAGE with 300 touched bytes, the table of
    accesses that closes its Memory tokens, and a private-input
    GLOBAL_MEMORY page (OFFSET fixed alone) in one transcript at blowup 2:
    the GPU proofs equal the oracle replay byte for byte, the fixed-columns
    roots are the reference's static zero-page and private-page commitments,
    and the proofs verify."""
    import json
    import os

    from lambda_vm_amd import proof_io, prover
    from lambda_vm_amd import vm_tables as V
    from lambda_vm_amd.prover import AirTableProof
    from tests.test_vm_table_page import access_chain_table

    print("[gpu-air] PAGE 2^18 x 5 + accesses + GLOBAL_MEMORY 2^18 x 4, blowup 2", flush=True)
    rng = np.random.default_rng(12)
    base = 0x3_0004_0000
    final = {base + int(o): (int(t), int(v)) for o, t, v in
             zip(rng.choice(V.PAGE_SIZE, 300, replace=False), rng.integers(4, 2**40, 300), rng.integers(0, 256, 300))}
    page = V.page_table_air(base)
    page.bus_descriptor = A.descriptor(page)
    chain_air, chain_tr = access_chain_table(base, b"", final)
    gm = V.global_memory_table_air(base, b"\x01\x02\x03", private=True)
    gm.bus_descriptor = A.descriptor(gm)
    gm_tr = V.generate_global_memory_trace(base, b"\x01\x02\x03", {base + 2: (7, 5)})
    tables = [(page, V.generate_page_trace(base, None, final)), (chain_air, chain_tr), (gm, gm_tr)]
    tables = [(a, np.ascontiguousarray(tr, dtype=np.uint64)) for a, tr in tables]
    opts = prover.ProofOptions(blowup=2, coset_offset=3, fri_final_log_degree=4, num_queries=6)
    dev = [(a, gpu.DeviceBuffer.from_array(tr), tr.shape[0]) for a, tr in tables]
    got = prover.prove_air_tables(dev, opts, label=b"gpu-page")
    for _, d, _ in dev:
        d.free()
    kats = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kats.json")))["preprocessed_commitments"]
    assert got[0].precomputed_root == bytes(kats["page_zero"]["2"])
    assert got[2].precomputed_root == bytes(kats["page_private"]["2"])
    ref = replay_air_tables(tables, 2, 3, 4, 6, 0, b"gpu-page")
    assert proof_io.serialize_multi_proof(got) == proof_io.serialize_multi_proof([AirTableProof(**r) for r in ref])
    assert verify_air_tables([page, chain_air, gm], got, 2, 3, 4, 6, 0, b"gpu-page")[0]


def test_gpu_hint_table_proof(gpu):
    """HINT (41 columns, 27 interactions: 14 aux columns) beside the LT table
    that answers its range checks, with grinding: the GPU proofs equal the
    oracle replay byte for byte and verify."""
    from lambda_vm_amd import proof_io, prover
    from lambda_vm_amd import vm_tables as V
    from lambda_vm_amd.prover import AirTableProof
    from tests.test_vm_table_hint import _lt_ops, hint_ops

    print("[gpu-air] HINT 128 x 41 + LT, blowup 4, grinding 4", flush=True)
    ops = hint_ops(100, 21)
    hint, lt = V.hint_table_air(), V.lt_table_air()
    for a in (hint, lt):
        a.bus_descriptor = A.descriptor(a)
    tables = [(hint, V.generate_hint_trace(ops)), (lt, V.generate_lt_trace(_lt_ops(ops)))]
    tables = [(a, np.ascontiguousarray(tr, dtype=np.uint64)) for a, tr in tables]
    opts = prover.ProofOptions(blowup=4, coset_offset=3, fri_final_log_degree=2, num_queries=8, grinding_factor=4)
    dev = [(a, gpu.DeviceBuffer.from_array(tr), tr.shape[0]) for a, tr in tables]
    got = prover.prove_air_tables(dev, opts, label=b"gpu-hint")
    for _, d, _ in dev:
        d.free()
    ref = replay_air_tables(tables, 4, 3, 2, 8, 4, b"gpu-hint")
    assert proof_io.serialize_multi_proof(got) == proof_io.serialize_multi_proof([AirTableProof(**r) for r in ref])
    assert verify_air_tables([hint, lt], got, 4, 3, 2, 8, 4, b"gpu-hint")[0]


def test_gpu_keccak_table_proof(gpu):
    """KECCAK core (486 columns, 52 constraints, 84 interactions: 42 aux Fp3
    columns) beside the IS_WH table that answers its lane-address range
    checks: the GPU proofs equal the oracle replay byte for byte and verify."""
    from lambda_vm_amd import proof_io, prover
    from lambda_vm_amd import vm_tables as V
    from lambda_vm_amd.prover import AirTableProof
    from tests.test_vm_table_keccak import keccak_ops

    print("[gpu-air] KECCAK 64 x 486 + IS_WH, blowup 4", flush=True)
    ops = keccak_ops(50, 31)
    kec, iswh = V.keccak_table_air(), V.is_wh_table_air()
    for a in (kec, iswh):
        a.bus_descriptor = A.descriptor(a)
    tables = [(kec, V.generate_keccak_trace(ops)),
              (iswh, V.generate_is_wh_trace(V.dedup_is_wh_rows(V.keccak_is_wh_keys(ops))))]
    tables = [(a, np.ascontiguousarray(tr, dtype=np.uint64)) for a, tr in tables]
    opts = prover.ProofOptions(blowup=4, coset_offset=3, fri_final_log_degree=2, num_queries=8)
    dev = [(a, gpu.DeviceBuffer.from_array(tr), tr.shape[0]) for a, tr in tables]
    got = prover.prove_air_tables(dev, opts, label=b"gpu-keccak")
    for _, d, _ in dev:
        d.free()
    ref = replay_air_tables(tables, 4, 3, 2, 8, 0, b"gpu-keccak")
    assert proof_io.serialize_multi_proof(got) == proof_io.serialize_multi_proof([AirTableProof(**r) for r in ref])
    assert verify_air_tables([kec, iswh], got, 4, 3, 2, 8, 0, b"gpu-keccak")[0]


def test_gpu_memw_aligned_page_proof(gpu):
    """MEMW_A (aligned accesses replayed by a memory model) with the PAGE whose
    tokens close its memory chains: the GPU proofs equal the oracle replay byte
    for byte and verify."""
    from lambda_vm_amd import proof_io, prover
    from lambda_vm_amd import vm_tables as V
    from lambda_vm_amd.prover import AirTableProof
    from tests.test_vm_table_memw_aligned import BASE, INIT, PAGE, access_chain

    print("[gpu-air] MEMW_A 256 x 29 + PAGE 1024 x 5, blowup 4", flush=True)
    ops, _, cells = access_chain(200, 41, INIT)
    memw, page = V.memw_aligned_table_air(), V.page_table_air(BASE, INIT, page_size=PAGE)
    for a in (memw, page):
        a.bus_descriptor = A.descriptor(a)
    tables = [(memw, V.generate_memw_aligned_trace(ops)), (page, V.generate_page_trace(BASE, INIT, cells, PAGE))]
    tables = [(a, np.ascontiguousarray(tr, dtype=np.uint64)) for a, tr in tables]
    opts = prover.ProofOptions(blowup=4, coset_offset=3, fri_final_log_degree=2, num_queries=8)
    dev = [(a, gpu.DeviceBuffer.from_array(tr), tr.shape[0]) for a, tr in tables]
    got = prover.prove_air_tables(dev, opts, label=b"gpu-memw-a")
    for _, d, _ in dev:
        d.free()
    ref = replay_air_tables(tables, 4, 3, 2, 8, 0, b"gpu-memw-a")
    assert proof_io.serialize_multi_proof(got) == proof_io.serialize_multi_proof([AirTableProof(**r) for r in ref])
    assert verify_air_tables([memw, page], got, 4, 3, 2, 8, 0, b"gpu-memw-a")[0]


def test_gpu_mul_table_proof(gpu):
    """MUL (26 columns, degree-2 convolution constraints, 25 interactions) beside
    HINT: the GPU proofs equal the oracle replay byte for byte and verify."""
    from lambda_vm_amd import proof_io, prover
    from lambda_vm_amd import vm_tables as V
    from lambda_vm_amd.prover import AirTableProof
    from tests.test_vm_table_hint import hint_ops
    from tests.test_vm_table_mul import mul_ops

    print("[gpu-air] MUL 512 x 26 + HINT, blowup 2", flush=True)
    mul, hint = V.mul_table_air(), V.hint_table_air()
    for a in (mul, hint):
        a.bus_descriptor = A.descriptor(a)
    tables = [(mul, V.generate_mul_trace(mul_ops(400, 51))), (hint, V.generate_hint_trace(hint_ops(9, 52)))]
    tables = [(a, np.ascontiguousarray(tr, dtype=np.uint64)) for a, tr in tables]
    opts = prover.ProofOptions(blowup=2, coset_offset=3, fri_final_log_degree=3, num_queries=10)
    dev = [(a, gpu.DeviceBuffer.from_array(tr), tr.shape[0]) for a, tr in tables]
    got = prover.prove_air_tables(dev, opts, label=b"gpu-mul")
    for _, d, _ in dev:
        d.free()
    ref = replay_air_tables(tables, 2, 3, 3, 10, 0, b"gpu-mul")
    assert proof_io.serialize_multi_proof(got) == proof_io.serialize_multi_proof([AirTableProof(**r) for r in ref])
    assert verify_air_tables([mul, hint], got, 2, 3, 3, 10, 0, b"gpu-mul")[0]


def test_gpu_memw_page_proof(gpu):
    """MEMW (unaligned accesses across a 2^32 low-word boundary) with the two
    PAGE tables that close its memory chains: the GPU proofs equal the oracle
    replay byte for byte and verify."""
    from lambda_vm_amd import proof_io, prover
    from lambda_vm_amd import vm_tables as V
    from lambda_vm_amd.prover import AirTableProof
    from tests.test_vm_table_memw import _pages, unaligned_chain

    print("[gpu-air] MEMW 256 x 49 + 2 PAGE 1024 x 5, blowup 4", flush=True)
    ops, finals = unaligned_chain(200, 61)
    memw = V.memw_table_air()
    memw.bus_descriptor = A.descriptor(memw)
    tables = [(memw, V.generate_memw_trace(ops))] + _pages(finals, memory_only=False)
    tables = [(a, np.ascontiguousarray(tr, dtype=np.uint64)) for a, tr in tables]
    opts = prover.ProofOptions(blowup=4, coset_offset=3, fri_final_log_degree=2, num_queries=8)
    dev = [(a, gpu.DeviceBuffer.from_array(tr), tr.shape[0]) for a, tr in tables]
    got = prover.prove_air_tables(dev, opts, label=b"gpu-memw")
    for _, d, _ in dev:
        d.free()
    ref = replay_air_tables(tables, 4, 3, 2, 8, 0, b"gpu-memw")
    assert proof_io.serialize_multi_proof(got) == proof_io.serialize_multi_proof([AirTableProof(**r) for r in ref])
    assert verify_air_tables([a for a, _ in tables], got, 4, 3, 2, 8, 0, b"gpu-memw")[0]


def test_gpu_shift_table_proof(gpu):
    """SHIFT (29 columns, 22 degree-3 constraints, 17 interactions) beside MUL:
    the GPU proofs equal the oracle replay byte for byte and verify."""
    from lambda_vm_amd import proof_io, prover
    from lambda_vm_amd import vm_tables as V
    from lambda_vm_amd.prover import AirTableProof
    from tests.test_vm_table_mul import mul_ops
    from tests.test_vm_table_shift import shift_ops

    print("[gpu-air] SHIFT 512 x 29 + MUL, blowup 4", flush=True)
    shift, mul = V.shift_table_air(), V.mul_table_air()
    for a in (shift, mul):
        a.bus_descriptor = A.descriptor(a)
    tables = [(shift, V.generate_shift_trace(shift_ops(500, 71))), (mul, V.generate_mul_trace(mul_ops(60, 72)))]
    tables = [(a, np.ascontiguousarray(tr, dtype=np.uint64)) for a, tr in tables]
    opts = prover.ProofOptions(blowup=4, coset_offset=3, fri_final_log_degree=2, num_queries=8)
    dev = [(a, gpu.DeviceBuffer.from_array(tr), tr.shape[0]) for a, tr in tables]
    got = prover.prove_air_tables(dev, opts, label=b"gpu-shift")
    for _, d, _ in dev:
        d.free()
    ref = replay_air_tables(tables, 4, 3, 2, 8, 0, b"gpu-shift")
    assert proof_io.serialize_multi_proof(got) == proof_io.serialize_multi_proof([AirTableProof(**r) for r in ref])
    assert verify_air_tables([shift, mul], got, 4, 3, 2, 8, 0, b"gpu-shift")[0]


def test_gpu_dvrm_table_proof(gpu):
    """DVRM (34 columns, 19 constraints, 35 interactions) beside MUL: the GPU
    proofs equal the oracle replay byte for byte and verify."""
    from lambda_vm_amd import proof_io, prover
    from lambda_vm_amd import vm_tables as V
    from lambda_vm_amd.prover import AirTableProof
    from tests.test_vm_table_dvrm import dvrm_ops
    from tests.test_vm_table_mul import mul_ops

    print("[gpu-air] DVRM 512 x 34 + MUL, blowup 4", flush=True)
    dvrm, mul = V.dvrm_table_air(), V.mul_table_air()
    for a in (dvrm, mul):
        a.bus_descriptor = A.descriptor(a)
    tables = [(dvrm, V.generate_dvrm_trace(dvrm_ops(480, 81))), (mul, V.generate_mul_trace(mul_ops(60, 82)))]
    tables = [(a, np.ascontiguousarray(tr, dtype=np.uint64)) for a, tr in tables]
    opts = prover.ProofOptions(blowup=4, coset_offset=3, fri_final_log_degree=2, num_queries=8)
    dev = [(a, gpu.DeviceBuffer.from_array(tr), tr.shape[0]) for a, tr in tables]
    got = prover.prove_air_tables(dev, opts, label=b"gpu-dvrm")
    for _, d, _ in dev:
        d.free()
    ref = replay_air_tables(tables, 4, 3, 2, 8, 0, b"gpu-dvrm")
    assert proof_io.serialize_multi_proof(got) == proof_io.serialize_multi_proof([AirTableProof(**r) for r in ref])
    assert verify_air_tables([dvrm, mul], got, 4, 3, 2, 8, 0, b"gpu-dvrm")[0]


def test_gpu_keccak_rnd_table_proof(gpu):
    """KECCAK_RND (1480 columns, 1031 interactions, 516 aux Fp3 columns; its
    constraint program exceeds the run-time compiler's node cap and runs on
    the interpreter): the GPU proof of two permutations equals the oracle
    replay byte for byte and verifies."""
    from lambda_vm_amd import proof_io, prover
    from lambda_vm_amd import vm_tables as V
    from lambda_vm_amd.prover import AirTableProof
    from tests.test_vm_table_keccak_rnd import rnd_ops

    print("[gpu-air] KECCAK_RND 64 x 1480, blowup 2", flush=True)
    rnd = V.keccak_rnd_table_air()
    rnd.bus_descriptor = A.descriptor(rnd)
    tables = [(rnd, np.ascontiguousarray(V.generate_keccak_rnd_trace(rnd_ops(2, 91)), dtype=np.uint64))]
    opts = prover.ProofOptions(blowup=2, coset_offset=3, fri_final_log_degree=2, num_queries=6)
    dev = [(a, gpu.DeviceBuffer.from_array(tr), tr.shape[0]) for a, tr in tables]
    got = prover.prove_air_tables(dev, opts, label=b"gpu-keccak-rnd")
    for _, d, _ in dev:
        d.free()
    ref = replay_air_tables(tables, 2, 3, 2, 6, 0, b"gpu-keccak-rnd")
    assert proof_io.serialize_multi_proof(got) == proof_io.serialize_multi_proof([AirTableProof(**r) for r in ref])
    assert verify_air_tables([rnd], got, 2, 3, 2, 6, 0, b"gpu-keccak-rnd")[0]


def test_gpu_cpu32_table_proof(gpu):
    """CPU32 (38 columns, 32 degree-3 constraints, 23 interactions) beside the
    SHIFT chip its word shifts use: the GPU proofs equal the oracle replay byte
    for byte and verify."""
    from lambda_vm_amd import proof_io, prover
    from lambda_vm_amd import vm_tables as V
    from lambda_vm_amd.prover import AirTableProof
    from tests.test_vm_table_cpu32 import chip_ops, cpu32_ops

    print("[gpu-air] CPU32 512 x 38 + SHIFT, blowup 4", flush=True)
    ops, _ = cpu32_ops(500, 101)
    cpu, sh = V.cpu32_table_air(), V.shift_table_air()
    for a in (cpu, sh):
        a.bus_descriptor = A.descriptor(a)
    tables = [(cpu, V.generate_cpu32_trace(ops)), (sh, V.generate_shift_trace(chip_ops(ops)[0]))]
    tables = [(a, np.ascontiguousarray(tr, dtype=np.uint64)) for a, tr in tables]
    opts = prover.ProofOptions(blowup=4, coset_offset=3, fri_final_log_degree=2, num_queries=8)
    dev = [(a, gpu.DeviceBuffer.from_array(tr), tr.shape[0]) for a, tr in tables]
    got = prover.prove_air_tables(dev, opts, label=b"gpu-cpu32")
    for _, d, _ in dev:
        d.free()
    ref = replay_air_tables(tables, 4, 3, 2, 8, 0, b"gpu-cpu32")
    assert proof_io.serialize_multi_proof(got) == proof_io.serialize_multi_proof([AirTableProof(**r) for r in ref])
    assert verify_air_tables([cpu, sh], got, 4, 3, 2, 8, 0, b"gpu-cpu32")[0]


def test_gpu_decode_table_proof(gpu):
    """DECODE (five fixed columns of a decoded RV64IM program, one receiver)
    beside the CPU32 rows of its word instructions: the GPU proofs equal the
    oracle replay byte for byte, DECODE's precomputed root is the verifier's own
    commitment and both verify."""
    from lambda_vm_amd import proof_io, prover
    from lambda_vm_amd import vm_tables as V
    from lambda_vm_amd.prover import AirTableProof
    from oracle.air_proof import precomputed_commitment
    from tests.test_vm_table_decode import word_program

    print("[gpu-air] DECODE 512 x 6 + CPU32, blowup 4", flush=True)
    words, ops = word_program(400, 111)
    entries = V.decode_entries_from_words([(0x2000, words)])
    dec, cpu = V.decode_table_air(entries), V.cpu32_table_air()
    for a in (dec, cpu):
        a.bus_descriptor = A.descriptor(a)
    tables = [(dec, V.generate_decode_trace(entries, [o.pc for o in ops])), (cpu, V.generate_cpu32_trace(ops))]
    tables = [(a, np.ascontiguousarray(tr, dtype=np.uint64)) for a, tr in tables]
    opts = prover.ProofOptions(blowup=4, coset_offset=3, fri_final_log_degree=2, num_queries=8)
    dev = [(a, gpu.DeviceBuffer.from_array(tr), tr.shape[0]) for a, tr in tables]
    got = prover.prove_air_tables(dev, opts, label=b"gpu-decode")
    for _, d, _ in dev:
        d.free()
    ref = replay_air_tables(tables, 4, 3, 2, 8, 0, b"gpu-decode")
    assert got[0].precomputed_root == precomputed_commitment(dec, 4, 3)
    assert proof_io.serialize_multi_proof(got) == proof_io.serialize_multi_proof([AirTableProof(**r) for r in ref])
    assert verify_air_tables([dec, cpu], got, 4, 3, 2, 8, 0, b"gpu-decode")[0]


def test_gpu_cpu_table_proof(gpu):
    """CPU (its TIMESTAMP column fixed, 38 constraints, 20 interactions) over
    an executed RV64IM program, beside the program's DECODE table and the
    CPU32 rows of its word instructions: the GPU proofs equal the oracle
    replay byte for byte, the fixed-column roots are the verifier's own and
    all verify."""
    from lambda_vm_amd import proof_io, prover
    from lambda_vm_amd import vm_tables as V
    from lambda_vm_amd.prover import AirTableProof
    from oracle.air_proof import precomputed_commitment
    from tests.test_vm_table_cpu import ENTRY, cpu32_rows, cpu_ops, program, run

    print("[gpu-air] CPU 512 x 37 + DECODE + CPU32, blowup 4", flush=True)
    _, logs = run(121, steps=512)
    ops = cpu_ops(logs)
    entries = V.decode_entries_from_words([(ENTRY, program(121, length=1024))])
    cpu_tr = V.generate_cpu_trace(ops)
    cpu, dec, c32 = V.cpu_table_air(cpu_tr.shape[0]), V.decode_table_air(entries), V.cpu32_table_air()
    for a in (cpu, dec, c32):
        a.bus_descriptor = A.descriptor(a)
    tables = [(cpu, cpu_tr), (dec, V.generate_decode_trace(entries, [o.pc for o in ops])),
              (c32, V.generate_cpu32_trace(cpu32_rows(ops)))]
    tables = [(a, np.ascontiguousarray(tr, dtype=np.uint64)) for a, tr in tables]
    opts = prover.ProofOptions(blowup=4, coset_offset=3, fri_final_log_degree=2, num_queries=8)
    dev = [(a, gpu.DeviceBuffer.from_array(tr), tr.shape[0]) for a, tr in tables]
    got = prover.prove_air_tables(dev, opts, label=b"gpu-cpu")
    for _, d, _ in dev:
        d.free()
    ref = replay_air_tables(tables, 4, 3, 2, 8, 0, b"gpu-cpu")
    assert got[0].precomputed_root == precomputed_commitment(cpu, 4, 3)
    assert got[1].precomputed_root == precomputed_commitment(dec, 4, 3)
    assert proof_io.serialize_multi_proof(got) == proof_io.serialize_multi_proof([AirTableProof(**r) for r in ref])
    assert verify_air_tables([cpu, dec, c32], got, 4, 3, 2, 8, 0, b"gpu-cpu")[0]


def test_gpu_ecdas_table_proof(gpu):
    """ECDAS (521 columns, 200 degree-3 convolution-carry constraints, 388
    interactions) over the double-and-add steps of two secp256k1 scalar
    multiplications: the GPU proof equals the oracle replay byte for byte and
    verifies."""
    from lambda_vm_amd import proof_io, prover
    from lambda_vm_amd import vm_tables as V
    from lambda_vm_amd.prover import AirTableProof
    from tests.test_vm_table_ecdas import ecdas_ops

    print("[gpu-air] ECDAS 1024 x 521, blowup 4", flush=True)
    air = V.ecdas_table_air()
    air.bus_descriptor = A.descriptor(air)
    k1, k2 = V.SECP_N - 12345, 0x7FFF_FFFF_FFFF_FFFF_0123_4567_89AB_CDEF_0011_2233_4455_6677_8899_AABB_CCDD_EEFF
    tables = [(air, np.ascontiguousarray(V.generate_ecdas_trace(ecdas_ops([k1, k2])), dtype=np.uint64))]
    opts = prover.ProofOptions(blowup=4, coset_offset=3, fri_final_log_degree=2, num_queries=8)
    dev = [(a, gpu.DeviceBuffer.from_array(tr), tr.shape[0]) for a, tr in tables]
    got = prover.prove_air_tables(dev, opts, label=b"gpu-ecdas")
    for _, d, _ in dev:
        d.free()
    ref = replay_air_tables(tables, 4, 3, 2, 8, 0, b"gpu-ecdas")
    assert proof_io.serialize_multi_proof(got) == proof_io.serialize_multi_proof([AirTableProof(**r) for r in ref])
    assert verify_air_tables([air], got, 4, 3, 2, 8, 0, b"gpu-ecdas")[0]


def test_gpu_ecsm_ecdas_proof(gpu):
    """ECSM (667 columns, 413 constraints, 579 interactions) with the ECDAS
    chain of the same three scalar multiplications: the GPU proofs equal the
    oracle replay byte for byte and verify."""
    from lambda_vm_amd import proof_io, prover
    from lambda_vm_amd import vm_tables as V
    from lambda_vm_amd.prover import AirTableProof
    from tests.test_vm_table_ecsm import KS, ecdas_rows, ecsm_ops

    print("[gpu-air] ECSM 4 x 667 + ECDAS 1024 x 521, blowup 4", flush=True)
    ops = ecsm_ops(KS)
    ecsm, ecdas = V.ecsm_table_air(), V.ecdas_table_air()
    for a in (ecsm, ecdas):
        a.bus_descriptor = A.descriptor(a)
    tables = [(ecsm, V.generate_ecsm_trace(ops)), (ecdas, V.generate_ecdas_trace(ecdas_rows(ops)))]
    tables = [(a, np.ascontiguousarray(tr, dtype=np.uint64)) for a, tr in tables]
    opts = prover.ProofOptions(blowup=4, coset_offset=3, fri_final_log_degree=1, num_queries=8)
    dev = [(a, gpu.DeviceBuffer.from_array(tr), tr.shape[0]) for a, tr in tables]
    got = prover.prove_air_tables(dev, opts, label=b"gpu-ecsm")
    for _, d, _ in dev:
        d.free()
    ref = replay_air_tables(tables, 4, 3, 1, 8, 0, b"gpu-ecsm")
    assert proof_io.serialize_multi_proof(got) == proof_io.serialize_multi_proof([AirTableProof(**r) for r in ref])
    assert verify_air_tables([ecsm, ecdas], got, 4, 3, 1, 8, 0, b"gpu-ecsm")[0]


def test_gpu_l2g_proofs(gpu):
    """LOCAL_TO_GLOBAL: the epoch-local AIR (Memory bus, range and ordering
    checks) and the global AIR (GlobalMemory bus) over one epoch's boundary
    trace: the GPU proofs equal the oracle replay byte for byte and verify."""
    from lambda_vm_amd import proof_io, prover
    from lambda_vm_amd import vm_tables as V
    from lambda_vm_amd.prover import AirTableProof
    from tests.test_vm_table_l2g import PAGE, scenario

    print("[gpu-air] L2G memory + global over 32 rows, blowup 4", flush=True)
    init, touched = scenario(131, epochs=3)
    rows = V.epoch_boundaries({PAGE + i: v for i, v in enumerate(init)}, touched)[2]
    tr = np.ascontiguousarray(V.generate_l2g_trace(rows), dtype=np.uint64)
    mem, glob = V.l2g_memory_air(3), V.l2g_global_air(3)
    for a in (mem, glob):
        a.bus_descriptor = A.descriptor(a)
    tables = [(mem, tr), (glob, tr)]
    opts = prover.ProofOptions(blowup=4, coset_offset=3, fri_final_log_degree=2, num_queries=8)
    dev = [(a, gpu.DeviceBuffer.from_array(t), t.shape[0]) for a, t in tables]
    got = prover.prove_air_tables(dev, opts, label=b"gpu-l2g")
    for _, d, _ in dev:
        d.free()
    ref = replay_air_tables(tables, 4, 3, 2, 8, 0, b"gpu-l2g")
    assert proof_io.serialize_multi_proof(got) == proof_io.serialize_multi_proof([AirTableProof(**r) for r in ref])
    assert verify_air_tables([mem, glob], got, 4, 3, 2, 8, 0, b"gpu-l2g")[0]


def test_gpu_copy_table_proof(gpu):
    """COPY (41 columns, 40 degree-3 constraints, 33 interactions, two Linear
    read/write multiplicities) over memcpy/memmove, memset and write calls of
    0..64 bytes: the GPU proof equals the oracle replay byte for byte and
    verifies."""
    from lambda_vm_amd import proof_io, prover
    from lambda_vm_amd import vm_tables as V
    from lambda_vm_amd.prover import AirTableProof
    from tests.test_vm_table_copy import run

    _, rows, _, _, _ = run(141)
    print(f"[gpu-air] COPY {len(rows)} rows x 41, blowup 4", flush=True)
    air = V.copy_table_air()
    air.bus_descriptor = A.descriptor(air)
    tables = [(air, np.ascontiguousarray(V.generate_copy_trace(rows), dtype=np.uint64))]
    opts = prover.ProofOptions(blowup=4, coset_offset=3, fri_final_log_degree=2, num_queries=8)
    dev = [(a, gpu.DeviceBuffer.from_array(tr), tr.shape[0]) for a, tr in tables]
    got = prover.prove_air_tables(dev, opts, label=b"gpu-copy")
    for _, d, _ in dev:
        d.free()
    ref = replay_air_tables(tables, 4, 3, 2, 8, 0, b"gpu-copy")
    assert proof_io.serialize_multi_proof(got) == proof_io.serialize_multi_proof([AirTableProof(**r) for r in ref])
    assert verify_air_tables([air], got, 4, 3, 2, 8, 0, b"gpu-copy")[0]
