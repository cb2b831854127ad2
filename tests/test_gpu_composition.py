"""GPU constraint evaluation and composition polynomial (lvm_eval_constraints /
lvm_eval_composition) against the oracle on the same seeded inputs:
FibonacciMultiColumn, Quadratic and a 4-bus LogUp table, run both as
run-time compiled kernels and on the interpreter (value slots in LDS and in
global scratch), bit-reversed and natural output order, a low-degree check of
the GPU output for valid traces, and program validation."""
import random

import numpy as np
import pytest

from lambda_vm_amd import air as A
from lambda_vm_amd import workload
from test_oracle_composition import G, T, _high_coeffs, rand3, scale3
from test_oracle_logup import alpha_powers

pytestmark = pytest.mark.gpu
P = workload.P


@pytest.fixture(params=[True, False], ids=["compiled", "interpreted"])
def jit(request, gpu):
    """Constraint programs as compiled kernels or on the interpreter; checks
    that the chosen kernel is the one that ran."""
    gpu.set_constraint_jit(request.param)
    gpu.prof_enable(True)
    gpu.prof_reset()
    yield request.param
    names = ("composition_compiled", "constraint_evals_compiled") if request.param else ("composition",
                                                                                      "constraint_evals")
    ran = sum(gpu.prof_query(nm)[0] for nm in names)
    other = ("composition", "constraint_evals") if request.param else ("composition_compiled",
                                                                       "constraint_evals_compiled")
    stray = sum(gpu.prof_query(nm)[0] for nm in other)
    gpu.prof_enable(False)
    gpu.set_constraint_jit(True)
    assert ran > 0 and stray == 0


def canon(a):
    a = np.asarray(a, dtype=np.uint64)
    return np.where(a >= np.uint64(P), a - np.uint64(P), a)


def _bitrev_perm(N):
    bits = N.bit_length() - 1
    return np.array([int(format(i, f"0{bits}b")[::-1], 2) if bits else 0 for i in range(N)], dtype=np.int64)


def _case(gpu, ora, which, n, B, m, rng, valid=True):
    w = ora.lde_weights(n, G)
    aux_tr = None
    rap = alpha = toff = None
    if which == "fib":
        air = A.fibonacci_multi_column(m)
        if valid:
            seeds = np.array([[rng.randrange(P), rng.randrange(P)] for _ in range(m)], dtype=np.uint64)
            trace = A.fibonacci_trace(n, m, seeds)
        else:
            trace = workload.trace_table(7, 1, n, m)
        A.with_boundaries_from_trace(air, trace)
    elif which == "quad":
        air = A.quadratic_multi(m)
        trace = A.quadratic_trace(n, m, np.array([rng.randrange(P) for _ in range(m)], dtype=np.uint64))
        A.with_boundaries_from_trace(air, trace)
    else:
        air = A.logup_air(24, A.four_bus_interactions())
        d = A.descriptor(air)
        trace = workload.trace_table(8, 2, n, 24)
        z, al = rand3(rng), rand3(rng)
        ap = alpha_powers(al, d.alpha_powers_len)
        aux, L, _ = ora.logup_aux(d, trace, ap, z)
        aux_tr = aux.reshape(n, -1)
        rap, alpha = [z, al], ap
        toff = scale3(T(L), pow(n, P - 2, P))
    main_h, main_lde = gpu.coset_lde_row_major_with_merkle_tree_keep(trace, n, trace.shape[1], B, w)
    aux_h = aux_lde = None
    if aux_tr is not None:
        aux_h, aux_lde = gpu.coset_lde_ext3_row_major_with_merkle_tree_keep(aux_tr, n, aux_tr.shape[1] // 3, B, w)
    groups, ex = air.groups()
    betas = [rand3(rng) for _ in range(air.num_constraints)]
    bnd = [(b.col, b.is_aux, b.step, b.value) for b in air.boundary]
    b_betas = [rand3(rng) for _ in bnd]
    return dict(air=air, main_h=main_h, aux_h=aux_h, main_lde=main_lde, aux_lde=aux_lde, groups=groups, ex=ex,
                betas=betas, bnd=bnd, b_betas=b_betas, rap=rap, alpha=alpha, toff=toff)


@pytest.mark.parametrize("which,n,B,m", [("fib", 64, 4, 5), ("fib", 256, 2, 80), ("quad", 128, 8, 3),
                                         ("logup", 64, 4, 0), ("logup", 1024, 4, 0)])
def test_composition_matches_oracle(gpu, ora, jit, which, n, B, m):
    rng = random.Random(hash((which, n, B, m)) & 0xFFFFFF)
    c = _case(gpu, ora, which, n, B, m, rng)
    air = c["air"]
    args = (B, n, G, c["betas"], c["groups"], c["ex"], c["bnd"], c["b_betas"], c["rap"], c["alpha"], c["toff"])
    ref = ora.eval_composition(air.device, c["main_lde"], c["aux_lde"], *args)
    hn = gpu.eval_composition(air.device, c["main_h"], c["aux_h"], *args, natural_order=True)
    got = hn.to_array(np.uint64, (n * B, 3))
    assert np.array_equal(canon(got), canon(ref))
    hb = gpu.eval_composition(air.device, c["main_h"], c["aux_h"], *args)
    got_b = hb.to_array(np.uint64, (n * B, 3))
    assert np.array_equal(canon(got_b), canon(ref)[_bitrev_perm(n * B)])
    if which != "logup" or n <= 64:
        ev = gpu.eval_constraints(air.device, c["main_h"], c["aux_h"], B, c["rap"], c["alpha"], c["toff"])
        ev_ref = ora.eval_constraints(air.device, c["main_lde"], c["aux_lde"], B, c["rap"], c["alpha"], c["toff"])
        assert np.array_equal(canon(ev), canon(ev_ref))


@pytest.mark.parametrize("which", ["fib", "logup"])
def test_composition_many_boundary_steps(gpu, ora, jit, which):
    """Boundaries on seven steps (the compiled kernel keeps four steps' sums
    live through the program and sums the rest afterwards), several per step,
    on main and aux columns, with arbitrary values."""
    rng = random.Random(77)
    n, B = 64, 4
    c = _case(gpu, ora, which, n, B, 5, rng)
    air = c["air"]
    w = 5 if which == "fib" else 24
    bnd = [(col, False, s, rand3(rng)) for s in (0, 3, 9, 63, 1, 17, 40) for col in range(0, w, 2)]
    if which == "logup":
        bnd += [(0, True, 9, rand3(rng)), (1, True, 40, rand3(rng)), (0, True, 5, rand3(rng))]
    bb = [rand3(rng) for _ in bnd]
    args = (B, n, G, c["betas"], c["groups"], c["ex"], bnd, bb, c["rap"], c["alpha"], c["toff"])
    ref = ora.eval_composition(air.device, c["main_lde"], c["aux_lde"], *args)
    got = gpu.eval_composition(air.device, c["main_h"], c["aux_h"], *args, natural_order=True).to_array(
        np.uint64, (n * B, 3))
    assert np.array_equal(canon(got), canon(ref))


@pytest.mark.parametrize("which", ["fib", "quad", "logup"])
def test_gpu_composition_low_degree(gpu, ora, jit, which):
    rng = random.Random(5)
    n, B = 64, 4
    c = _case(gpu, ora, which, n, B, 4, rng)
    air = c["air"]
    args = (B, n, G, c["betas"], c["groups"], c["ex"], c["bnd"], c["b_betas"], c["rap"], c["alpha"], c["toff"])
    H = gpu.eval_composition(air.device, c["main_h"], c["aux_h"], *args, natural_order=True).to_array(
        np.uint64, (n * B, 3))
    deg = max(max(air.max_degree * (n - 1) + e - n for e in c["ex"]), n - 2)
    assert not _high_coeffs(ora, H, deg).any()


def test_composition_rejects_bad_programs(gpu, ora):
    rng = random.Random(9)
    c = _case(gpu, ora, "fib", 64, 4, 2, rng)
    air = c["air"]
    big = A.fibonacci_multi_column(5)  # reads columns the 2-column table lacks
    with pytest.raises(gpu.LvmError):
        gpu.eval_composition(big.device, c["main_h"], None, 4, 64, G, [rand3(rng)] * 5, [0] * 5, [2])
    with pytest.raises(gpu.LvmError):  # zerofier group out of range
        gpu.eval_composition(air.device, c["main_h"], None, 4, 64, G, c["betas"], [1, 1], [2])
    logup = A.logup_air(24, A.four_bus_interactions())
    with pytest.raises(gpu.LvmError):  # aux program without an aux table
        gpu.eval_composition(logup.device, c["main_h"], None, 4, 64, G, [rand3(rng)] * logup.num_constraints,
                             [0] * logup.num_constraints, [0])


@pytest.mark.parametrize("which,n,B", [("quad", 64, 4), ("logup", 256, 4), ("fib", 128, 2)])
def test_composition_d2_commit_matches_oracle(gpu, ora, jit, which, n, B):
    rng = random.Random(hash((which, n)) & 0xFFFF)
    c = _case(gpu, ora, which, n, B, 3, rng)
    air = c["air"]
    args = (B, n, G, c["betas"], c["groups"], c["ex"], c["bnd"], c["b_betas"], c["rap"], c["alpha"], c["toff"])
    N = n * B
    hb = gpu.eval_composition(air.device, c["main_h"], c["aux_h"], *args)  # bit-reversed rows
    parts = gpu.composition_commit_d2(hb, N, n, G)
    ref_h = ora.eval_composition(air.device, c["main_lde"], c["aux_lde"], *args)
    halves = ora.decompose_d2(ref_h, G)
    ref_lde, ref_nodes = ora.lde_commit(halves, 2, pow(G, P - 2, P))
    assert parts.root == bytes(ref_nodes[0])
    assert np.array_equal(canon(parts.download()), canon(ref_lde))


@pytest.mark.parametrize("which", ["fib", "logup"])
def test_pinned_reference_style_program(gpu, ora, jit, which):
    """Programs lowered with pinned root slots and no root node list (the
    reference lowering's form): the library finds each constraint's node."""
    from lambda_vm_amd import constraint_ir as ir
    rng = random.Random(17)
    c = _case(gpu, ora, which, 64, 4, 6, rng)
    air = c["air"]
    dev = ir.lower(air.program)  # pinned roots
    dev.root_nodes = None
    args = (4, 64, G, c["betas"], c["groups"], c["ex"], c["bnd"], c["b_betas"], c["rap"], c["alpha"], c["toff"])
    ref = ora.eval_composition(air.device, c["main_lde"], c["aux_lde"], *args)
    got = gpu.eval_composition(dev, c["main_h"], c["aux_h"], *args, natural_order=True).to_array(np.uint64,
                                                                                                (256, 3))
    assert np.array_equal(canon(got), canon(ref))
    ev = gpu.eval_constraints(dev, c["main_h"], c["aux_h"], 4, c["rap"], c["alpha"], c["toff"])
    ev_ref = ora.eval_constraints(dev, c["main_lde"], c["aux_lde"], 4, c["rap"], c["alpha"], c["toff"])
    assert np.array_equal(canon(ev), canon(ev_ref))


@pytest.mark.parametrize("num_parts,n,B", [(1, 64, 4), (2, 128, 4), (3, 64, 4), (4, 256, 8), (5, 32, 8)])
def test_composition_commit_parts_matches_oracle(gpu, ora, num_parts, n, B):
    """lvm_composition_commit_parts for 1..5 parts against the oracle's
    (H itself, the degree-2 split, break_in_parts on the coset): the parts
    LDE, its root, and the parts' values at a point by the barycentric
    formula over the n trace-coset rows."""
    from oracle.air_proof import break_in_parts_lde
    N = n * B
    H = workload.trace_table(13, num_parts, N, 3)  # natural-order values
    perm = _bitrev_perm(N)
    hb = gpu.DeviceBuffer.from_array(np.ascontiguousarray(H[perm]))
    parts = gpu.composition_commit_parts(hb, N, n, num_parts, G)
    if num_parts == 1:
        ref = H
    elif num_parts == 2:
        ref, _ = ora.lde_commit(ora.decompose_d2(H, G), 2, pow(G, P - 2, P))
    else:
        ref = break_in_parts_lde(H, num_parts, G)
    nodes = ora.commit_rows_bit_reversed(ref)
    assert parts.root == bytes(nodes[0])
    assert np.array_equal(canon(parts.download()), canon(ref))
    z = np.array([5, 7, 11], dtype=np.uint64)
    got = gpu.barycentric_base_multi(parts, G, z.reshape(1, 3))[0]
    want = ora.barycentric_base(ref, n, B, G, z)
    assert np.array_equal(canon(got), canon(want))
    hb.free()
    parts.close()


def test_composition_commit_parts_rejects_bad_shapes(gpu):
    hb = gpu.DeviceBuffer(64 * 24)
    with pytest.raises(gpu.LvmError):
        gpu.composition_commit_parts(hb, 64, 16, 9, G)   # more than 8 parts
    with pytest.raises(gpu.LvmError):
        gpu.composition_commit_parts(hb, 64, 64, 3, G)   # trace rows above N/2
    with pytest.raises(gpu.LvmError):
        gpu.composition_commit_parts(hb, 64, 16, 0, G)
    hb.free()



@pytest.mark.parametrize("natural", [False, True])
def test_batched_composition_equals_per_table(gpu, ora, natural):
    """lvm_eval_composition_batch over tables of different programs, sizes and
    blowups (one LDE under 256 rows, two tables sharing one program) equals
    lvm_eval_composition per table, and the oracle, and runs as one compiled
    launch."""
    rng = random.Random(404 + natural)
    specs = [("fib", 256, 4, 9), ("quad", 128, 8, 3), ("logup", 512, 4, 0), ("fib", 16, 4, 9), ("fib", 1024, 2, 9)]
    cases, comps, refs = [], [], []
    for which, n, B, m in specs:
        c = _case(gpu, ora, which, n, B, m, rng)
        air = c["air"]
        args = (B, n, G, c["betas"], c["groups"], c["ex"], c["bnd"], c["b_betas"], c["rap"], c["alpha"], c["toff"])
        ref = canon(ora.eval_composition(air.device, c["main_lde"], c["aux_lde"], *args))
        refs.append(ref if natural else ref[_bitrev_perm(n * B)])
        single = gpu.eval_composition(air.device, c["main_h"], c["aux_h"], *args, natural_order=natural)
        assert np.array_equal(canon(single.to_array(np.uint64, (n * B, 3))), refs[-1])
        single.free()
        cases.append(c)
        comps.append(dict(dev=air.device, main=c["main_h"], aux=c["aux_h"], next_step=B, trace_rows=n, coset_offset=G,
                          betas=c["betas"], root_group=c["groups"], group_exemptions=c["ex"], boundary=c["bnd"],
                          b_betas=c["b_betas"], rap=c["rap"], alpha=c["alpha"], table_offset=c["toff"]))
    gpu.prof_enable(True)
    gpu.prof_reset()
    outs = gpu.eval_composition_batch(comps, natural_order=natural)
    launches = gpu.prof_query("composition_compiled")[0]
    gpu.prof_enable(False)
    assert launches == 1
    for (which, n, B, m), o, ref in zip(specs, outs, refs):
        assert np.array_equal(canon(o.to_array(np.uint64, (n * B, 3))), ref), (which, n, B)
        o.free()
