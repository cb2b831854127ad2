"""Synthetic AIRs of the constraint + LogUp stage, captured as constraint
programs, with their zerofier groups and boundary constraints.

Restated from the reference's example AIRs and LogUp constraint emitters:
  - FibonacciMultiColumn (crypto/stark/src/examples/fibonacci_multi_column.rs
    :40-55): per column a[i+2] - a[i+1] - a[i], two end exemptions;
    boundaries a[0], a[1] per column (:150-170);
  - Quadratic (crypto/stark/src/examples/quadratic_air.rs:36-41): x[i+1] - x[i]^2,
    one end exemption, widened here to every column; boundary x[0];
  - LogUp (crypto/stark/src/lookup.rs:2160-2325 emit_fingerprint,
    emit_logup_batched_term, emit_logup_accumulated, emit_logup_constraints):
    one degree-3 batched-term constraint per committed pair, then the
    accumulated constraint over the absorbed 1-2 interactions; boundary
    acc[0] = 0 (:1330-1345).
Frame offsets read trace rows i + offset; on the LDE a trace step is `blowup`
rows (next_step = blowup).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import constraint_ir as ir
from .logup import BusInteraction, Packing, build_fingerprint_descriptor, split_interactions

P = ir.P
SHIFT_8, SHIFT_16, SHIFT_24 = 1 << 8, 1 << 16, 1 << 24


@dataclass
class Boundary:
    col: int
    is_aux: bool
    step: int
    value: tuple  # Fp3


@dataclass
class AirProgram:
    """A captured AIR: lowered program, per-constraint end exemptions, boundaries.

    bus: an AirWithBuses table (crypto/stark/src/lookup.rs:1000-1085; every VM
    table and the synthetic LogUp table) rather than one of the reference's
    example AIRs. The two kinds differ in three pieces of public shape
    metadata that fix the proof layout:
      - composition parts = composition_poly_degree_bound(n) / n: for bus AIRs
        (max_degree - 1).max(1) with max_degree the larger of the base
        constraints' and LogUp's (lookup.rs:1063-1078, logup_max_degree
        :2290-2301); for an example AIR its own bound (`example_parts`:
        fibonacci_multi_column.rs:93 gives 1, quadratic_air.rs:149 gives 2);
      - frame rows: bus AIRs have transition offsets [0, 1] (lookup.rs:950);
        example AIRs their own (fibonacci_multi_column.rs:188 [0, 1, 2],
        quadratic_air.rs:88 [0, 1]);
      - columns opened at the next rows: bus AIRs only the LogUp accumulated
        column (lookup.rs:1042-1053), none without interactions; example AIRs
        every column (the trait default, traits.rs:217-220).
    """
    name: str
    num_main: int
    num_aux: int  # Fp3 columns
    program: ir.ConstraintProgram
    device: ir.DeviceProgram
    exemptions: list  # per constraint
    boundary: list = field(default_factory=list)
    interactions: list = field(default_factory=list)
    max_degree: int = 1
    bus: bool = True
    example_parts: int | None = None
    # preprocessed tables (prover.rs:1194-1290): the first num_precomputed
    # columns are fixed per table, produced by precomputed() as (n, k), and
    # commit to a tree of their own; the rest (multiplicities) to the main tree
    num_precomputed: int = 0
    precomputed: object = None

    @property
    def num_constraints(self) -> int:
        return len(self.program.roots)

    @property
    def num_parts(self) -> int:
        """Composition parts: degree bound / trace length (prover.rs:1705)."""
        if self.bus:
            return max(self.max_degree - 1, 1)
        if self.example_parts is None:
            raise ValueError(f"{self.name}: an example AIR states its composition degree bound")
        return self.example_parts

    @property
    def frame_rows(self) -> int:
        """Rows of the OOD frame: the number of transition offsets."""
        return 2 if self.bus else int(self.device.max_offset) + 1

    def next_row_cols(self) -> list:
        """Full-width ([main | aux]) indices of the columns opened at the next
        rows, ascending."""
        if self.bus:
            # the accumulated column is the last aux column (after the committed pairs)
            return [self.num_main + self.num_aux - 1] if self.interactions else []
        return list(range(self.num_main + self.num_aux))

    def groups(self):
        """(group of each constraint, exemption count of each group)."""
        ex = sorted(set(self.exemptions))
        return [ex.index(e) for e in self.exemptions], ex


def fibonacci_multi_column(m: int) -> AirProgram:
    b = ir.IrBuilder()
    for c in range(m):
        a0, a1, a2 = b.main(0, c), b.main(1, c), b.main(2, c)
        b.emit(c, a2 - a1 - a0)
    prog = b.finish(m)
    return AirProgram(f"fibonacci_multi_column_{m}", m, 0, prog, ir.lower(prog, pin_roots=False), [2] * m,
                      max_degree=1, bus=False, example_parts=1)


def quadratic_multi(m: int) -> AirProgram:
    b = ir.IrBuilder()
    for c in range(m):
        x, x_next = b.main(0, c), b.main(1, c)
        b.emit(c, x_next - x * x)
    prog = b.finish(m)
    return AirProgram(f"quadratic_{m}", m, 0, prog, ir.lower(prog, pin_roots=False), [1] * m, max_degree=2,
                      bus=False, example_parts=2)


# ---- LogUp constraint emission (lookup.rs:2160-2325) ----

def _packing_elements(b: ir.IrBuilder, packing: Packing, c: int, off: int):
    w2 = lambda x: b.main(off, x) + b.main(off, x + 1) * SHIFT_16  # noqa: E731
    w4 = lambda x: (b.main(off, x) + b.main(off, x + 1) * SHIFT_8 + b.main(off, x + 2) * SHIFT_16  # noqa: E731
                    + b.main(off, x + 3) * SHIFT_24)
    v = packing.value
    col = lambda x: b.main(off, x)  # noqa: E731
    if v == "Direct":
        return [col(c)]
    if v == "Word2L":
        return [w2(c)]
    if v == "Word4L":
        return [w4(c)]
    if v == "DWordWL":
        return [col(c), col(c + 1)]
    if v == "DWordHHW":
        return [col(c), w2(c + 1)]
    if v == "DWordWHH":
        return [w2(c), col(c + 2)]
    if v == "DWordHL":
        return [w2(c), w2(c + 2)]
    if v == "DWordBL":
        return [w4(c), w4(c + 4)]
    if v == "QuadHL":
        return [w2(c + 2 * i) for i in range(4)]
    return [col(c + i) for i in range(4)]


def _linear(b: ir.IrBuilder, terms, off: int):
    acc = b.const_base(0)
    for t in terms:
        if t.column is None:
            acc = acc + b.const_base(t.coefficient % P)
        else:
            acc = acc + b.main(off, t.column) * b.const_base(t.coefficient % P)
    return acc


def emit_fingerprint(b: ir.IrBuilder, it: BusInteraction, off: int = 0):
    """z - (bus_id + sum_j alpha^j v_j) (lookup.rs:2160-2185), subtraction distributed."""
    fp = -(b.const_base(it.bus_id % P) - b.challenge(0))
    a = 1
    for bv in it.values:
        elems = (_packing_elements(b, bv.packing, bv.start_column, off) if bv.packing is not None
                 else [_linear(b, bv.linear, off)])
        for e in elems:
            fp = fp - e * b.alpha_power(a)
            a += 1
    return fp


def emit_multiplicity(b: ir.IrBuilder, m, off: int = 0):
    col = lambda c: b.main(off, c)  # noqa: E731
    k = m.kind
    if k == "One":
        return b.const_base(1)
    if k == "Column":
        return col(m.cols[0])
    if k == "Sum":
        return col(m.cols[0]) + col(m.cols[1])
    if k == "Negated":
        return b.const_base(1) - col(m.cols[0])
    if k == "Diff":
        return col(m.cols[0]) - col(m.cols[1])
    if k == "Sum3":
        return col(m.cols[0]) + col(m.cols[1]) + col(m.cols[2])
    return _linear(b, m.linear, off)


def emit_logup_constraints(b: ir.IrBuilder, interactions, idx_base: int) -> int:
    """Batched-term constraints, then the accumulated one; returns the next index."""
    pairs, absorbed = split_interactions(len(interactions))
    idx = idx_base
    for p in range(pairs):
        ia, ib = interactions[2 * p], interactions[2 * p + 1]
        c = b.aux(0, p)
        ma, mb = emit_multiplicity(b, ia.multiplicity), emit_multiplicity(b, ib.multiplicity)
        fa, fb = emit_fingerprint(b, ia), emit_fingerprint(b, ib)
        ta = ma * fb
        ta = ta if ia.is_sender else -ta
        tb = mb * fa
        tb = tb if ib.is_sender else -tb
        b.emit(idx, c * fa * fb - ta - tb)
        idx += 1
    acc_idx = pairs
    delta = b.aux(1, acc_idx) - b.aux(0, acc_idx)
    for i in range(pairs):
        delta = delta - b.aux(0, i)
    delta = delta + b.table_offset()
    ab = interactions[len(interactions) - absorbed:]
    if absorbed == 1:
        m = emit_multiplicity(b, ab[0].multiplicity)
        f = emit_fingerprint(b, ab[0])
        mt = m if ab[0].is_sender else -m
        root = -(mt - delta * f)
    else:
        m1, m2 = emit_multiplicity(b, ab[0].multiplicity), emit_multiplicity(b, ab[1].multiplicity)
        f1, f2 = emit_fingerprint(b, ab[0]), emit_fingerprint(b, ab[1])
        t1 = m1 * f2
        t1 = t1 if ab[0].is_sender else -t1
        t2 = m2 * f1
        t2 = t2 if ab[1].is_sender else -t2
        root = delta * f1 * f2 - t1 - t2
    b.emit(idx, root)
    return idx + 1


def logup_air(num_main: int, interactions) -> AirProgram:
    """An AirWithBuses table whose only transition constraints are LogUp's."""
    b = ir.IrBuilder()
    n_next = emit_logup_constraints(b, interactions, 0)
    prog = b.finish(0)
    pairs, absorbed = split_interactions(len(interactions))
    bnd = [Boundary(pairs, True, 0, (0, 0, 0))]
    deg = max(3 if pairs else 0, 1 + absorbed)
    return AirProgram(f"logup_{len(interactions)}", num_main, pairs + 1, prog, ir.lower(prog, pin_roots=False),
                      [0] * n_next, boundary=bnd, interactions=list(interactions), max_degree=deg)


def product_bus_air(interactions) -> AirProgram:
    """A bus table of 25 main columns with one degree-4 base constraint
    (columns 20..23 multiply to column 24 on every row) beside the LogUp
    constraints of `interactions`: max degree 4, so three composition parts
    (the break_in_parts path of crypto/stark/src/prover.rs:1858-1900)."""
    b = ir.IrBuilder()
    x = [b.main(0, c) for c in range(20, 25)]
    b.emit(0, x[0] * x[1] * x[2] * x[3] - x[4])
    n_total = emit_logup_constraints(b, interactions, 1) if interactions else 1
    prog = b.finish(25)
    pairs, absorbed = split_interactions(len(interactions)) if interactions else (0, 0)
    num_aux = pairs + 1 if interactions else 0
    bnd = [Boundary(pairs, True, 0, (0, 0, 0))] if interactions else []
    deg = max(4, 3 if pairs else 0, 1 + absorbed if interactions else 0)
    return AirProgram(f"product_bus_{len(interactions)}", 25, num_aux, prog, ir.lower(prog, pin_roots=False),
                      [0] * n_total, boundary=bnd, interactions=list(interactions), max_degree=deg)


def product_trace(n: int, cfg: int = 9) -> np.ndarray:
    """A trace of product_bus_air: uniform columns, column 24 the product of 20..23."""
    from .workload import trace_table
    t = trace_table(cfg, 0, n, 25)
    prod = np.ones(n, dtype=object)
    for c in range(20, 24):
        prod = prod * t[:, c].astype(object) % P
    t[:, 24] = prod.astype(np.uint64)
    return t


def four_bus_interactions():
    """4 buses over a 24-column table: bus b sends (col 2b, col 2b+1) once per
    row and receives (col 8+2b, col 9+2b) with multiplicity col 16+b."""
    from .logup import BusValue, Multiplicity, ONE
    its = []
    for bus in range(4):
        its.append(BusInteraction.sender(bus, ONE, [BusValue.column(2 * bus), BusValue.column(2 * bus + 1)]))
        its.append(BusInteraction.receiver(bus, Multiplicity("Column", (16 + bus,)),
                                           [BusValue.column(8 + 2 * bus), BusValue.column(9 + 2 * bus)]))
    return its


def with_boundaries_from_trace(air: AirProgram, trace: np.ndarray) -> AirProgram:
    """Boundary constraints of the base AIRs, valued from the trace itself."""
    if air.name.startswith("fibonacci"):
        air.boundary = [Boundary(c, False, s, (int(trace[s, c]) % P, 0, 0)) for c in range(air.num_main)
                        for s in (0, 1)]
    elif air.name.startswith("quadratic"):
        air.boundary = [Boundary(c, False, 0, (int(trace[0, c]) % P, 0, 0)) for c in range(air.num_main)]
    return air


def fibonacci_trace(n: int, m: int, seeds: np.ndarray) -> np.ndarray:
    """Valid trace: column c is the Fibonacci sequence from (seeds[c,0], seeds[c,1])."""
    t = np.zeros((n, m), dtype=object)
    t[0] = [int(x) % P for x in seeds[:, 0]]
    t[1] = [int(x) % P for x in seeds[:, 1]]
    for i in range(2, n):
        t[i] = (t[i - 1] + t[i - 2]) % P
    return t.astype(np.uint64)


def quadratic_trace(n: int, m: int, seeds: np.ndarray) -> np.ndarray:
    t = np.zeros((n, m), dtype=object)
    t[0] = [int(x) % P for x in seeds]
    for i in range(1, n):
        t[i] = (t[i - 1] * t[i - 1]) % P
    return t.astype(np.uint64)


def descriptor(air: AirProgram):
    return build_fingerprint_descriptor(air.interactions)
