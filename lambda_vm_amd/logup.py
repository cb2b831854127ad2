"""Host-side LogUp interaction model and its flat device descriptor.

Mirrors the reference's bus-interaction types (crypto/stark/src/lookup.rs:
`Packing` :150-210, `BusValue`, `LinearTerm`, `Multiplicity`,
`BusInteraction`, `split_interactions` :117-125) and the compiler that
flattens a table's interactions into the CSR descriptor the device kernels
walk (crypto/stark/src/logup_gpu.rs:100-290 `build_fingerprint_descriptor`,
`encode_signed_multiplicity` :54-98, `i64_to_canonical` :37-43). The
descriptor is what crosses the C ABI (lvm_logup_desc).

Fingerprint of interaction k at row i:
    fp = z - (bus_id + sum_j alpha^j * v_j(i))      (v_j the bus elements, j >= 1)
and its signed term m(i) / fp, negated for receivers.
"""
from __future__ import annotations

import enum
from dataclasses import dataclass, field

import numpy as np

P = 0xFFFFFFFF00000001
SHIFT_8, SHIFT_16, SHIFT_24 = 1 << 8, 1 << 16, 1 << 24


def i64_to_canonical(c: int) -> int:
    return c % P


def neg_canonical(x: int) -> int:
    return 0 if x == 0 else P - x


class Packing(enum.Enum):
    """How consecutive trace columns combine into bus elements (lookup.rs:150-210)."""
    DIRECT = "Direct"      # 1 column -> 1 element
    WORD2L = "Word2L"      # h0 + 2^16 h1
    WORD4L = "Word4L"      # b0 + 2^8 b1 + 2^16 b2 + 2^24 b3
    DWORD_WL = "DWordWL"   # 2 x Direct
    DWORD_HHW = "DWordHHW"  # Direct + Word2L
    DWORD_WHH = "DWordWHH"  # Word2L + Direct
    DWORD_HL = "DWordHL"   # 2 x Word2L
    DWORD_BL = "DWordBL"   # 2 x Word4L
    QUAD_HL = "QuadHL"     # 4 x Word2L
    QUAD_WL = "QuadWL"     # 4 x Direct

    def num_columns(self) -> int:
        return {"Direct": 1, "Word2L": 2, "Word4L": 4, "DWordWL": 2, "DWordHHW": 3, "DWordWHH": 3,
                "DWordHL": 4, "DWordBL": 8, "QuadHL": 8, "QuadWL": 4}[self.value]

    def num_bus_elements(self) -> int:
        return {"Direct": 1, "Word2L": 1, "Word4L": 1, "DWordWL": 2, "DWordHHW": 2, "DWordWHH": 2,
                "DWordHL": 2, "DWordBL": 2, "QuadHL": 4, "QuadWL": 4}[self.value]

    def elements(self, c: int) -> list[list[tuple[int, int]]]:
        """The bus elements of this packing at start column c, each a list of
        (coefficient, column) terms (logup_gpu.rs:172-250)."""
        w2 = lambda x: [(1, x), (SHIFT_16, x + 1)]  # noqa: E731
        w4 = lambda x: [(1, x), (SHIFT_8, x + 1), (SHIFT_16, x + 2), (SHIFT_24, x + 3)]  # noqa: E731
        v = self.value
        if v == "Direct":
            return [[(1, c)]]
        if v == "Word2L":
            return [w2(c)]
        if v == "Word4L":
            return [w4(c)]
        if v == "DWordWL":
            return [[(1, c)], [(1, c + 1)]]
        if v == "DWordHHW":
            return [[(1, c)], w2(c + 1)]
        if v == "DWordWHH":
            return [w2(c), [(1, c + 2)]]
        if v == "DWordHL":
            return [w2(c), w2(c + 2)]
        if v == "DWordBL":
            return [w4(c), w4(c + 4)]
        if v == "QuadHL":
            return [w2(c + 2 * i) for i in range(4)]
        if v == "QuadWL":
            return [[(1, c + i)] for i in range(4)]
        raise ValueError(v)


@dataclass(frozen=True)
class LinearTerm:
    """`LinearTerm::{Column, ColumnUnsigned, Constant}`: column given => a
    coefficient * column term, else a constant."""
    coefficient: int
    column: int | None = None
    unsigned: bool = False


@dataclass(frozen=True)
class BusValue:
    """Either a packed run of columns or a linear form (`BusValue::{Packed, Linear}`)."""
    packing: Packing | None = None
    start_column: int = 0
    linear: tuple[LinearTerm, ...] = ()

    @staticmethod
    def packed(packing: Packing, start_column: int) -> "BusValue":
        return BusValue(packing=packing, start_column=start_column)

    @staticmethod
    def column(c: int) -> "BusValue":
        return BusValue(packing=Packing.DIRECT, start_column=c)

    @staticmethod
    def of_linear(terms) -> "BusValue":
        return BusValue(linear=tuple(terms))

    def num_bus_elements(self) -> int:
        return self.packing.num_bus_elements() if self.packing is not None else 1


@dataclass(frozen=True)
class Multiplicity:
    """`Multiplicity::{One, Column, Sum, Negated, Diff, Sum3, Linear}`."""
    kind: str = "One"
    cols: tuple[int, ...] = ()
    linear: tuple[LinearTerm, ...] = ()


ONE = Multiplicity("One")


@dataclass
class BusInteraction:
    bus_id: int
    is_sender: bool
    multiplicity: Multiplicity = ONE
    values: list[BusValue] = field(default_factory=list)

    @staticmethod
    def sender(bus_id, multiplicity, values):
        return BusInteraction(bus_id, True, multiplicity, list(values))

    @staticmethod
    def receiver(bus_id, multiplicity, values):
        return BusInteraction(bus_id, False, multiplicity, list(values))

    def num_bus_elements(self) -> int:
        """1 (the bus id) + the values' elements (lookup.rs:1546-1552)."""
        return 1 + sum(v.num_bus_elements() for v in self.values)


def split_interactions(k: int) -> tuple[int, int]:
    """(committed pairs, absorbed count) for k interactions (lookup.rs:117-125)."""
    if k <= 2:
        return 0, k
    if k % 2 == 1:
        return (k - 1) // 2, 1
    return (k - 2) // 2, 2


def encode_signed_multiplicity(m: Multiplicity, is_sender: bool):
    """(constant, [(coef, col)]) with the receiver sign folded in (logup_gpu.rs:54-98)."""
    cst = 0
    terms: list[tuple[int, int]] = []
    if m.kind == "One":
        cst = 1
    elif m.kind == "Column":
        terms.append((1, m.cols[0]))
    elif m.kind == "Sum":
        terms += [(1, m.cols[0]), (1, m.cols[1])]
    elif m.kind == "Negated":
        cst = 1
        terms.append((neg_canonical(1), m.cols[0]))
    elif m.kind == "Diff":
        terms += [(1, m.cols[0]), (neg_canonical(1), m.cols[1])]
    elif m.kind == "Sum3":
        terms += [(1, c) for c in m.cols[:3]]
    elif m.kind == "Linear":
        for t in m.linear:
            if t.column is None:
                cst += i64_to_canonical(t.coefficient)
            else:
                terms.append(((t.coefficient % P) if t.unsigned else i64_to_canonical(t.coefficient), t.column))
    else:
        raise ValueError(m.kind)
    cst %= P
    if not is_sender:
        cst = neg_canonical(cst)
        terms = [(neg_canonical(c), col) for c, col in terms]
    return cst, terms


@dataclass
class FingerprintDescriptor:
    """Flat CSR descriptor (logup_gpu.rs:100-142); all coefficients canonical."""
    num_interactions: int
    alpha_powers_len: int
    bus_ids: np.ndarray
    elem_offsets: np.ndarray
    elem_alpha_idx: np.ndarray
    elem_const: np.ndarray
    term_offsets: np.ndarray
    term_coef: np.ndarray
    term_col: np.ndarray
    num_out_cols: int
    out_col_offsets: np.ndarray
    out_col_interactions: np.ndarray
    mult_const: np.ndarray
    mult_term_offsets: np.ndarray
    mult_term_coef: np.ndarray
    mult_term_col: np.ndarray

    @property
    def num_committed(self) -> int:
        return self.num_out_cols - 1

    def max_column(self) -> int:
        cols = np.concatenate([self.term_col, self.mult_term_col]).astype(np.int64)
        return int(cols.max()) if cols.size else -1


def build_fingerprint_descriptor(interactions: list[BusInteraction]) -> FingerprintDescriptor:
    """Compile a table's interactions (logup_gpu.rs:256-290): element alpha
    indices start at 1 (alpha^0 is the bus id); committed pair p = interactions
    {2p, 2p+1}; the trailing 1-2 absorbed interactions form the last column."""
    bus_ids, elem_offsets, elem_alpha, elem_const = [], [0], [], []
    term_offsets, term_coef, term_col = [0], [], []
    mult_const, mult_off, mult_coef, mult_col = [], [0], [], []
    max_elems = 0
    for it in interactions:
        bus_ids.append(it.bus_id % P)
        max_elems = max(max_elems, it.num_bus_elements())
        a = 1
        for bv in it.values:
            if bv.packing is not None:
                for j, el in enumerate(bv.packing.elements(bv.start_column)):
                    elem_alpha.append(a + j)
                    elem_const.append(0)
                    for coef, col in el:
                        term_coef.append(coef)
                        term_col.append(col)
                    term_offsets.append(len(term_coef))
                a += bv.packing.num_bus_elements()
            else:
                cst = 0
                for t in bv.linear:
                    if t.column is None:
                        cst += i64_to_canonical(t.coefficient)
                    else:
                        term_coef.append((t.coefficient % P) if t.unsigned else i64_to_canonical(t.coefficient))
                        term_col.append(t.column)
                elem_alpha.append(a)
                elem_const.append(cst % P)
                term_offsets.append(len(term_coef))
                a += 1
        elem_offsets.append(len(elem_alpha))
        cst, terms = encode_signed_multiplicity(it.multiplicity, it.is_sender)
        mult_const.append(cst)
        for coef, col in terms:
            mult_coef.append(coef)
            mult_col.append(col)
        mult_off.append(len(mult_coef))
    pairs, absorbed = split_interactions(len(interactions))
    out_off, out_int = [0], []
    for p in range(pairs):
        out_int += [2 * p, 2 * p + 1]
        out_off.append(len(out_int))
    out_int += list(range(len(interactions) - absorbed, len(interactions)))
    out_off.append(len(out_int))
    u64 = lambda x: np.array(x, dtype=np.uint64)  # noqa: E731
    u32 = lambda x: np.array(x, dtype=np.uint32)  # noqa: E731
    return FingerprintDescriptor(
        num_interactions=len(interactions), alpha_powers_len=max_elems,
        bus_ids=u64(bus_ids), elem_offsets=u32(elem_offsets), elem_alpha_idx=u32(elem_alpha),
        elem_const=u64(elem_const), term_offsets=u32(term_offsets), term_coef=u64(term_coef),
        term_col=u32(term_col), num_out_cols=pairs + 1, out_col_offsets=u32(out_off),
        out_col_interactions=u32(out_int), mult_const=u64(mult_const), mult_term_offsets=u32(mult_off),
        mult_term_coef=u64(mult_coef), mult_term_col=u32(mult_col))
