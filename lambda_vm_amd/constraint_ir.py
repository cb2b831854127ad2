"""Host-side constraint IR: capture, and lowering to the flat device program.

Mirrors the reference's constraint IR (crypto/stark/src/constraint_ir/ir.rs
`Op`/`Dim`/`ConstraintProgram`, builder.rs `IrBuilder`: hash-consed nodes,
node 0 = base zero, constants deduplicated into side tables) and its device
lowering (constraint_ir/device.rs:130-330 `DeviceProgram::lower`): 16-byte
nodes (op, a, b, res), operands encoded kind << 29 | payload, dim-split
base/ext value slots reused at an operand's last use, row-invariant leaves
(constants, challenges, alpha powers, the table offset) referenced directly
instead of materialised, root slots pinned. The lowered program is what
crosses the C ABI (lvm_program); its wire tags match device.rs:60-110.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

P = 0xFFFFFFFF00000001

# op tags (device.rs:60-85)
OP_CONST_BASE, OP_CONST_EXT, OP_VAR, OP_RAP_CHALLENGE, OP_ALPHA_POW, OP_TABLE_OFFSET = 0, 1, 2, 3, 4, 5
OP_ADD, OP_SUB, OP_MUL, OP_NEG, OP_EMBED = 6, 7, 8, 9, 10
# operand kinds (device.rs:88-104)
OPK_SHIFT = 29
OPK_PAYLOAD_MASK = (1 << OPK_SHIFT) - 1
OPK_BASE_SLOT, OPK_EXT_SLOT, OPK_BASE_CONST, OPK_EXT_CONST, OPK_RAP, OPK_ALPHA, OPK_OFFSET = range(7)
RES_EXT_BIT = 1 << 31

BASE, EXT = "base", "ext"
_UNIFORM = {"const_base", "const_ext", "rap", "alpha", "offset"}


def pack_var(main: bool, offset: int, row: int, col: int) -> tuple[int, int]:
    """(a, b) words of a Var leaf (device.rs:118-124)."""
    return col & 0xFFFF, (int(main) << 16) | ((offset & 0xFF) << 8) | (row & 0xFF)


class Expr:
    """A node handle; arithmetic operators append hash-consed nodes."""
    __slots__ = ("b", "id", "dim")

    def __init__(self, b: "IrBuilder", id_: int, dim: str):
        self.b, self.id, self.dim = b, id_, dim

    def _lift(self, o):
        return o if isinstance(o, Expr) else self.b.const_base(o)

    def __add__(self, o):
        return self.b.add(self, self._lift(o))

    def __radd__(self, o):
        return self.b.add(self._lift(o), self)

    def __sub__(self, o):
        return self.b.sub(self, self._lift(o))

    def __rsub__(self, o):
        return self.b.sub(self._lift(o), self)

    def __mul__(self, o):
        return self.b.mul(self, self._lift(o))

    def __rmul__(self, o):
        return self.b.mul(self._lift(o), self)

    def __neg__(self):
        return self.b.neg(self)


@dataclass
class ConstraintProgram:
    nodes: list            # tuples: (kind, ...)
    dims: list
    base_consts: list
    ext_consts: list
    roots: list
    num_base: int


class IrBuilder:
    """Capture front-end (builder.rs): node 0 is the base zero."""

    def __init__(self):
        self.nodes, self.dims, self.cse = [], [], {}
        self.base_consts, self.ext_consts, self.roots = [], [], []
        z = self.const_base(0)
        assert z.id == 0

    def _push(self, op: tuple, dim: str) -> Expr:
        key = (op, dim)
        if key in self.cse:
            return Expr(self, self.cse[key], dim)
        i = len(self.nodes)
        self.nodes.append(op)
        self.dims.append(dim)
        self.cse[key] = i
        return Expr(self, i, dim)

    # leaves
    def main(self, offset: int, col: int) -> Expr:
        return self._push(("var", True, offset, 0, col), BASE)

    def aux(self, offset: int, col: int) -> Expr:
        return self._push(("var", False, offset, 0, col), EXT)

    def challenge(self, idx: int) -> Expr:
        return self._push(("rap", idx), EXT)

    def alpha_power(self, idx: int) -> Expr:
        return self._push(("alpha", idx), EXT)

    def table_offset(self) -> Expr:
        return self._push(("offset",), EXT)

    def const_base(self, v: int) -> Expr:
        v %= P
        if v not in self.base_consts:
            self.base_consts.append(v)
        return self._push(("const_base", self.base_consts.index(v)), BASE)

    def const_ext(self, v) -> Expr:
        v = tuple(int(x) % P for x in v)
        if v not in self.ext_consts:
            self.ext_consts.append(v)
        return self._push(("const_ext", self.ext_consts.index(v)), EXT)

    # arithmetic
    @staticmethod
    def _join(a: str, b: str) -> str:
        return BASE if a == BASE and b == BASE else EXT

    def add(self, a: Expr, b: Expr) -> Expr:
        return self._push(("add", a.id, b.id), self._join(a.dim, b.dim))

    def sub(self, a: Expr, b: Expr) -> Expr:
        return self._push(("sub", a.id, b.id), self._join(a.dim, b.dim))

    def mul(self, a: Expr, b: Expr) -> Expr:
        return self._push(("mul", a.id, b.id), self._join(a.dim, b.dim))

    def neg(self, a: Expr) -> Expr:
        return self._push(("neg", a.id), a.dim)

    def embed(self, a: Expr) -> Expr:
        return self._push(("embed", a.id), EXT)

    def emit(self, idx: int, e: Expr) -> None:
        while len(self.roots) <= idx:
            self.roots.append(0)
        self.roots[idx] = e.id

    def finish(self, num_base: int) -> ConstraintProgram:
        return ConstraintProgram(self.nodes, self.dims, self.base_consts, self.ext_consts, self.roots, num_base)


def _operands(op: tuple) -> list[int]:
    if op[0] in ("add", "sub", "mul"):
        return [op[1], op[2]]
    if op[0] in ("neg", "embed"):
        return [op[1]]
    return []


@dataclass
class DeviceProgram:
    """Lowered program (device.rs:126-160): nodes as (op, a, b, res) u32 rows."""
    nodes: np.ndarray        # (num_nodes, 4) uint32
    base_consts: np.ndarray  # uint64
    ext_consts: np.ndarray   # (k, 3) uint64
    roots: np.ndarray        # uint32 (slot | RES_EXT_BIT)
    num_base: int
    num_base_slots: int
    num_ext_slots: int
    max_offset: int          # largest frame offset a Var reads
    root_nodes: np.ndarray = None  # per constraint: index (in `nodes`) of the node computing it

    def packed_nodes(self) -> np.ndarray:
        """2 u64 per node: op | a << 32, b | res << 32 (constraint_ir/gpu_interp.rs:30-38)."""
        n = self.nodes.astype(np.uint64)
        out = np.empty(2 * len(n), dtype=np.uint64)
        out[0::2] = n[:, 0] | (n[:, 1] << np.uint64(32))
        out[1::2] = n[:, 2] | (n[:, 3] << np.uint64(32))
        return out


def lower(prog: ConstraintProgram, pin_roots: bool = True) -> DeviceProgram:
    """DeviceProgram::lower (device.rs:200-330). With pin_roots=False a
    constraint's slot is reused once its value has been consumed: the consumer
    must take each constraint value when its node runs (root_nodes), which the
    GPU composition kernel does; the slot set is then the program's live set."""
    n = len(prog.nodes)
    assert n <= OPK_PAYLOAD_MASK
    used, last_use = [False] * n, [0] * n
    for i, op in enumerate(prog.nodes):
        for j in _operands(op):
            used[j] = True
            last_use[j] = i
    is_root = [False] * n
    for r in prog.roots:
        is_root[r] = True
        used[r] = True
    emitted = [used[i] and (prog.nodes[i][0] not in _UNIFORM or is_root[i]) for i in range(n)]

    def enc_uniform(op):
        k = op[0]
        if k == "const_base":
            return (OPK_BASE_CONST << OPK_SHIFT) | op[1]
        if k == "const_ext":
            return (OPK_EXT_CONST << OPK_SHIFT) | op[1]
        if k == "rap":
            return (OPK_RAP << OPK_SHIFT) | op[1]
        if k == "alpha":
            return (OPK_ALPHA << OPK_SHIFT) | op[1]
        if k == "offset":
            return OPK_OFFSET << OPK_SHIFT
        raise AssertionError(op)

    slot_of = [None] * n
    node_index, root_slot = [None] * n, [None] * n
    free_base, free_ext = [], []
    nb = ne = 0
    rows = []
    max_offset = 0
    for i in range(n):
        if not emitted[i]:
            continue
        op, dim = prog.nodes[i], prog.dims[i]

        def enc(j):
            if not emitted[j]:
                return enc_uniform(prog.nodes[j])
            s = slot_of[j]
            assert s is not None, "operand before definition"
            return ((OPK_BASE_SLOT if prog.dims[j] == BASE else OPK_EXT_SLOT) << OPK_SHIFT) | s

        k = op[0]
        if k == "const_base":
            tag, a, b = OP_CONST_BASE, op[1], 0
        elif k == "const_ext":
            tag, a, b = OP_CONST_EXT, op[1], 0
        elif k == "var":
            a, b = pack_var(op[1], op[2], op[3], op[4])
            tag = OP_VAR
            max_offset = max(max_offset, op[2])
        elif k == "rap":
            tag, a, b = OP_RAP_CHALLENGE, op[1], 0
        elif k == "alpha":
            tag, a, b = OP_ALPHA_POW, op[1], 0
        elif k == "offset":
            tag, a, b = OP_TABLE_OFFSET, 0, 0
        elif k in ("add", "sub", "mul"):
            tag = {"add": OP_ADD, "sub": OP_SUB, "mul": OP_MUL}[k]
            a, b = enc(op[1]), enc(op[2])
        elif k == "neg":
            tag, a, b = OP_NEG, enc(op[1]), 0
        else:
            tag, a, b = OP_EMBED, enc(op[1]), 0
        # free operand slots at their last use (before allocating the result)
        for j in _operands(op):
            if emitted[j] and (not is_root[j] or not pin_roots) and last_use[j] == i and slot_of[j] is not None:
                (free_base if prog.dims[j] == BASE else free_ext).append(slot_of[j])
                slot_of[j] = None
        if dim == BASE:
            if free_base:
                s = free_base.pop()
            else:
                s, nb = nb, nb + 1
        else:
            if free_ext:
                s = free_ext.pop()
            else:
                s, ne = ne, ne + 1
        slot_of[i] = s
        node_index[i] = len(rows)
        rows.append((tag, a, b, s if dim == BASE else (s | RES_EXT_BIT)))
        root_slot[i] = s if dim == BASE else (s | RES_EXT_BIT)
        if is_root[i] and not pin_roots and last_use[i] <= i:
            # consumed at emission and never read again
            (free_base if dim == BASE else free_ext).append(s)
            slot_of[i] = None
    roots = [root_slot[r] for r in prog.roots]
    root_nodes = [node_index[r] for r in prog.roots]
    return DeviceProgram(
        nodes=np.array(rows, dtype=np.uint32).reshape(-1, 4),
        base_consts=np.array(prog.base_consts, dtype=np.uint64),
        ext_consts=np.array(prog.ext_consts, dtype=np.uint64).reshape(-1, 3),
        roots=np.array(roots, dtype=np.uint32),
        num_base=prog.num_base, num_base_slots=nb, num_ext_slots=ne, max_offset=max_offset,
        root_nodes=np.array(root_nodes, dtype=np.uint32))
