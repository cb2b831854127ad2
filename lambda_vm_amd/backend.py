"""Host-side mirror of the reference GPU backend surface (crypto/math-cuda)
over the C ABI of liblvm_gpu.so.

Names, argument meaning and error behaviour follow the reference:
  math_cuda::lde::coset_lde_row_major_with_merkle_tree_keep
      (crypto/math-cuda/src/lde.rs:674)          -> coset_lde_row_major_with_merkle_tree_keep
  math_cuda::lde::coset_lde_ext3_row_major_with_merkle_tree_keep
      (crypto/math-cuda/src/lde.rs:855)          -> coset_lde_ext3_row_major_with_merkle_tree_keep
  math_cuda::ntt::{forward, inverse} (ntt.rs:50, :57) -> ntt_forward, ntt_inverse
  math_cuda::merkle::build_merkle_tree_on_device (merkle.rs:316) -> build_merkle_tree
  math_cuda::merkle::gather_merkle_paths_dev (merkle.rs:358) -> GpuLde.open
  math_cuda::{gl_add, gl_mul, ext3_mul, ...} (lib.rs:34-197) -> gl_op, fp3_op
A failing call raises LvmError, the counterpart of the reference's
`Err(DriverError)` that its `try_*` callers turn into `None` (CPU arm).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from ._native import LvmError  # noqa: F401  (raised by every failing call)

OPS = {"add": 0, "sub": 1, "mul": 2, "neg": 3, "inv": 4, "sqr": 5}


def _u64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint64)


def _p64(a: np.ndarray):
    return a.ctypes.data_as(N.u64p)


def _p8(a: np.ndarray):
    return a.ctypes.data_as(N.u8p)


def gl_op(op: str, a, b=None) -> np.ndarray:
    N.ensure_device()
    a = _u64(a)
    bb = _u64(b) if b is not None else None
    out = np.empty_like(a)
    N.call("lvm_gl_op", OPS[op], _p64(a), _p64(bb) if bb is not None else None, _p64(out), a.size)
    return out


def fp3_op(op: str, a, b=None) -> np.ndarray:
    N.ensure_device()
    a = _u64(a).reshape(-1, 3)
    bb = _u64(b).reshape(-1, 3) if b is not None else None
    out = np.empty_like(a)
    N.call("lvm_fp3_op", OPS[op], _p64(a), _p64(bb) if bb is not None else None, _p64(out), a.shape[0])
    return out


def ntt_forward(x) -> np.ndarray:
    """Evaluations of each column at w_n^i (natural order); x: (n,) or (n, m)."""
    N.ensure_device()
    a = _u64(x)
    shape = a.shape
    a2 = a.reshape(shape[0], -1)
    out = np.empty_like(a2)
    N.call("lvm_ntt_forward", _p64(a2), _p64(out), a2.shape[0], a2.shape[1])
    return out.reshape(shape)


def ntt_inverse(x) -> np.ndarray:
    """Unnormalised inverse (no 1/n), natural order in and out."""
    N.ensure_device()
    a = _u64(x)
    shape = a.shape
    a2 = a.reshape(shape[0], -1)
    out = np.empty_like(a2)
    N.call("lvm_ntt_inverse", _p64(a2), _p64(out), a2.shape[0], a2.shape[1])
    return out.reshape(shape)


class GpuLde:
    """Device-resident LDE table + its Keccak Merkle tree (GpuLdeBase/GpuLdeExt3
    + GpuMerkleTree, crypto/math-cuda/src/lde.rs:923-988)."""

    def __init__(self, handle: int):
        self.h = C.c_void_p(handle)
        rows, w, n = C.c_size_t(), C.c_size_t(), C.c_size_t()
        N.call("lvm_lde_shape", self.h, C.byref(rows), C.byref(w), C.byref(n))
        self.lde_rows, self.words_per_row, self.trace_rows = rows.value, w.value, n.value
        r = np.zeros(32, dtype=np.uint8)
        N.call("lvm_lde_root", self.h, _p8(r))
        self.root = r.tobytes()

    def close(self) -> None:
        if self.h is not None and self.h.value:
            N.lib().lvm_lde_release(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def download(self) -> np.ndarray:
        """Natural-order row-major host copy, (lde_rows, words_per_row)."""
        out = np.empty((self.lde_rows, self.words_per_row), dtype=np.uint64)
        N.call("lvm_lde_download", self.h, _p64(out))
        return out

    def device_ptr(self) -> int:
        p = C.c_void_p()
        N.call("lvm_lde_device_ptr", self.h, C.byref(p))
        return p.value

    def nodes(self) -> np.ndarray:
        cnt = C.c_size_t()
        N.call("lvm_lde_tree_nodes", self.h, None, C.byref(cnt))
        out = np.zeros((cnt.value, 32), dtype=np.uint8)
        N.call("lvm_lde_tree_nodes", self.h, _p8(out), C.byref(cnt))
        return out

    def open(self, positions):
        """Auth paths (count, depth, 32) and the two opened rows per position
        (count, 2, words_per_row): rows br(2p), br(2p+1) of the natural LDE."""
        pos = _u64(positions)
        L = self.lde_rows // 2
        depth = L.bit_length() - 1
        paths = np.zeros((pos.size, max(depth, 1), 32), dtype=np.uint8)
        rows = np.zeros((pos.size, 2, self.words_per_row), dtype=np.uint64)
        N.call("lvm_lde_open", self.h, _p64(pos), pos.size, _p8(paths), _p64(rows))
        return paths[:, :depth], rows


def coset_lde_row_major_with_merkle_tree_keep(row_major, n: int, m: int, blowup_factor: int, weights,
                                              retain_host_lde: bool = True):
    """Row-major n x m trace -> (GpuLde handle, natural-order row-major LDE or None)."""
    N.ensure_device()
    tr = _u64(row_major).reshape(-1)
    if tr.size != n * m:
        raise ValueError("row_major must hold n*m elements")
    w = _u64(weights)
    if w.size != n:
        raise ValueError("weights length must match n")
    root = np.zeros(32, dtype=np.uint8)
    host = np.empty((n * blowup_factor, m), dtype=np.uint64) if retain_host_lde else None
    h = C.c_void_p()
    N.call("lvm_coset_lde_row_major_with_merkle_tree_keep", _p64(tr), n, m, blowup_factor, _p64(w),
           _p64(host) if host is not None else None, _p8(root), C.byref(h))
    return GpuLde(h.value), host


def coset_lde_ext3_row_major_with_merkle_tree_keep(row_major, n: int, m: int, blowup_factor: int, weights,
                                                   retain_host_lde: bool = True):
    """Row-major n x m Fp3 trace (3m u64 per row) -> (GpuLde, host LDE (N, 3m) or None)."""
    N.ensure_device()
    tr = _u64(row_major).reshape(-1)
    if tr.size != n * m * 3:
        raise ValueError("row_major must hold n*m Fp3 elements")
    w = _u64(weights)
    root = np.zeros(32, dtype=np.uint8)
    host = np.empty((n * blowup_factor, 3 * m), dtype=np.uint64) if retain_host_lde else None
    h = C.c_void_p()
    N.call("lvm_coset_lde_ext3_row_major_with_merkle_tree_keep", _p64(tr), n, m, blowup_factor, _p64(w),
           _p64(host) if host is not None else None, _p8(root), C.byref(h))
    return GpuLde(h.value), host


class DeviceBuffer:
    """Raw device allocation for resident pipelines (the reference's `predev`)."""

    def __init__(self, nbytes: int):
        N.ensure_device()
        p = C.c_void_p()
        N.call("lvm_dev_alloc", C.byref(p), nbytes)
        self.ptr = p.value
        self.nbytes = nbytes

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        N.call("lvm_memcpy_htod", C.c_void_p(b.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes)
        return b

    def to_array(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        N.call("lvm_memcpy_dtoh", out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), out.nbytes)
        return out

    def free(self) -> None:
        if self.ptr:
            N.lib().lvm_dev_free(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def coset_lde_dev(trace: DeviceBuffer, n: int, m: int, blowup_factor: int, weights: DeviceBuffer) -> GpuLde:
    """Device-resident row-major trace + weights -> GpuLde (no host copies)."""
    N.ensure_device()
    root = np.zeros(32, dtype=np.uint8)
    h = C.c_void_p()
    N.call("lvm_coset_lde_dev", C.c_void_p(trace.ptr), n, m, blowup_factor, C.c_void_p(weights.ptr), _p8(root),
           C.byref(h))
    return GpuLde(h.value)


def coset_lde_split_trees_dev(trace: DeviceBuffer, n: int, m: int, blowup_factor: int, weights: DeviceBuffer,
                              split_col: int, want_precomputed_nodes: bool = True):
    """A preprocessed table's commitment (crypto/math-cuda/src/lde.rs:726):
    (GpuLde with the multiplicity-columns tree, precomputed root, precomputed
    node array (2L-1, 32) or None)."""
    N.ensure_device()
    root = np.zeros(32, dtype=np.uint8)
    pre_root = np.zeros(32, dtype=np.uint8)
    L = n * blowup_factor // 2
    pre_nodes = np.zeros((2 * L - 1, 32), dtype=np.uint8) if want_precomputed_nodes else None
    h = C.c_void_p()
    N.call("lvm_coset_lde_split_trees_dev", C.c_void_p(trace.ptr), n, m, blowup_factor, C.c_void_p(weights.ptr),
           split_col, pre_nodes.ctypes.data_as(C.c_void_p) if pre_nodes is not None else None, _p8(pre_root),
           _p8(root), C.byref(h))
    return GpuLde(h.value), pre_root.tobytes(), pre_nodes


def merkle_paths_from_nodes(nodes: np.ndarray, positions) -> np.ndarray:
    """Authentication paths (Q, depth, 32) of leaves `positions` in a host
    node array (2L-1, 32) (crypto/crypto/src/merkle_tree/utils.rs:7-21)."""
    L = (nodes.shape[0] + 1) // 2
    depth = L.bit_length() - 1
    out = np.zeros((len(positions), depth, 32), dtype=np.uint8)
    for q, pos in enumerate(positions):
        idx = int(pos) + L - 1
        for k in range(depth):
            sib = idx - 1 if idx % 2 == 0 else idx + 1
            out[q, k] = nodes[sib]
            idx = (idx - 1) // 2 if idx % 2 == 0 else idx // 2
    return out


def dev_copy(dst: DeviceBuffer, src: DeviceBuffer, nbytes: int) -> None:
    """Device-to-device copy by the library's plain copy kernel."""
    N.call("lvm_dev_copy", C.c_void_p(dst.ptr), C.c_void_p(src.ptr), nbytes)


def coset_weights_dev(n: int, offset: int) -> DeviceBuffer:
    """Device buffer of w_i = n^-1 * offset^i (prover.rs:451-467)."""
    b = DeviceBuffer(8 * n)
    N.call("lvm_coset_weights", n, offset, C.c_void_p(b.ptr))
    return b


def _descs(tables) -> "C.Array":
    """(trace DeviceBuffer, weights DeviceBuffer, n, m, blowup) tuples -> lvm_table_desc[]."""
    arr = (N.TableDesc * len(tables))()
    for i, (tr, w, n, m, b) in enumerate(tables):
        arr[i] = N.TableDesc(tr.ptr, w.ptr, n, m, b)
    return arr


class BatchPlan:
    """Batched ragged commitment of many tables (lvm_batch_*): one launch per
    stage step across all tables, optionally replayed as a captured HIP graph.
    Each table's result equals coset_lde_dev on it."""

    def __init__(self, tables, use_graph: bool = True):
        N.ensure_device()
        self.count = len(tables)
        self._keep = list(tables)
        descs = _descs(tables)
        h = C.c_void_p()
        N.call("lvm_batch_new", C.cast(descs, C.c_void_p), self.count, int(use_graph), C.byref(h))
        self.h = h

    def bind(self, tables) -> None:
        """Point the plan at other inputs of the same shapes."""
        descs = _descs(tables)
        N.call("lvm_batch_bind", self.h, C.cast(descs, C.c_void_p), len(tables))
        self._keep = list(tables)

    def run(self) -> list[bytes]:
        roots = np.zeros((self.count, 32), dtype=np.uint8)
        N.call("lvm_batch_run", self.h, _p8(roots))
        return [bytes(r) for r in roots]

    def launches(self) -> list[tuple[str, float, int]]:
        """(profiler name, compulsory HBM bytes, blocks) of each launch in order."""
        cnt = C.c_size_t()
        N.call("lvm_batch_launch_count", self.h, C.byref(cnt))
        out = []
        for i in range(cnt.value):
            name = C.create_string_buffer(64)
            by, bl = C.c_double(), C.c_uint32()
            N.call("lvm_batch_launch_info", self.h, i, name, 64, C.byref(by), C.byref(bl))
            out.append((name.value.decode(), by.value, bl.value))
        return out

    def take(self, index: int) -> GpuLde:
        h = C.c_void_p()
        N.call("lvm_batch_take", self.h, index, C.byref(h))
        return GpuLde(h.value)

    def close(self) -> None:
        if self.h is not None and self.h.value:
            N.lib().lvm_batch_release(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_lde_commit(tables) -> tuple[list[bytes], list[GpuLde]]:
    """One-shot batched commitment: roots and the resident handle of each table."""
    N.ensure_device()
    descs = _descs(tables)
    roots = np.zeros((len(tables), 32), dtype=np.uint8)
    hs = (C.c_void_p * len(tables))()
    N.call("lvm_batch_lde_commit", C.cast(descs, C.c_void_p), len(tables), _p8(roots),
           C.cast(hs, N.vpp))
    return [bytes(r) for r in roots], [GpuLde(h) for h in hs]


def set_ntt_plan(log_n: int, radices) -> None:
    """Install a pass split for size 2^log_n (empty restores the default)."""
    arr = (C.c_int * max(1, len(radices)))(*radices)
    N.call("lvm_set_ntt_plan", log_n, arr, len(radices))


def get_ntt_plan(log_n: int) -> list[int]:
    arr = (C.c_int * 8)()
    cnt = C.c_int()
    N.call("lvm_get_ntt_plan", log_n, arr, C.byref(cnt))
    return list(arr[:cnt.value])


_LOGUP_ARRAYS = [("bus_ids", np.uint64), ("elem_offsets", np.uint32), ("elem_alpha_idx", np.uint32),
                 ("elem_const", np.uint64), ("term_offsets", np.uint32), ("term_coef", np.uint64),
                 ("term_col", np.uint32), ("out_col_offsets", np.uint32), ("out_col_interactions", np.uint32),
                 ("mult_const", np.uint64), ("mult_term_offsets", np.uint32), ("mult_term_coef", np.uint64),
                 ("mult_term_col", np.uint32)]


def _logup_desc(d):
    """ctypes lvm_logup_desc over a FingerprintDescriptor (arrays kept alive in the result)."""
    s = N.LogupDesc()
    s.num_interactions = d.num_interactions
    s.num_out_cols = d.num_out_cols
    keep = []
    for name, dt in _LOGUP_ARRAYS:
        a = np.ascontiguousarray(getattr(d, name), dtype=dt)
        if a.size == 0:
            a = np.zeros(1, dtype=dt)
        keep.append(a)
        setattr(s, name, a.ctypes.data)
    return s, keep


def logup_aux_dev(main: DeviceBuffer, num_rows: int, num_cols: int, desc, alpha_powers, z,
                  column_major: bool = False, want_terms: bool = False):
    """Resident LogUp aux build (crypto/math-cuda/src/logup.rs:311
    `logup_aux_resident`). main: device trace, row-major (num_rows, num_cols)
    or column-major. Returns (aux DeviceBuffer of num_rows x num_out_cols Fp3,
    row-major; L as 3 ints; term columns DeviceBuffer or None)."""
    N.ensure_device()
    s, keep = _logup_desc(desc)
    al = _u64(alpha_powers).reshape(-1)
    zz = _u64(z).reshape(3)
    C_ = desc.num_out_cols
    aux = DeviceBuffer(num_rows * C_ * 24)
    terms = DeviceBuffer(num_rows * C_ * 24) if want_terms else None
    L = np.zeros(3, dtype=np.uint64)
    rs, cs = (1, num_rows) if column_major else (num_cols, 1)
    N.call("lvm_logup_aux_dev", C.c_void_p(main.ptr), rs, cs, num_rows, num_cols, C.byref(s), _p64(al),
           al.size // 3, _p64(zz), C.c_void_p(aux.ptr), C.c_void_p(terms.ptr) if terms else None, _p64(L))
    return aux, [int(x) for x in L], terms


def logup_term_columns(main_cols, desc, alpha_powers, z) -> np.ndarray:
    """Host column-major main (num_cols, num_rows) -> term columns (num_out_cols,
    num_rows, 3) (crypto/math-cuda/src/logup.rs:133 `logup_term_columns`)."""
    N.ensure_device()
    mc = _u64(main_cols)
    num_cols, num_rows = mc.shape
    s, keep = _logup_desc(desc)
    al = _u64(alpha_powers).reshape(-1)
    zz = _u64(z).reshape(3)
    out = np.zeros((desc.num_out_cols, num_rows, 3), dtype=np.uint64)
    N.call("lvm_logup_term_columns", _p64(mc), num_rows, num_cols, C.byref(s), _p64(al), al.size // 3, _p64(zz),
           _p64(out))
    return out


def _arr_ptr(x, dt, keep):
    a = np.ascontiguousarray(x if x is not None and len(x) else [0], dtype=dt).reshape(-1)
    keep.append(a)
    return a.ctypes.data


def _program_struct(dev, rap, alpha, table_offset, keep):
    """lvm_program over a lowered DeviceProgram (lambda_vm_amd.constraint_ir)."""
    s = N.Program()
    s.nodes = _arr_ptr(dev.packed_nodes(), np.uint64, keep)
    s.num_nodes = len(dev.nodes)
    s.num_base_slots = dev.num_base_slots
    s.num_ext_slots = dev.num_ext_slots
    s.base_consts = _arr_ptr(dev.base_consts, np.uint64, keep)
    s.num_base_consts = len(dev.base_consts)
    ec = np.asarray(dev.ext_consts, dtype=np.uint64).reshape(-1)
    s.ext_consts = _arr_ptr(ec, np.uint64, keep)
    s.num_ext_consts = ec.size // 3
    s.roots = _arr_ptr(np.asarray(dev.roots, dtype=np.uint64), np.uint64, keep)
    s.num_roots = len(dev.roots)
    r = np.asarray(rap if rap is not None else [], dtype=np.uint64).reshape(-1)
    a = np.asarray(alpha if alpha is not None else [], dtype=np.uint64).reshape(-1)
    s.rap_challenges = _arr_ptr(r, np.uint64, keep)
    s.num_rap = r.size // 3
    s.alpha_powers = _arr_ptr(a, np.uint64, keep)
    s.num_alpha = a.size // 3
    s.table_offset = _arr_ptr(table_offset, np.uint64, keep) if table_offset is not None else None
    s.root_nodes = (_arr_ptr(dev.root_nodes, np.uint32, keep) if getattr(dev, "root_nodes", None) is not None
                    else None)
    return s


def eval_composition(dev, main: GpuLde, aux, next_step: int, trace_rows: int, coset_offset: int, betas,
                     root_group, group_exemptions, boundary=(), b_betas=None, rap=None, alpha=None,
                     table_offset=None, natural_order: bool = False) -> DeviceBuffer:
    """Composition polynomial H on the LDE (crypto/math-cuda/src/constraint_interp.rs:409
    `eval_composition_on_device_keep`): a device buffer of N Fp3, rows in the
    tables' bit-reversed order unless natural_order. boundary: (col, is_aux,
    step, value3) tuples."""
    N.ensure_device()
    keep = []
    s, c = _composition_structs(dev, trace_rows, coset_offset, betas, root_group, group_exemptions, boundary,
                                b_betas, rap, alpha, table_offset, keep)
    out = DeviceBuffer(main.lde_rows * 24)
    N.call("lvm_eval_composition", C.byref(s), main.h, aux.h if aux is not None else None, next_step, C.byref(c),
           int(natural_order), C.c_void_p(out.ptr))
    return out


def _composition_structs(dev, trace_rows, coset_offset, betas, root_group, group_exemptions, boundary, b_betas, rap,
                         alpha, table_offset, keep):
    s = _program_struct(dev, rap, alpha, table_offset, keep)
    c = N.Composition()
    c.trace_rows = trace_rows
    c.coset_offset = coset_offset
    c.beta = _arr_ptr(np.asarray(betas, dtype=np.uint64).reshape(-1), np.uint64, keep)
    c.root_group = _arr_ptr(root_group, np.uint32, keep)
    c.group_exemptions = _arr_ptr(group_exemptions, np.uint32, keep)
    c.num_groups = len(group_exemptions)
    c.num_boundary = len(boundary)
    c.b_col = _arr_ptr([b[0] for b in boundary], np.uint32, keep)
    c.b_is_aux = _arr_ptr([int(b[1]) for b in boundary], np.uint32, keep)
    c.b_step = _arr_ptr([b[2] for b in boundary], np.uint64, keep)
    c.b_value = _arr_ptr(np.asarray([b[3] for b in boundary], dtype=np.uint64).reshape(-1) if boundary else None,
                         np.uint64, keep)
    c.b_beta = _arr_ptr(np.asarray(b_betas, dtype=np.uint64).reshape(-1) if boundary else None, np.uint64, keep)
    keep.extend([s, c])
    return s, c


def eval_composition_batch(tables, natural_order: bool = False) -> list:
    """The composition polynomials of several tables in shared launches
    (lvm_eval_composition_batch). tables: one dict per table with the
    arguments of eval_composition (dev, main, aux, next_step, trace_rows,
    coset_offset, betas, root_group, group_exemptions and optionally boundary,
    b_betas, rap, alpha, table_offset). Returns one device buffer per table,
    each equal to eval_composition's."""
    N.ensure_device()
    T = len(tables)
    if T == 0:
        return []
    keep, progs, comps, mains, auxs, steps, outs = [], [], [], [], [], [], []
    for t in tables:
        s, c = _composition_structs(t["dev"], t["trace_rows"], t["coset_offset"], t["betas"], t["root_group"],
                                    t["group_exemptions"], t.get("boundary", ()), t.get("b_betas"), t.get("rap"),
                                    t.get("alpha"), t.get("table_offset"), keep)
        progs.append(C.addressof(s))
        comps.append(C.addressof(c))
        mains.append(t["main"].h.value if isinstance(t["main"].h, C.c_void_p) else t["main"].h)
        aux = t.get("aux")
        auxs.append(None if aux is None else (aux.h.value if isinstance(aux.h, C.c_void_p) else aux.h))
        steps.append(t["next_step"])
        outs.append(DeviceBuffer(t["main"].lde_rows * 24))
    N.call("lvm_eval_composition_batch", (C.c_void_p * T)(*progs), (C.c_void_p * T)(*mains),
           (C.c_void_p * T)(*auxs), (C.c_size_t * T)(*steps), (C.c_void_p * T)(*comps), T, int(natural_order),
           (C.c_void_p * T)(*[o.ptr for o in outs]))
    return outs


def eval_constraints(dev, main: GpuLde, aux, next_step: int, rap=None, alpha=None, table_offset=None) -> np.ndarray:
    """(num_roots, N, 3) per-constraint values, natural row order
    (crypto/math-cuda/src/constraint_interp.rs:53 `eval_constraints_on_device`)."""
    N.ensure_device()
    keep = []
    s = _program_struct(dev, rap, alpha, table_offset, keep)
    out = DeviceBuffer(len(dev.roots) * main.lde_rows * 24)
    N.call("lvm_eval_constraints", C.byref(s), main.h, aux.h if aux is not None else None, next_step,
           C.c_void_p(out.ptr))
    res = out.to_array(np.uint64, (len(dev.roots), main.lde_rows, 3))
    out.free()
    return res


def prepare_constraints(dev, root_group=None, num_rap: int = 0, num_alpha: int = 0,
                        has_table_offset: bool = False) -> None:
    """Compile the program's kernel ahead of its first evaluation (no device
    needed): the composition form for the given zerofier group of each
    constraint, the per-constraint form when root_group is None. Raises
    LvmError (LVM_ERR_UNSUPPORTED) when the program is not compiled; it then
    runs on the interpreter kernel."""
    keep = []
    zero3 = np.zeros(3, dtype=np.uint64)
    s = _program_struct(dev, np.zeros((num_rap, 3), dtype=np.uint64), np.zeros((num_alpha, 3), dtype=np.uint64),
                        zero3 if has_table_offset else None, keep)
    c = None
    if root_group is not None:
        c = N.Composition()
        c.root_group = _arr_ptr(root_group, np.uint32, keep)
        c.num_groups = int(max(root_group)) + 1 if len(root_group) else 0
    N.call("lvm_prepare_constraints", C.byref(s), C.byref(c) if c is not None else None)


def set_constraint_jit(on: bool) -> None:
    """Run constraint programs as compiled kernels (default) or on the interpreter."""
    N.call("lvm_set_constraint_jit", int(bool(on)))


def commit_rows_dev(dev_rows: DeviceBuffer, lde_rows: int, words_per_row: int, trace_rows: int) -> GpuLde:
    """Commit device rows that already are an LDE in bit-reversed row order
    (the one-part composition), as a handle with its Merkle tree."""
    root = np.zeros(32, dtype=np.uint8)
    h = C.c_void_p()
    N.call("lvm_commit_rows_dev", C.c_void_p(dev_rows.ptr), lde_rows, words_per_row, trace_rows, _p8(root),
           C.byref(h))
    return GpuLde(h.value)


def composition_commit_d2(dev_h: DeviceBuffer, lde_rows: int, trace_rows: int, coset_offset: int) -> GpuLde:
    """Degree-2 split of the bit-reversed composition values, extension of both
    halves to the LDE and their commitment: the parts LDE (lde_rows rows of
    [H0 | H1]) over a table of trace_rows rows."""
    root = np.zeros(32, dtype=np.uint8)
    h = C.c_void_p()
    N.call("lvm_composition_commit_d2", C.c_void_p(dev_h.ptr), lde_rows, trace_rows, coset_offset, _p8(root),
           C.byref(h))
    return GpuLde(h.value)


def composition_commit_parts(dev_h: DeviceBuffer, lde_rows: int, trace_rows: int, num_parts: int,
                             coset_offset: int) -> GpuLde:
    """The composition polynomial's num_parts parts on the LDE, committed
    (crypto/stark/src/prover.rs:1683-1900): lde_rows rows of num_parts Fp3."""
    root = np.zeros(32, dtype=np.uint8)
    h = C.c_void_p()
    N.call("lvm_composition_commit_parts", C.c_void_p(dev_h.ptr), lde_rows, trace_rows, num_parts, coset_offset,
           _p8(root), C.byref(h))
    return GpuLde(h.value)


def generate_nonce(seed: bytes, grinding_factor: int, max_nonce: int = 1 << 40) -> int:
    """Smallest valid proof-of-work nonce (crypto/math-cuda/src/grinding.rs:29)."""
    N.ensure_device()
    sd = np.frombuffer(bytes(seed), dtype=np.uint8).copy()
    out = np.zeros(1, dtype=np.uint64)
    N.call("lvm_generate_nonce", _p8(sd), grinding_factor, max_nonce, _p64(out))
    return int(out[0])


def splitmix_table_dev(n: int, m: int, cfg: int, table: int) -> DeviceBuffer:
    """Device-generated synthetic table (workload.trace_table_modp on the GPU)."""
    b = DeviceBuffer(max(8, n * m * 8))
    N.call("lvm_fill_splitmix_table", C.c_void_p(b.ptr), n, m, cfg, table)
    return b


def build_merkle_tree(leaves) -> np.ndarray:
    """(num_leaves, 32) digests -> (2L-1, 32) nodes, root first."""
    N.ensure_device()
    lv = np.ascontiguousarray(leaves, dtype=np.uint8).reshape(-1, 32)
    L = 1
    while L < lv.shape[0]:
        L <<= 1
    out = np.zeros((2 * L - 1, 32), dtype=np.uint8)
    N.call("lvm_build_merkle_tree", _p8(lv), lv.shape[0], _p8(out))
    return out


def batch_inverse_ext3(a) -> np.ndarray:
    """Elementwise Fp3 inverses (crypto/math-cuda/src/inverse.rs:53)."""
    N.ensure_device()
    x = _u64(a).reshape(-1, 3)
    out = np.empty_like(x)
    N.call("lvm_batch_inverse_ext3", _p64(x), _p64(out), x.shape[0])
    return out


def barycentric_base_multi(h: GpuLde, coset_offset: int, z_points) -> np.ndarray:
    """(K, words_per_row, 3): every column of the resident table at each point
    (crypto/math-cuda/src/barycentric.rs:379)."""
    z = _u64(z_points).reshape(-1, 3)
    out = np.zeros((z.shape[0], h.words_per_row, 3), dtype=np.uint64)
    N.call("lvm_barycentric_base_multi", h.h, coset_offset, _p64(z), z.shape[0], _p64(out))
    return out


def deep_composition(main: GpuLde, coset_offset: int, z_points, trace_ood, gammas_tr, aux: GpuLde | None = None,
                     parts: GpuLde | None = None, z_parts=None, parts_ood=None, gammas_h=None) -> DeviceBuffer:
    """DEEP codeword (bit-reversed order, N Fp3) in a new device buffer
    (crypto/math-cuda/src/deep.rs:316)."""
    zp = _u64(z_points).reshape(-1, 3)
    tood = _u64(trace_ood).reshape(-1, 3)
    gtr = _u64(gammas_tr).reshape(-1, 3)
    out = DeviceBuffer(main.lde_rows * 24)
    zpa = _u64(z_parts).reshape(3) if z_parts is not None else None
    pood = _u64(parts_ood).reshape(-1, 3) if parts_ood is not None else None
    gh = _u64(gammas_h).reshape(-1, 3) if gammas_h is not None else None
    N.call("lvm_deep_composition", main.h, aux.h if aux else None, parts.h if parts else None, coset_offset,
           _p64(zp), zp.shape[0], _p64(zpa) if zpa is not None else None, _p64(tood),
           _p64(pood) if pood is not None else None, _p64(gtr), _p64(gh) if gh is not None else None,
           C.c_void_p(out.ptr))
    return out


class Transcript:
    """DefaultTranscript (crypto/crypto/src/fiat_shamir/default_transcript.rs:
    31-146) implemented in the library's host code."""

    def __init__(self, data: bytes = b"", _handle=None):
        if _handle is not None:
            self.h = _handle
            return
        buf = np.frombuffer(bytes(data) + b"\0", dtype=np.uint8).copy()
        h = C.c_void_p()
        N.call("lvm_transcript_new", _p8(buf), len(data), C.byref(h))
        self.h = h

    def __del__(self):
        try:
            if self.h is not None and self.h.value:
                N.lib().lvm_transcript_release(self.h)
        except Exception:
            pass

    def clone(self) -> "Transcript":
        h = C.c_void_p()
        N.call("lvm_transcript_clone", self.h, C.byref(h))
        return Transcript(_handle=h)

    def append_bytes(self, data: bytes) -> None:
        buf = np.frombuffer(bytes(data) + b"\0", dtype=np.uint8).copy()
        N.call("lvm_transcript_append_bytes", self.h, _p8(buf), len(data))

    def append_base(self, v: int) -> None:
        N.call("lvm_transcript_append_base", self.h, int(v))

    def append_fp3(self, v) -> None:
        a = _u64(v)
        N.call("lvm_transcript_append_fp3", self.h, _p64(a))

    def append_fp3_many(self, vals) -> None:
        """Append Fp3 elements in order (rows of a (count, 3) array)."""
        a = np.ascontiguousarray(_u64(vals).reshape(-1, 3))
        N.call("lvm_transcript_append_fp3_many", self.h, _p64(a), a.shape[0])

    def sample(self) -> bytes:
        out = np.zeros(32, dtype=np.uint8)
        N.call("lvm_transcript_sample", self.h, _p8(out))
        return out.tobytes()

    def state(self) -> bytes:
        out = np.zeros(32, dtype=np.uint8)
        N.call("lvm_transcript_state", self.h, _p8(out))
        return out.tobytes()

    def sample_base(self) -> int:
        v = C.c_uint64()
        N.call("lvm_transcript_sample_base", self.h, C.byref(v))
        return v.value

    def sample_fp3(self) -> np.ndarray:
        out = np.zeros(3, dtype=np.uint64)
        N.call("lvm_transcript_sample_fp3", self.h, _p64(out))
        return out

    def sample_u64(self, ub: int) -> int:
        v = C.c_uint64()
        N.call("lvm_transcript_sample_u64", self.h, ub, C.byref(v))
        return v.value


class GpuFri:
    """Device-resident FRI state (FriCommitState, crypto/math-cuda/src/fri.rs)."""

    def __init__(self, evals, coset_offset: int, on_device: DeviceBuffer | None = None, length: int | None = None,
                 adopt: bool = False):
        N.ensure_device()
        h = C.c_void_p()
        if on_device is not None:
            N.call("lvm_fri_new", C.c_void_p(on_device.ptr), length, coset_offset, 2 if adopt else 1, C.byref(h))
            if adopt:
                on_device.ptr = 0  # owned by the FRI state from now on
        else:
            ev = _u64(evals).reshape(-1, 3)
            N.call("lvm_fri_new", ev.ctypes.data_as(C.c_void_p), ev.shape[0], coset_offset, 0, C.byref(h))
        self.h = h

    def close(self) -> None:
        if self.h is not None and self.h.value:
            N.lib().lvm_fri_release(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def fold_and_commit_layer(self, zeta) -> bytes:
        z = _u64(zeta)
        root = np.zeros(32, dtype=np.uint8)
        N.call("lvm_fri_fold_and_commit_layer", self.h, _p64(z), _p8(root))
        return root.tobytes()

    def fold(self, zeta) -> None:
        z = _u64(zeta)
        N.call("lvm_fri_fold", self.h, _p64(z))

    def current(self):
        ln, off = C.c_size_t(), C.c_uint64()
        N.call("lvm_fri_current", self.h, None, C.byref(ln), C.byref(off))
        out = np.zeros((ln.value, 3), dtype=np.uint64)
        N.call("lvm_fri_current", self.h, _p64(out), C.byref(ln), C.byref(off))
        return out, off.value

    def num_layers(self) -> int:
        c = C.c_size_t()
        N.call("lvm_fri_num_layers", self.h, C.byref(c))
        return c.value

    def layer(self, i: int):
        ln = C.c_size_t()
        N.call("lvm_fri_layer", self.h, i, None, None, C.byref(ln))
        ev = np.zeros((ln.value, 3), dtype=np.uint64)
        L = 1
        while L < ln.value // 2:
            L <<= 1
        nodes = np.zeros((2 * L - 1, 32), dtype=np.uint8)
        N.call("lvm_fri_layer", self.h, i, _p64(ev), _p8(nodes), C.byref(ln))
        return ev, nodes

    def query(self, iotas):
        io = _u64(iotas)
        nl = self.num_layers()
        sym = np.zeros((nl, io.size, 3), dtype=np.uint64)
        depths = []
        for i in range(nl):
            ln = C.c_size_t()
            N.call("lvm_fri_layer", self.h, i, None, None, C.byref(ln))
            L = 1
            while L < ln.value // 2:
                L <<= 1
            depths.append(L.bit_length() - 1)
        paths = np.zeros(max(sum(depths) * io.size * 32, 1), dtype=np.uint8)
        N.call("lvm_fri_query", self.h, _p64(io), io.size, _p64(sym), _p8(paths))
        out, off = [], 0
        for d in depths:
            out.append(paths[off:off + io.size * d * 32].reshape(io.size, d, 32))
            off += io.size * d * 32
        return sym, out

    def commit_phase(self, transcript: Transcript, blowup_log: int, k: int):
        """roots (num_committed, 32) and terminal coefficients (2^k_eff, 3)."""
        ln, off = C.c_size_t(), C.c_uint64()
        N.call("lvm_fri_current", self.h, None, C.byref(ln), C.byref(off))
        lde_log = ln.value.bit_length() - 1
        roots = np.zeros((max(lde_log, 1), 32), dtype=np.uint8)
        coeffs = np.zeros((1 << max(lde_log, 0), 3), dtype=np.uint64)
        nc, ncf = C.c_size_t(), C.c_size_t()
        N.call("lvm_fri_commit_phase", self.h, transcript.h, blowup_log, k, _p8(roots), C.byref(nc), _p64(coeffs),
               C.byref(ncf))
        return roots[: nc.value].copy(), coeffs[: ncf.value].copy()


def fri_commit_phase_batch(fris, transcripts, blowup_logs, k: int):
    """The FRI commit phase of several tables in shared launches
    (lvm_fri_commit_phase_batch): [(roots (num_committed, 32), terminal
    coefficients (2^k_eff, 3))] per table, each transcript advanced as by its
    own GpuFri.commit_phase."""
    T = len(fris)
    if T == 0:
        return []
    fs = (C.c_void_p * T)(*[f.h.value for f in fris])
    ts = (C.c_void_p * T)(*[t.h.value for t in transcripts])
    bl = (C.c_uint * T)(*blowup_logs)
    roots, coeffs = [], []
    for f in fris:
        ln, off = C.c_size_t(), C.c_uint64()
        N.call("lvm_fri_current", f.h, None, C.byref(ln), C.byref(off))
        lde_log = ln.value.bit_length() - 1
        roots.append(np.zeros((max(lde_log, 1), 32), dtype=np.uint8))
        coeffs.append(np.zeros((1 << max(lde_log, 0), 3), dtype=np.uint64))
    rp = (C.c_void_p * T)(*[r.ctypes.data for r in roots])
    cp = (C.c_void_p * T)(*[c.ctypes.data for c in coeffs])
    nc = (C.c_size_t * T)()
    ncf = (C.c_size_t * T)()
    N.call("lvm_fri_commit_phase_batch", fs, ts, T, bl, k, rp, nc, cp, ncf)
    return [(roots[t][: nc[t]].copy(), coeffs[t][: ncf[t]].copy()) for t in range(T)]


def inject_fault_after(calls: int) -> None:
    N.lib().lvm_inject_fault_after(calls)


def prof_enable(on: bool = True) -> None:
    N.call("lvm_prof_enable", int(on))


def prof_reset() -> None:
    N.call("lvm_prof_reset")


def prof_query(name: str | None = None):
    cnt, ms = C.c_uint64(), C.c_double()
    N.call("lvm_prof_query", name.encode() if name else None, C.byref(cnt), C.byref(ms))
    return cnt.value, ms.value


def sync() -> None:
    N.call("lvm_sync")


def fp3_powers(base, count: int) -> np.ndarray:
    """(count, 3) canonical powers base^0 .. base^(count-1)."""
    b = _u64(base).reshape(3).copy()
    out = np.zeros((count, 3), dtype=np.uint64)
    N.call("lvm_fp3_powers", _p64(b), count, _p64(out))
    return out


def trim() -> None:
    """Hand the library's cached device buffers back to the driver pool."""
    N.call("lvm_trim")


def prove_committed_table(h: GpuLde, transcript: Transcript, n: int, m: int, blowup: int, offset: int,
                          fri_k: int = 6, num_queries: int = 30):
    """Rounds 3-4 of a committed table in one call (lvm_prove_committed_table):
    a pipeline.TableProof."""
    from .pipeline import TableProof
    N.ensure_device()
    Nl = n * blowup
    lde_log = Nl.bit_length() - 1
    blog = blowup.bit_length() - 1
    terminal_log = min(blog + fri_k, lde_log)
    total_folds = lde_log - terminal_log
    layers = total_folds - 1 if total_folds else 0
    depths = [lde_log - l - 2 for l in range(layers)]
    Q = num_queries
    ood = np.zeros((2, m, 3), dtype=np.uint64)
    roots = np.zeros((max(lde_log, 1), 32), dtype=np.uint8)
    coeffs = np.zeros((1 << terminal_log, 3), dtype=np.uint64)
    iotas = np.zeros(max(Q, 1), dtype=np.uint64)
    sym = np.zeros((max(layers, 1), max(Q, 1), 3), dtype=np.uint64)
    fpaths = np.zeros(max(sum(depths) * Q * 32, 1), dtype=np.uint8)
    tdepth = (Nl // 2).bit_length() - 1
    tpaths = np.zeros((max(Q, 1), max(tdepth, 1), 32), dtype=np.uint8)
    trows = np.zeros((max(Q, 1), 2, m), dtype=np.uint64)
    r = N.TableRounds()
    r.ood, r.fri_roots, r.fri_coeffs = ood.ctypes.data, roots.ctypes.data, coeffs.ctypes.data
    r.iotas, r.fri_sym, r.fri_paths = iotas.ctypes.data, sym.ctypes.data, fpaths.ctypes.data
    r.trace_paths, r.trace_rows = tpaths.ctypes.data, trows.ctypes.data
    N.call("lvm_prove_committed_table", h.h, transcript.h, n, blowup, offset, fri_k, Q, C.byref(r))
    nl = r.num_fri_roots
    if nl != layers:
        raise LvmError("lvm_prove_committed_table", 5, f"FRI committed {nl} layers, expected {layers}")
    paths, off = [], 0
    for d in depths:
        paths.append(fpaths[off:off + Q * d * 32].reshape(Q, d, 32))
        off += Q * d * 32
    return TableProof(h.root, [int(v) for v in r.z], ood, roots[:nl].copy(), coeffs[:r.num_fri_coeffs].copy(),
                      iotas[:Q].copy(), sym[:nl, :Q].copy(), paths, tpaths[:Q, :tdepth], trows[:Q],
                      transcript.state())
