"""ORACLE — TEST INFRASTRUCTURE ONLY.

ctypes bindings for ``oracle/liblvm_oracle.so``, the C restatement of the
reference CPU prover's STARK hot path (Goldilocks/Fp3 arithmetic, Keccak-256,
the Fiat-Shamir transcript, the two-half batched FFT and coset LDE, the
row-pair Merkle commitment, FRI, barycentric OOD, batch inversion, grinding).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this package, as the checker or as the timed
CPU baseline. The product package ``lambda_vm_amd`` never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblvm_oracle.so")

P = 0xFFFFFFFF00000001

u64p = C.POINTER(C.c_uint64)
u8p = C.POINTER(C.c_uint8)
szp = C.POINTER(C.c_size_t)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    if force and os.path.exists(LIB_PATH):
        os.remove(LIB_PATH)
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _declare(C.CDLL(LIB_PATH))
    return _lib


def _declare(L):
    sig = {
        "ora_set_threads": (None, [C.c_int]),
        "ora_get_threads": (C.c_int, []),
        "ora_gl_op": (None, [C.c_int, u64p, u64p, u64p, C.c_size_t]),
        "ora_fp3_op": (None, [C.c_int, u64p, u64p, u64p, C.c_size_t]),
        "ora_gl_root_of_unity": (C.c_uint64, [C.c_uint]),
        "ora_gl_pow": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "ora_keccak256": (None, [u8p, C.c_size_t, u8p]),
        "ora_keccak_pad": (None, [u8p, C.c_size_t, C.c_uint8, u8p]),
        "ora_keccak_f1600": (None, [u64p]),
        "ora_transcript_new": (C.c_void_p, [u8p, C.c_size_t]),
        "ora_transcript_clone": (C.c_void_p, [C.c_void_p]),
        "ora_transcript_free": (None, [C.c_void_p]),
        "ora_transcript_append_bytes": (None, [C.c_void_p, u8p, C.c_size_t]),
        "ora_transcript_append_base": (None, [C.c_void_p, C.c_uint64]),
        "ora_transcript_append_fp3": (None, [C.c_void_p, u64p]),
        "ora_transcript_sample": (None, [C.c_void_p, u8p]),
        "ora_transcript_state": (None, [C.c_void_p, u8p]),
        "ora_transcript_sample_base": (C.c_uint64, [C.c_void_p]),
        "ora_transcript_sample_fp3": (None, [C.c_void_p, u64p]),
        "ora_transcript_sample_u64": (C.c_uint64, [C.c_void_p, C.c_uint64]),
        "ora_bitrev_permute_row_major": (None, [u64p, C.c_size_t, C.c_size_t, C.c_size_t]),
        "ora_fft_two_half": (C.c_int, [u64p, C.c_size_t, C.c_size_t, C.c_int]),
        "ora_lde_weights": (None, [C.c_size_t, C.c_uint64, u64p]),
        "ora_coset_lde_row_major": (C.c_int, [u64p, C.c_size_t, C.c_size_t, C.c_size_t, u64p, u64p]),
        "ora_commit_rows_bit_reversed": (
            C.c_int, [u64p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, u8p]),
        "ora_merkle_padded_leaves": (C.c_size_t, [C.c_size_t]),
        "ora_merkle_build": (C.c_int, [u8p, C.c_size_t, u8p]),
        "ora_merkle_path": (C.c_size_t, [u8p, C.c_size_t, C.c_size_t, u8p]),
        "ora_merkle_verify": (C.c_int, [u8p, u8p, C.c_size_t, u8p, C.c_size_t]),
        "ora_fri_inv_twiddles": (None, [C.c_uint64, C.c_size_t, u64p]),
        "ora_fri_fold": (None, [u64p, C.c_size_t, u64p, u64p, u64p]),
        "ora_fri_update_twiddles": (None, [u64p, C.c_size_t, u64p]),
        "ora_fri_layout": (None, [C.c_uint, C.c_uint, C.c_uint, C.POINTER(C.c_uint), szp, szp,
                                  C.POINTER(C.c_uint)]),
        "ora_fri_layer_leaves": (None, [u64p, C.c_size_t, u8p]),
        "ora_fri_commit": (C.c_int, [u64p, C.c_size_t, C.c_void_p, C.c_uint64, C.c_uint, C.c_uint,
                                     u8p, u64p, u64p]),
        "ora_fri_terminal_coeffs": (None, [u64p, C.c_size_t, C.c_uint64, C.c_uint, u64p]),
        "ora_barycentric_base": (None, [u64p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_uint64, u64p, u64p]),
        "ora_frame_points": (None, [u64p, szp, C.c_size_t, C.c_uint64, C.c_size_t, u64p]),
        "ora_batch_inverse_fp3": (C.c_int, [u64p, C.c_size_t]),
        "ora_batch_inverse_base": (C.c_int, [u64p, C.c_size_t]),
        "ora_is_valid_nonce": (C.c_int, [u8p, C.c_uint64, C.c_uint8]),
        "ora_generate_nonce": (C.c_uint64, [u8p, C.c_uint8, C.c_uint64]),
        "ora_lde_commit": (C.c_int, [u64p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_uint64, u64p, u8p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def _p64(a: np.ndarray):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(u64p)


def _p8(a):
    if isinstance(a, (bytes, bytearray)):
        a = np.frombuffer(bytes(a), dtype=np.uint8).copy()
    assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(u8p), a


def set_threads(t: int) -> None:
    lib().ora_set_threads(int(t))


def get_threads() -> int:
    return lib().ora_get_threads()


OPS = {"add": 0, "sub": 1, "mul": 2, "neg": 3, "inv": 4, "sqr": 5, "canon": 6}


def gl_op(op: str, a, b=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = a if b is None else np.ascontiguousarray(b, dtype=np.uint64)
    out = np.empty_like(a)
    lib().ora_gl_op(OPS[op], _p64(a), _p64(b), _p64(out), a.size)
    return out


def fp3_op(op: str, a, b=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 3)
    out = np.empty_like(a)
    if b is None:
        lib().ora_fp3_op(OPS[op], _p64(a), None, _p64(out), a.shape[0])
    else:
        b = np.ascontiguousarray(b, dtype=np.uint64).reshape(-1, 3)
        lib().ora_fp3_op(OPS[op], _p64(a), _p64(b), _p64(out), a.shape[0])
    return out


def root_of_unity(order: int) -> int:
    return lib().ora_gl_root_of_unity(order)


def gl_pow(a: int, e: int) -> int:
    return lib().ora_gl_pow(a, e)


def keccak256(data: bytes) -> bytes:
    out = np.zeros(32, dtype=np.uint8)
    buf = np.frombuffer(bytes(data) + b"\0", dtype=np.uint8).copy()
    lib().ora_keccak256(buf.ctypes.data_as(u8p), len(data), out.ctypes.data_as(u8p))
    return out.tobytes()


def keccak_pad(data: bytes, pad: int) -> bytes:
    out = np.zeros(32, dtype=np.uint8)
    buf = np.frombuffer(bytes(data) + b"\0", dtype=np.uint8).copy()
    lib().ora_keccak_pad(buf.ctypes.data_as(u8p), len(data), pad, out.ctypes.data_as(u8p))
    return out.tobytes()


class Transcript:
    """DefaultTranscript restatement (default_transcript.rs:31-146)."""

    def __init__(self, data: bytes = b"", _handle=None):
        if _handle is not None:
            self.h = _handle
        else:
            p, keep = _p8(np.frombuffer(bytes(data) + b"\0", dtype=np.uint8).copy())
            self.h = lib().ora_transcript_new(p, len(data))

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_transcript_free(self.h)
            self.h = None

    def clone(self) -> "Transcript":
        return Transcript(_handle=lib().ora_transcript_clone(self.h))

    def append_bytes(self, data: bytes) -> None:
        p, keep = _p8(np.frombuffer(bytes(data) + b"\0", dtype=np.uint8).copy())
        lib().ora_transcript_append_bytes(self.h, p, len(data))

    def append_base(self, v: int) -> None:
        lib().ora_transcript_append_base(self.h, int(v))

    def append_fp3(self, v) -> None:
        a = np.ascontiguousarray(v, dtype=np.uint64)
        lib().ora_transcript_append_fp3(self.h, _p64(a))

    def sample(self) -> bytes:
        out = np.zeros(32, dtype=np.uint8)
        lib().ora_transcript_sample(self.h, out.ctypes.data_as(u8p))
        return out.tobytes()

    def state(self) -> bytes:
        out = np.zeros(32, dtype=np.uint8)
        lib().ora_transcript_state(self.h, out.ctypes.data_as(u8p))
        return out.tobytes()

    def sample_base(self) -> int:
        return lib().ora_transcript_sample_base(self.h)

    def sample_fp3(self) -> np.ndarray:
        out = np.zeros(3, dtype=np.uint64)
        lib().ora_transcript_sample_fp3(self.h, _p64(out))
        return out

    def sample_u64(self, ub: int) -> int:
        return lib().ora_transcript_sample_u64(self.h, ub)


def fft_two_half(buf: np.ndarray, inverse: bool = False) -> np.ndarray:
    """buf: (n, m) uint64 row-major; returns a transformed copy."""
    out = np.ascontiguousarray(buf, dtype=np.uint64).copy()
    n, m = out.shape
    rc = lib().ora_fft_two_half(_p64(out), n, m, int(inverse))
    assert rc == 0
    return out


def lde_weights(n: int, offset: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint64)
    lib().ora_lde_weights(n, offset, _p64(out))
    return out


def coset_lde_row_major(trace: np.ndarray, blowup: int, offset: int, weights=None) -> np.ndarray:
    trace = np.ascontiguousarray(trace, dtype=np.uint64)
    n, m = trace.shape
    w = lde_weights(n, offset) if weights is None else np.ascontiguousarray(weights, dtype=np.uint64)
    out = np.empty((n * blowup, m), dtype=np.uint64)
    rc = lib().ora_coset_lde_row_major(_p64(trace), n, m, blowup, _p64(w), _p64(out))
    assert rc == 0, rc
    return out


def merkle_padded_leaves(n: int) -> int:
    return lib().ora_merkle_padded_leaves(n)


def commit_rows_bit_reversed(data: np.ndarray, elem_words: int = 1, c0: int = 0, c1=None) -> np.ndarray:
    """data: (rows, cols[*3]) row-major; returns the (2L-1, 32) node array."""
    data = np.ascontiguousarray(data, dtype=np.uint64)
    rows = data.shape[0]
    cols = data.shape[1] // elem_words
    c1 = cols if c1 is None else c1
    L = merkle_padded_leaves(rows // 2)
    nodes = np.zeros((2 * L - 1, 32), dtype=np.uint8)
    rc = lib().ora_commit_rows_bit_reversed(_p64(data), rows, cols, elem_words, c0, c1,
                                            nodes.ctypes.data_as(u8p))
    assert rc == 0, rc
    return nodes


def merkle_build(leaves: np.ndarray) -> np.ndarray:
    leaves = np.ascontiguousarray(leaves, dtype=np.uint8).reshape(-1, 32)
    L = merkle_padded_leaves(leaves.shape[0])
    nodes = np.zeros((2 * L - 1, 32), dtype=np.uint8)
    rc = lib().ora_merkle_build(leaves.ctypes.data_as(u8p), leaves.shape[0], nodes.ctypes.data_as(u8p))
    assert rc == 0
    return nodes


def merkle_path(nodes: np.ndarray, pos: int) -> np.ndarray:
    nodes = np.ascontiguousarray(nodes, dtype=np.uint8)
    L = (nodes.shape[0] + 1) // 2
    out = np.zeros((64, 32), dtype=np.uint8)
    k = lib().ora_merkle_path(nodes.ctypes.data_as(u8p), L, pos, out.ctypes.data_as(u8p))
    return out[:k].copy()


def merkle_verify(root: bytes, leaf: bytes, pos: int, path: np.ndarray) -> bool:
    r, kr = _p8(root)
    lf, kl = _p8(leaf)
    path = np.ascontiguousarray(path, dtype=np.uint8)
    return bool(lib().ora_merkle_verify(r, lf, pos, path.ctypes.data_as(u8p), path.shape[0]))


def fri_inv_twiddles(offset: int, domain_size: int) -> np.ndarray:
    out = np.empty(domain_size // 2, dtype=np.uint64)
    lib().ora_fri_inv_twiddles(offset, domain_size, _p64(out))
    return out


def fri_fold(evals: np.ndarray, zeta, inv_tw: np.ndarray) -> np.ndarray:
    evals = np.ascontiguousarray(evals, dtype=np.uint64).reshape(-1, 3)
    z = np.ascontiguousarray(zeta, dtype=np.uint64)
    tw = np.ascontiguousarray(inv_tw, dtype=np.uint64)
    out = np.empty((evals.shape[0] // 2, 3), dtype=np.uint64)
    lib().ora_fri_fold(_p64(evals), evals.shape[0], _p64(z), _p64(tw), _p64(out))
    return out


def fri_layout(lde_log: int, blowup_log: int, k: int):
    tf, ek = C.c_uint(), C.c_uint()
    nc, tl = C.c_size_t(), C.c_size_t()
    lib().ora_fri_layout(lde_log, blowup_log, k, C.byref(tf), C.byref(nc), C.byref(tl), C.byref(ek))
    return tf.value, nc.value, tl.value, ek.value


def fri_layer_leaves(evals: np.ndarray) -> np.ndarray:
    evals = np.ascontiguousarray(evals, dtype=np.uint64).reshape(-1, 3)
    out = np.zeros((evals.shape[0] // 2, 32), dtype=np.uint8)
    lib().ora_fri_layer_leaves(_p64(evals), evals.shape[0], out.ctypes.data_as(u8p))
    return out


def fri_commit(evals: np.ndarray, transcript: Transcript, offset: int, blowup_log: int, k: int,
               keep_layers: bool = False):
    ev = np.ascontiguousarray(evals, dtype=np.uint64).reshape(-1, 3).copy()
    n = ev.shape[0]
    tf, nc, tl, ek = fri_layout(n.bit_length() - 1, blowup_log, k)
    roots = np.zeros((max(nc, 1), 32), dtype=np.uint8)
    coeffs = np.zeros((1 << ek, 3), dtype=np.uint64)
    layers = None
    if keep_layers:
        tot = sum(n >> (i + 1) for i in range(nc))
        layers = np.zeros((max(tot, 1), 3), dtype=np.uint64)
    rc = lib().ora_fri_commit(_p64(ev), n, transcript.h, offset, blowup_log, k,
                              roots.ctypes.data_as(u8p), _p64(coeffs),
                              _p64(layers) if layers is not None else None)
    assert rc == 0
    return roots[:nc].copy(), coeffs, layers


def fri_terminal_coeffs(codeword_bitrev: np.ndarray, terminal_offset: int, k: int) -> np.ndarray:
    cw = np.ascontiguousarray(codeword_bitrev, dtype=np.uint64).reshape(-1, 3)
    out = np.zeros((1 << k, 3), dtype=np.uint64)
    lib().ora_fri_terminal_coeffs(_p64(cw), cw.shape[0], terminal_offset, k, _p64(out))
    return out


def barycentric_base(lde: np.ndarray, n: int, blowup: int, offset: int, z) -> np.ndarray:
    lde = np.ascontiguousarray(lde, dtype=np.uint64)
    m = lde.shape[1]
    zz = np.ascontiguousarray(z, dtype=np.uint64)
    out = np.zeros((m, 3), dtype=np.uint64)
    lib().ora_barycentric_base(_p64(lde), n, m, blowup, offset, _p64(zz), _p64(out))
    return out


def frame_points(z, frame_offsets, trace_root: int, step_size: int) -> np.ndarray:
    zz = np.ascontiguousarray(z, dtype=np.uint64)
    fo = (C.c_size_t * len(frame_offsets))(*frame_offsets)
    out = np.zeros((len(frame_offsets) * step_size, 3), dtype=np.uint64)
    lib().ora_frame_points(_p64(zz), fo, len(frame_offsets), trace_root, step_size, _p64(out))
    return out


def batch_inverse_fp3(v: np.ndarray) -> np.ndarray:
    out = np.ascontiguousarray(v, dtype=np.uint64).reshape(-1, 3).copy()
    rc = lib().ora_batch_inverse_fp3(_p64(out), out.shape[0])
    assert rc == 0
    return out


def batch_inverse_base(v: np.ndarray) -> np.ndarray:
    out = np.ascontiguousarray(v, dtype=np.uint64).copy()
    rc = lib().ora_batch_inverse_base(_p64(out), out.shape[0])
    assert rc == 0
    return out


def is_valid_nonce(seed: bytes, nonce: int, g: int) -> bool:
    p, keep = _p8(seed)
    return bool(lib().ora_is_valid_nonce(p, nonce, g))


def generate_nonce(seed: bytes, g: int, limit_search: int = 1 << 40) -> int:
    p, keep = _p8(seed)
    return lib().ora_generate_nonce(p, g, limit_search)


def lde_commit(trace: np.ndarray, blowup: int, offset: int, want_lde: bool = True):
    trace = np.ascontiguousarray(trace, dtype=np.uint64)
    n, m = trace.shape
    N = n * blowup
    L = merkle_padded_leaves(N // 2)
    nodes = np.zeros((2 * L - 1, 32), dtype=np.uint8)
    lde = np.empty((N, m), dtype=np.uint64) if want_lde else None
    rc = lib().ora_lde_commit(_p64(trace), n, m, blowup, offset,
                              _p64(lde) if lde is not None else None, nodes.ctypes.data_as(u8p))
    assert rc == 0, rc
    return lde, nodes


def deep_composition(main: np.ndarray, offset: int, z_points, trace_ood, gammas_tr, aux=None, parts=None,
                     z_parts=None, parts_ood=None, gammas_h=None) -> np.ndarray:
    """Natural-order DEEP codeword (N, 3). main: (N, m) base; aux/parts:
    (N, 3*ma) / (N, 3*np) Fp3 row-major."""
    L = lib()
    if not hasattr(L.ora_deep_composition, "_declared"):
        L.ora_deep_composition.restype = C.c_int
        L.ora_deep_composition.argtypes = [u64p, C.c_size_t, u64p, C.c_size_t, u64p, C.c_size_t, C.c_size_t,
                                           C.c_uint64, u64p, C.c_size_t, u64p, u64p, u64p, u64p, u64p, u64p]
        L.ora_deep_composition._declared = True
    main = np.ascontiguousarray(main, dtype=np.uint64)
    N, m = main.shape
    zp = np.ascontiguousarray(z_points, dtype=np.uint64).reshape(-1, 3)
    K = zp.shape[0]
    aux_a = np.ascontiguousarray(aux, dtype=np.uint64) if aux is not None else None
    parts_a = np.ascontiguousarray(parts, dtype=np.uint64) if parts is not None else None
    ma = aux_a.shape[1] // 3 if aux_a is not None else 0
    npp = parts_a.shape[1] // 3 if parts_a is not None else 0
    arr = lambda x: np.ascontiguousarray(x, dtype=np.uint64) if x is not None else None  # noqa: E731
    tood, gtr, zpa, pood, gh = arr(trace_ood), arr(gammas_tr), arr(z_parts), arr(parts_ood), arr(gammas_h)
    out = np.zeros((N, 3), dtype=np.uint64)
    p = lambda x: _p64(x) if x is not None else None  # noqa: E731
    rc = L.ora_deep_composition(_p64(main), m, p(aux_a), ma, p(parts_a), npp, N, offset, _p64(zp), K, p(zpa),
                                _p64(tood), p(pood), _p64(gtr), p(gh), _p64(out))
    assert rc == 0
    return out


class _LogupDesc(C.Structure):
    _fields_ = [("num_interactions", C.c_size_t), ("bus_ids", C.c_void_p), ("elem_offsets", C.c_void_p),
                ("elem_alpha_idx", C.c_void_p), ("elem_const", C.c_void_p), ("term_offsets", C.c_void_p),
                ("term_coef", C.c_void_p), ("term_col", C.c_void_p), ("num_out_cols", C.c_size_t),
                ("out_col_offsets", C.c_void_p), ("out_col_interactions", C.c_void_p), ("mult_const", C.c_void_p),
                ("mult_term_offsets", C.c_void_p), ("mult_term_coef", C.c_void_p), ("mult_term_col", C.c_void_p)]


_DESC_ARRAYS = [("bus_ids", np.uint64), ("elem_offsets", np.uint32), ("elem_alpha_idx", np.uint32),
                ("elem_const", np.uint64), ("term_offsets", np.uint32), ("term_coef", np.uint64),
                ("term_col", np.uint32), ("out_col_offsets", np.uint32), ("out_col_interactions", np.uint32),
                ("mult_const", np.uint64), ("mult_term_offsets", np.uint32), ("mult_term_coef", np.uint64),
                ("mult_term_col", np.uint32)]


def _logup_desc(d):
    """ctypes view of a descriptor object with the FingerprintDescriptor fields;
    returns (struct, arrays kept alive)."""
    keep = {}
    s = _LogupDesc()
    s.num_interactions = d.num_interactions
    s.num_out_cols = d.num_out_cols
    for name, dt in _DESC_ARRAYS:
        a = np.ascontiguousarray(getattr(d, name), dtype=dt)
        if a.size == 0:
            a = np.zeros(1, dtype=dt)
        keep[name] = a
        setattr(s, name, a.ctypes.data)
    return s, keep


def _declare_logup(L):
    if not hasattr(L.ora_logup_aux, "_declared"):
        L.ora_logup_aux.restype = C.c_int
        L.ora_logup_aux.argtypes = [C.c_void_p, u64p, C.c_size_t, C.c_size_t, C.c_size_t, u64p, u64p, u64p, u64p,
                                    u64p]
        L.ora_logup_fingerprints.restype = C.c_int
        L.ora_logup_fingerprints.argtypes = [C.c_void_p, u64p, C.c_size_t, C.c_size_t, C.c_size_t, u64p, u64p, u64p]
        L.ora_logup_aux._declared = True


def logup_fingerprints(desc, main: np.ndarray, alpha_powers, z) -> np.ndarray:
    """(K, n, 3) fingerprints of a row-major (n, m) main trace."""
    L = lib()
    _declare_logup(L)
    main = np.ascontiguousarray(main, dtype=np.uint64)
    n, m = main.shape
    s, keep = _logup_desc(desc)
    al = np.ascontiguousarray(alpha_powers, dtype=np.uint64).reshape(-1)
    zz = np.ascontiguousarray(z, dtype=np.uint64)
    out = np.zeros((desc.num_interactions, n, 3), dtype=np.uint64)
    rc = L.ora_logup_fingerprints(C.addressof(s), _p64(main), m, 1, n, _p64(al), _p64(zz), _p64(out))
    assert rc == 0
    return out


def logup_aux(desc, main: np.ndarray, alpha_powers, z):
    """(aux (n, C, 3), L (3,), terms (C, n, 3)) for a row-major (n, m) main
    trace: committed term columns then the accumulated column."""
    L = lib()
    _declare_logup(L)
    main = np.ascontiguousarray(main, dtype=np.uint64)
    n, m = main.shape
    s, keep = _logup_desc(desc)
    al = np.ascontiguousarray(alpha_powers, dtype=np.uint64).reshape(-1)
    zz = np.ascontiguousarray(z, dtype=np.uint64)
    C_ = desc.num_out_cols
    aux = np.zeros((n, C_, 3), dtype=np.uint64)
    terms = np.zeros((C_, n, 3), dtype=np.uint64)
    Lo = np.zeros(3, dtype=np.uint64)
    rc = L.ora_logup_aux(C.addressof(s), _p64(main), m, 1, n, _p64(al), _p64(zz), _p64(terms), _p64(aux), _p64(Lo))
    assert rc == 0, rc
    return aux, Lo, terms


class _Program(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("num_nodes", C.c_size_t), ("num_base_slots", C.c_size_t),
                ("num_ext_slots", C.c_size_t), ("base_consts", C.c_void_p), ("ext_consts", C.c_void_p),
                ("roots", C.c_void_p), ("num_roots", C.c_size_t), ("rap", C.c_void_p), ("alpha", C.c_void_p),
                ("table_offset", C.c_void_p), ("root_nodes", C.c_void_p)]


class _Composition(C.Structure):
    _fields_ = [("trace_rows", C.c_size_t), ("coset_offset", C.c_uint64), ("beta", C.c_void_p),
                ("root_group", C.c_void_p), ("group_exemptions", C.c_void_p), ("num_groups", C.c_size_t),
                ("num_boundary", C.c_size_t), ("b_col", C.c_void_p), ("b_is_aux", C.c_void_p),
                ("b_step", C.c_void_p), ("b_value", C.c_void_p), ("b_beta", C.c_void_p)]


def _arr(x, dt, keep):
    a = np.ascontiguousarray(x if x is not None and len(x) else [0], dtype=dt).reshape(-1)
    keep.append(a)
    return a.ctypes.data


def _program(dev, rap, alpha, table_offset, keep):
    s = _Program()
    s.nodes = _arr(dev.packed_nodes(), np.uint64, keep)
    s.num_nodes = len(dev.nodes)
    s.num_base_slots = dev.num_base_slots
    s.num_ext_slots = dev.num_ext_slots
    s.base_consts = _arr(dev.base_consts, np.uint64, keep)
    s.ext_consts = _arr(np.asarray(dev.ext_consts).reshape(-1), np.uint64, keep)
    s.roots = _arr(dev.roots.astype(np.uint64), np.uint64, keep)
    s.num_roots = len(dev.roots)
    s.rap = _arr(np.asarray(rap, dtype=np.uint64).reshape(-1) if rap is not None else None, np.uint64, keep)
    s.alpha = _arr(np.asarray(alpha, dtype=np.uint64).reshape(-1) if alpha is not None else None, np.uint64, keep)
    s.table_offset = _arr(np.asarray(table_offset if table_offset is not None else (0, 0, 0), dtype=np.uint64),
                          np.uint64, keep)
    rn = getattr(dev, "root_nodes", None)
    s.root_nodes = _arr(rn, np.uint32, keep) if rn is not None else None
    return s


def _declare_constraints(L):
    if not hasattr(L.ora_eval_composition, "_declared"):
        L.ora_eval_constraints.restype = C.c_int
        L.ora_eval_constraints.argtypes = [C.c_void_p, u64p, C.c_size_t, u64p, C.c_size_t, C.c_size_t, C.c_size_t,
                                           u64p]
        L.ora_eval_composition.restype = C.c_int
        L.ora_eval_composition.argtypes = [C.c_void_p, u64p, C.c_size_t, u64p, C.c_size_t, C.c_size_t, C.c_size_t,
                                           C.c_void_p, u64p]
        L.ora_eval_composition._declared = True


def eval_constraints(dev, main_lde, aux_lde, next_step, rap=None, alpha=None, table_offset=None) -> np.ndarray:
    """(num_roots, N, 3) constraint values on a natural-order LDE (main (N, m),
    aux (N, 3*ma) or None)."""
    L = lib()
    _declare_constraints(L)
    keep = []
    s = _program(dev, rap, alpha, table_offset, keep)
    main = np.ascontiguousarray(main_lde, dtype=np.uint64)
    N, m = main.shape
    aux = np.ascontiguousarray(aux_lde if aux_lde is not None else np.zeros((N, 3)), dtype=np.uint64)
    ma = aux.shape[1] // 3
    out = np.zeros((len(dev.roots), N, 3), dtype=np.uint64)
    rc = L.ora_eval_constraints(C.addressof(s), _p64(main), m, _p64(aux), ma, N, next_step, _p64(out))
    assert rc == 0
    return out


def eval_composition(dev, main_lde, aux_lde, next_step, trace_rows, coset_offset, betas, root_group,
                     group_exemptions, boundary=(), b_betas=None, rap=None, alpha=None, table_offset=None):
    """(N, 3) composition evaluations, natural order. boundary: (col, is_aux,
    step, value3) tuples."""
    L = lib()
    _declare_constraints(L)
    keep = []
    s = _program(dev, rap, alpha, table_offset, keep)
    main = np.ascontiguousarray(main_lde, dtype=np.uint64)
    N, m = main.shape
    aux = np.ascontiguousarray(aux_lde if aux_lde is not None else np.zeros((N, 3)), dtype=np.uint64)
    ma = aux.shape[1] // 3
    c = _Composition()
    c.trace_rows = trace_rows
    c.coset_offset = coset_offset
    c.beta = _arr(np.asarray(betas, dtype=np.uint64).reshape(-1), np.uint64, keep)
    c.root_group = _arr(root_group, np.uint32, keep)
    c.group_exemptions = _arr(group_exemptions, np.uint32, keep)
    c.num_groups = len(group_exemptions)
    c.num_boundary = len(boundary)
    c.b_col = _arr([b[0] for b in boundary], np.uint32, keep)
    c.b_is_aux = _arr([int(b[1]) for b in boundary], np.uint32, keep)
    c.b_step = _arr([b[2] for b in boundary], np.uint64, keep)
    c.b_value = _arr(np.asarray([b[3] for b in boundary], dtype=np.uint64).reshape(-1) if boundary else None,
                     np.uint64, keep)
    c.b_beta = _arr(np.asarray(b_betas, dtype=np.uint64).reshape(-1) if boundary else None, np.uint64, keep)
    out = np.zeros((N, 3), dtype=np.uint64)
    rc = L.ora_eval_composition(C.addressof(s), _p64(main), m, _p64(aux), ma, N, next_step, C.addressof(c),
                                _p64(out))
    assert rc == 0, rc
    return out


def decompose_d2(h: np.ndarray, offset: int) -> np.ndarray:
    """(N/2, 6) natural rows [H0 | H1] of a natural (N, 3) composition codeword."""
    L = lib()
    if not hasattr(L.ora_decompose_d2, "_declared"):
        L.ora_decompose_d2.restype = C.c_int
        L.ora_decompose_d2.argtypes = [u64p, C.c_size_t, C.c_uint64, u64p]
        L.ora_decompose_d2._declared = True
    h = np.ascontiguousarray(h, dtype=np.uint64)
    N = h.shape[0]
    out = np.zeros((N // 2, 6), dtype=np.uint64)
    assert L.ora_decompose_d2(_p64(h), N, offset, _p64(out)) == 0
    return out
