"""ctypes binding of liblvm_gpu.so (the C ABI declared in include/lvm_gpu.h).

The library is loaded from this package directory. There is no fallback: if
the shared object is missing, or no GPU is visible when a compute entry point
is called, the call raises instead of computing anything on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "liblvm_gpu.so")
HEADER = os.path.join(ROOT, "include", "lvm_gpu.h")

u64p = C.POINTER(C.c_uint64)
u8p = C.POINTER(C.c_uint8)
szp = C.POINTER(C.c_size_t)
vpp = C.POINTER(C.c_void_p)

LVM_OK = 0


class LvmError(RuntimeError):
    """A C ABI call returned a non-zero status (the caller's CPU-arm signal)."""

    def __init__(self, name: str, code: int, msg: str):
        super().__init__(f"{name} failed with status {code}: {msg}")
        self.code = code


# name -> (restype, argtypes)
SIGNATURES = {
    "lvm_init": (C.c_int, [C.c_int]),
    "lvm_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "lvm_last_error": (C.c_char_p, []),
    "lvm_trim": (C.c_int, []),
    "lvm_sync": (C.c_int, []),
    "lvm_inject_fault_after": (None, [C.c_long]),
    "lvm_dev_alloc": (C.c_int, [vpp, C.c_size_t]),
    "lvm_dev_free": (C.c_int, [C.c_void_p]),
    "lvm_memcpy_htod": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "lvm_memcpy_dtoh": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "lvm_dev_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "lvm_gl_op": (C.c_int, [C.c_int, u64p, u64p, u64p, C.c_size_t]),
    "lvm_fp3_op": (C.c_int, [C.c_int, u64p, u64p, u64p, C.c_size_t]),
    "lvm_ntt_forward": (C.c_int, [u64p, u64p, C.c_size_t, C.c_size_t]),
    "lvm_ntt_inverse": (C.c_int, [u64p, u64p, C.c_size_t, C.c_size_t]),
    "lvm_coset_lde_row_major_with_merkle_tree_keep": (
        C.c_int, [u64p, C.c_size_t, C.c_size_t, C.c_size_t, u64p, u64p, u8p, vpp]),
    "lvm_coset_lde_dev": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_void_p, u8p, vpp]),
    "lvm_coset_lde_split_trees_dev": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_void_p,
                                                C.c_size_t, C.c_void_p, u8p, u8p, vpp]),
    "lvm_coset_lde_ext3_row_major_with_merkle_tree_keep": (
        C.c_int, [u64p, C.c_size_t, C.c_size_t, C.c_size_t, u64p, u64p, u8p, vpp]),
    "lvm_lde_release": (C.c_int, [C.c_void_p]),
    "lvm_lde_root": (C.c_int, [C.c_void_p, u8p]),
    "lvm_lde_shape": (C.c_int, [C.c_void_p, szp, szp, szp]),
    "lvm_lde_download": (C.c_int, [C.c_void_p, u64p]),
    "lvm_lde_device_ptr": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "lvm_lde_tree_nodes": (C.c_int, [C.c_void_p, u8p, szp]),
    "lvm_lde_open": (C.c_int, [C.c_void_p, u64p, C.c_size_t, u8p, u64p]),
    "lvm_coset_weights": (C.c_int, [C.c_size_t, C.c_uint64, C.c_void_p]),
    "lvm_fill_splitmix_table": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, C.c_uint64, C.c_uint64]),
    "lvm_batch_new": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int, vpp]),
    "lvm_batch_bind": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "lvm_batch_run": (C.c_int, [C.c_void_p, u8p]),
    "lvm_batch_take": (C.c_int, [C.c_void_p, C.c_size_t, vpp]),
    "lvm_batch_release": (C.c_int, [C.c_void_p]),
    "lvm_batch_lde_commit": (C.c_int, [C.c_void_p, C.c_size_t, u8p, vpp]),
    "lvm_batch_launch_count": (C.c_int, [C.c_void_p, szp]),
    "lvm_batch_launch_info": (C.c_int, [C.c_void_p, C.c_size_t, C.c_char_p, C.c_size_t, C.POINTER(C.c_double),
                                        C.POINTER(C.c_uint32)]),
    "lvm_set_ntt_plan": (C.c_int, [C.c_uint, C.POINTER(C.c_int), C.c_int]),
    "lvm_get_ntt_plan": (C.c_int, [C.c_uint, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "lvm_build_merkle_tree": (C.c_int, [u8p, C.c_size_t, u8p]),
    "lvm_batch_inverse_ext3": (C.c_int, [u64p, u64p, C.c_size_t]),
    "lvm_batch_inverse_ext3_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "lvm_barycentric_base_multi": (C.c_int, [C.c_void_p, C.c_uint64, u64p, C.c_size_t, u64p]),
    "lvm_deep_composition": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, u64p, C.c_size_t, u64p,
                                       u64p, u64p, u64p, u64p, C.c_void_p]),
    "lvm_logup_aux_dev": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, C.c_void_p, u64p,
                                    C.c_size_t, u64p, C.c_void_p, C.c_void_p, u64p]),
    "lvm_logup_term_columns": (C.c_int, [u64p, C.c_size_t, C.c_size_t, C.c_void_p, u64p, C.c_size_t, u64p, u64p]),
    "lvm_eval_composition": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_int,
                                       C.c_void_p]),
    "lvm_eval_composition_batch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_size_t, C.c_int, C.c_void_p]),
    "lvm_eval_constraints": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "lvm_commit_rows_dev": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, u8p, vpp]),
    "lvm_prepare_constraints": (C.c_int, [C.c_void_p, C.c_void_p]),
    "lvm_set_constraint_jit": (C.c_int, [C.c_int]),
    "lvm_composition_commit_d2": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, C.c_uint64, u8p, vpp]),
    "lvm_composition_commit_parts": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_uint64, u8p,
                                               vpp]),
    "lvm_generate_nonce": (C.c_int, [u8p, C.c_uint8, C.c_uint64, u64p]),
    "lvm_generate_nonce_noalloc": (C.c_int, [u8p, C.c_uint8, C.c_uint64, u64p]),
    "lvm_transcript_new": (C.c_int, [u8p, C.c_size_t, vpp]),
    "lvm_transcript_clone": (C.c_int, [C.c_void_p, vpp]),
    "lvm_transcript_release": (C.c_int, [C.c_void_p]),
    "lvm_transcript_append_bytes": (C.c_int, [C.c_void_p, u8p, C.c_size_t]),
    "lvm_transcript_append_fp3_many": (C.c_int, [C.c_void_p, u64p, C.c_size_t]),
    "lvm_fp3_powers": (C.c_int, [u64p, C.c_size_t, u64p]),
    "lvm_transcript_append_base": (C.c_int, [C.c_void_p, C.c_uint64]),
    "lvm_transcript_append_fp3": (C.c_int, [C.c_void_p, u64p]),
    "lvm_transcript_sample": (C.c_int, [C.c_void_p, u8p]),
    "lvm_transcript_state": (C.c_int, [C.c_void_p, u8p]),
    "lvm_transcript_sample_base": (C.c_int, [C.c_void_p, u64p]),
    "lvm_transcript_sample_fp3": (C.c_int, [C.c_void_p, u64p]),
    "lvm_transcript_sample_u64": (C.c_int, [C.c_void_p, C.c_uint64, u64p]),
    "lvm_fri_new": (C.c_int, [C.c_void_p, C.c_size_t, C.c_uint64, C.c_int, vpp]),
    "lvm_fri_fold_and_commit_layer": (C.c_int, [C.c_void_p, u64p, u8p]),
    "lvm_fri_fold": (C.c_int, [C.c_void_p, u64p]),
    "lvm_fri_current": (C.c_int, [C.c_void_p, u64p, szp, u64p]),
    "lvm_fri_num_layers": (C.c_int, [C.c_void_p, szp]),
    "lvm_fri_layer": (C.c_int, [C.c_void_p, C.c_size_t, u64p, u8p, szp]),
    "lvm_fri_query": (C.c_int, [C.c_void_p, u64p, C.c_size_t, u64p, u8p]),
    "lvm_fri_release": (C.c_int, [C.c_void_p]),
    "lvm_fri_commit_phase": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint, C.c_uint, u8p, szp, u64p, szp]),
    "lvm_fri_commit_phase_batch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_uint, C.c_void_p,
                                             C.c_void_p, C.c_void_p, C.c_void_p]),
    "lvm_prove_committed_table": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_uint, C.c_uint64, C.c_uint,
                                            C.c_size_t, C.c_void_p]),
    "lvm_prof_enable": (C.c_int, [C.c_int]),
    "lvm_prof_reset": (C.c_int, []),
    "lvm_prof_query": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
}


class TableDesc(C.Structure):
    """lvm_table_desc: one table of a batched commitment (device pointers)."""
    _fields_ = [("trace", C.c_void_p), ("weights", C.c_void_p), ("n", C.c_size_t), ("m", C.c_size_t),
                ("blowup", C.c_size_t)]


class LogupDesc(C.Structure):
    """lvm_logup_desc: flat CSR descriptor of a table's bus interactions (host arrays)."""
    _fields_ = [("num_interactions", C.c_size_t), ("bus_ids", C.c_void_p), ("elem_offsets", C.c_void_p),
                ("elem_alpha_idx", C.c_void_p), ("elem_const", C.c_void_p), ("term_offsets", C.c_void_p),
                ("term_coef", C.c_void_p), ("term_col", C.c_void_p), ("num_out_cols", C.c_size_t),
                ("out_col_offsets", C.c_void_p), ("out_col_interactions", C.c_void_p), ("mult_const", C.c_void_p),
                ("mult_term_offsets", C.c_void_p), ("mult_term_coef", C.c_void_p), ("mult_term_col", C.c_void_p)]


class Program(C.Structure):
    """lvm_program: flat constraint program + per-proof uniforms (host arrays)."""
    _fields_ = [("nodes", C.c_void_p), ("num_nodes", C.c_size_t), ("num_base_slots", C.c_size_t),
                ("num_ext_slots", C.c_size_t), ("base_consts", C.c_void_p), ("num_base_consts", C.c_size_t),
                ("ext_consts", C.c_void_p), ("num_ext_consts", C.c_size_t), ("roots", C.c_void_p),
                ("num_roots", C.c_size_t), ("rap_challenges", C.c_void_p), ("num_rap", C.c_size_t),
                ("alpha_powers", C.c_void_p), ("num_alpha", C.c_size_t), ("table_offset", C.c_void_p),
                ("root_nodes", C.c_void_p)]


class Composition(C.Structure):
    """lvm_composition: combination coefficients, zerofier groups, boundary constraints."""
    _fields_ = [("trace_rows", C.c_size_t), ("coset_offset", C.c_uint64), ("beta", C.c_void_p),
                ("root_group", C.c_void_p), ("group_exemptions", C.c_void_p), ("num_groups", C.c_size_t),
                ("num_boundary", C.c_size_t), ("b_col", C.c_void_p), ("b_is_aux", C.c_void_p),
                ("b_step", C.c_void_p), ("b_value", C.c_void_p), ("b_beta", C.c_void_p)]


class TableRounds(C.Structure):
    """lvm_table_rounds: caller-provided outputs of lvm_prove_committed_table."""
    _fields_ = [("z", C.c_uint64 * 3), ("ood", C.c_void_p), ("fri_roots", C.c_void_p),
                ("num_fri_roots", C.c_size_t), ("fri_coeffs", C.c_void_p), ("num_fri_coeffs", C.c_size_t),
                ("iotas", C.c_void_p), ("fri_sym", C.c_void_p), ("fri_paths", C.c_void_p),
                ("trace_paths", C.c_void_p), ("trace_rows", C.c_void_p)]


def header_symbols() -> list[str]:
    """Every function the public header declares."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lvm_[a-z0-9_]+)\s*\(", text)))


_lib = None


def lib():
    """Load liblvm_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m lambda_vm_amd.build` "
                "(there is no CPU fallback for the GPU path)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(name: str, rc: int) -> None:
    if rc != LVM_OK:
        msg = lib().lvm_last_error()
        raise LvmError(name, rc, msg.decode(errors="replace") if msg else "")


def call(name: str, *args) -> None:
    check(name, getattr(lib(), name)(*args))


_initialised = False


def ensure_device(device: int = 0, tuned: bool = True) -> None:
    """Initialise the library on `device` and install the tuned NTT pass
    splits (lambda_vm_amd/tune.py) when a table for this GPU exists; raises
    when no GPU is visible."""
    global _initialised
    if not _initialised:
        call("lvm_init", device)
        if tuned:
            from . import tune
            tune.install(call)
        _initialised = True


def ensure_device_this_thread(device: int = 0) -> None:
    """Make `device` current for the calling host thread (HIP's current
    device is per thread); worker threads call this before their first call."""
    call("lvm_init", device)


def device_count() -> int:
    c = C.c_int(0)
    lib().lvm_device_count(C.byref(c))
    return c.value
