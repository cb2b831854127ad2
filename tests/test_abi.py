"""The C-ABI library loads and exports every symbol include/lvm_gpu.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os

import pytest

from lambda_vm_amd import _native


def test_header_declares_entry_points():
    syms = _native.header_symbols()
    assert "lvm_coset_lde_row_major_with_merkle_tree_keep" in syms
    assert len(syms) >= 20


def test_library_exports_every_header_symbol():
    if not os.path.exists(_native.LIB_PATH):
        pytest.fail("liblvm_gpu.so not built (run python -m lambda_vm_amd.build)")
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [s for s in _native.header_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    # the Python binding declares a signature for every exported entry point
    assert set(_native.header_symbols()) <= set(_native.SIGNATURES), \
        set(_native.header_symbols()) - set(_native.SIGNATURES)


def test_no_cpu_fallback_without_library(tmp_path, monkeypatch):
    monkeypatch.setattr(_native, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_native, "_lib", None)
    with pytest.raises(ImportError):
        _native.lib()


def test_product_package_never_imports_the_oracle():
    pkg = os.path.dirname(_native.__file__)
    for root, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h")):
                text = open(os.path.join(root, f), errors="replace").read()
                assert "import oracle" not in text and "from oracle" not in text, f
                assert "lvm_oracle" not in text and "liblvm_oracle" not in text, f


@pytest.mark.parametrize("n,m,blowup", [(1, 4, 4), (3, 4, 4), (16, 0, 4), (16, 4, 3), (1 << 31, 1, 4)])
def test_lde_shapes_refused_before_device_work(n, m, blowup):
    """Shapes the transform cannot take (one row, non-powers of two, no
    columns, a domain past 2^32) are refused with LVM_ERR_ARG before any
    device is touched, so the caller's CPU arm runs (gpu_lde.rs:313-335)."""
    lib = _native.lib()
    dummy = (ctypes.c_uint64 * 4)()
    h = ctypes.c_void_p()
    root = (ctypes.c_uint8 * 32)()
    for fn in ("lvm_coset_lde_dev", "lvm_coset_lde_row_major_with_merkle_tree_keep"):
        f = getattr(lib, fn)
        if fn == "lvm_coset_lde_dev":
            rc = f(dummy, ctypes.c_size_t(n), ctypes.c_size_t(m), ctypes.c_size_t(blowup), dummy, root,
                   ctypes.byref(h))
        else:
            rc = f(dummy, ctypes.c_size_t(n), ctypes.c_size_t(m), ctypes.c_size_t(blowup), dummy, None, root,
                   ctypes.byref(h))
        assert rc == 1, (fn, rc)  # LVM_ERR_ARG
        assert h.value is None


def test_batched_fri_arguments_refused_before_device_work():
    """lvm_fri_commit_phase_batch refuses missing arrays with LVM_ERR_ARG
    (no device touched); zero tables is a no-op."""
    lib = _native.lib()
    f = lib.lvm_fri_commit_phase_batch
    assert f(None, None, ctypes.c_size_t(0), None, ctypes.c_uint(2), None, None, None, None) == 0
    assert f(None, None, ctypes.c_size_t(1), None, ctypes.c_uint(2), None, None, None, None) == 1
