"""Run-time compilation of constraint programs (lvm_prepare_constraints):
the generated straight-line kernels of the synthetic AIRs compile with the
run-time compiler without a device, in the composition and per-constraint
forms, and programs the generator does not take are reported as such (they
run on the interpreter kernel). Values are checked on the GPU against the
oracle in test_gpu_composition.py, in both forms."""
import numpy as np
import pytest

from lambda_vm_amd import _native
from lambda_vm_amd import air as A
from lambda_vm_amd import backend as B
from lambda_vm_amd import constraint_ir as ir


@pytest.fixture(scope="module")
def lib():
    try:
        return _native.lib()
    except OSError as e:  # pragma: no cover - library not built
        pytest.skip(f"library not built: {e}")


@pytest.mark.parametrize("make", [lambda: A.fibonacci_multi_column(6), lambda: A.quadratic_multi(5),
                                  lambda: A.logup_air(24, A.four_bus_interactions())],
                         ids=["fib", "quad", "logup"])
def test_synthetic_airs_compile(lib, make):
    a = make()
    groups, _ = a.groups()
    B.prepare_constraints(a.device, groups, num_rap=2, num_alpha=8, has_table_offset=True)
    B.prepare_constraints(a.device, None, num_rap=2, num_alpha=8, has_table_offset=True)


def test_pinned_program_compiles(lib):
    a = A.logup_air(24, A.four_bus_interactions())
    dev = ir.lower(a.program)
    dev.root_nodes = None
    B.prepare_constraints(dev, [0] * a.num_constraints, num_rap=2, num_alpha=8, has_table_offset=True)


def _raw_program(nodes, roots, nbs, nes):
    return ir.DeviceProgram(nodes=np.array(nodes, dtype=np.uint32).reshape(-1, 4),
                            base_consts=np.array([5], dtype=np.uint64), ext_consts=np.zeros((0, 3), dtype=np.uint64),
                            roots=np.array(roots, dtype=np.uint32), num_base=1, num_base_slots=nbs,
                            num_ext_slots=nes, max_offset=0, root_nodes=np.array([len(nodes) - 1], dtype=np.uint32))


def test_unsupported_programs_reported(lib):
    slot = lambda s: (ir.OPK_BASE_SLOT << ir.OPK_SHIFT) | s  # noqa: E731
    # reads base slot 1 before any node writes it
    dev = _raw_program([(ir.OP_ADD, slot(1), slot(1), 0)], [0], 2, 0)
    with pytest.raises(B.LvmError) as e:
        B.prepare_constraints(dev, [0])
    assert e.value.code == 6  # LVM_ERR_UNSUPPORTED
    # an extension value stored into a base slot
    dev = _raw_program([(ir.OP_CONST_EXT, 0, 0, 0)], [0], 1, 0)
    dev.ext_consts = np.zeros((1, 3), dtype=np.uint64)
    with pytest.raises(B.LvmError):
        B.prepare_constraints(dev, [0])
    # a valid one-node program compiles
    dev = _raw_program([(ir.OP_CONST_BASE, 0, 0, 0)], [0], 1, 0)
    B.prepare_constraints(dev, [0])


def test_large_program_runs_on_the_interpreter(lib):
    """Programs above the compile size limit (the VM's KECCAK table: 1687
    nodes) are declined quickly instead of compiled; they run on the
    interpreter kernel."""
    import time

    from lambda_vm_amd import vm_tables as V

    a = V.keccak_table_air()
    groups, _ = a.groups()
    t0 = time.time()
    with pytest.raises(B.LvmError):
        B.prepare_constraints(a.device, groups, num_rap=2, num_alpha=A.descriptor(a).alpha_powers_len,
                              has_table_offset=True)
    assert time.time() - t0 < 5
