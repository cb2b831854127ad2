"""The reference VM's LOAD table (prover/src/tables/load.rs) captured as an
AIR: its fourteen constraints (flag and width bits, width implies mu, the
extension bytes above each width equal signed * sign_bit * 255, mu a bit) and
the LogUp constraints of its five bus interactions (the 24-element MEMW
read, three MSB8 sign lookups with width-selected multiplicities, the MEMORY
op) vanish on every row of a generated trace (LB, LH, LW, LD, LBU, LHU, LWU
of values with the top bit set and clear); a wrong extension byte or a
flipped sign bit breaks an extension constraint; and the oracle's replay of
a full proof is accepted."""
import numpy as np
import pytest

from lambda_vm_amd import air as A
from lambda_vm_amd import vm_tables as V
from oracle.air_proof import replay_air_tables, verify_air_tables
from test_vm_table_eq import _constraint_values

P = 0xFFFFFFFF00000001


def _ops(k, seed):
    rng = np.random.default_rng(seed)
    ops = []
    for i in range(k):
        width = (1, 2, 4, 8)[i % 4]
        signed = bool(i % 3) and width != 8
        raw = int(rng.integers(0, 2**64, dtype=np.uint64))
        ops.append(V.LoadOperation.of_memory(int(rng.integers(0, 2**40)), 100 + 4 * i, width, signed, raw))
    ops += [V.LoadOperation.of_memory(8, 1, 1, True, 0x80), V.LoadOperation.of_memory(8, 2, 2, True, 0x7FFF),
            V.LoadOperation.of_memory(8, 3, 4, False, 0xFFFFFFFF), V.LoadOperation.of_memory(8, 4, 4, True, 0x80000000)]
    return ops


def test_extension_semantics():
    op = V.LoadOperation.of_memory(0, 0, 1, True, 0x1280)
    assert op.res == (0x80,) + (0xFF,) * 7 and op.sign_bit() == 1
    op = V.LoadOperation.of_memory(0, 0, 2, False, 0xFFFF8001)
    assert op.res == (0x01, 0x80) + (0,) * 6 and op.sign_bit() == 1
    assert V.LoadOperation.of_memory(0, 0, 4, True, 0x7FFFFFFF).res[4:] == (0,) * 4


def test_constraints_vanish_on_generated_trace(ora):
    air = V.load_table_air()
    ev = _constraint_values(ora, air, V.generate_load_trace(_ops(29, 1)))
    assert not np.any(ev % np.uint64(P))


@pytest.mark.parametrize("case", ["extension_byte", "sign_bit"])
def test_wrong_extension_breaks_constraints(ora, case):
    air = V.load_table_air()
    ops = _ops(9, 2)
    tr = V.generate_load_trace(ops)
    row = len(ops) - 4  # LB of 0x80: res = 80 FF FF FF FF FF FF FF
    c = V.LoadCols
    if case == "extension_byte":
        tr[row, c.RES[5]] = 0
    else:
        tr[row, c.SIGN_BIT] = 0
    ev = _constraint_values(ora, air, tr)
    assert np.any(ev[6:13, row] % np.uint64(P))


def test_load_proof_verifies():
    air = V.load_table_air()
    air.bus_descriptor = A.descriptor(air)
    proofs = replay_air_tables([(air, V.generate_load_trace(_ops(40, 3)))], 4, 3, 2, 6, 0, b"load")
    ok, why = verify_air_tables([air], proofs, 4, 3, 2, 6, 0, b"load")
    assert ok, why
