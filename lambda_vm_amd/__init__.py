"""lambda_vm_amd — MI355X (gfx950) STARK proving hot path.

HIP kernels behind a C ABI (include/lvm_gpu.h, built into liblvm_gpu.so in
this directory) for the prover path of lambda_vm's crypto/stark: Goldilocks
coset LDE, Keccak-256 Merkle commitments, FRI folding, barycentric OOD
evaluation, LogUp running sums and AIR composition. `backend` mirrors the
reference's crypto/math-cuda surface on top of that ABI.
"""
from . import _native  # noqa: F401

__all__ = ["backend", "workload"]
