import numpy as np

from lambda_vm_amd import workload


def _seq(seed, count):
    s = seed
    out = []
    while len(out) < count:
        s = (s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        z ^= z >> 31
        if z < workload.P:
            out.append(z)
    return out


def test_vectorised_splitmix_matches_sequential_stream():
    for seed in (0, 1, 0x1A3BDA7A, 2**63 + 5):
        assert workload.field_stream(seed, 1000).tolist() == _seq(seed, 1000)


def test_trace_table_shape_and_range():
    t = workload.trace_table(0, 0, 64, 5)
    assert t.shape == (64, 5) and t.dtype == np.uint64
    assert (t < np.uint64(workload.P)).all()
    assert len({tuple(t[:, c]) for c in range(5)}) == 5
